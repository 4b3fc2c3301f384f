"""MI355X-native contact step (hand-written gfx950 HIP kernels behind a C ABI) behind the
reference ContactEngine protocol (antonbj3/3fold-motion-engine, src/motion_engine/contact_engine.py)."""
from .protocol import BodyState, ContactEngine
from .engines import ColorCacheContactEngine, GraphColoredContactEngine, RelaxedJacobiContactEngine
from . import adjoint, bodies, ncp, pgs, scenes
from .pgs import ExactConeContactEngine

__all__ = ["BodyState", "ContactEngine", "RelaxedJacobiContactEngine", "GraphColoredContactEngine",
           "ColorCacheContactEngine", "ExactConeContactEngine", "adjoint", "bodies", "ncp", "pgs", "scenes"]
