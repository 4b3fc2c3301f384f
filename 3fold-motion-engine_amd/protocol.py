"""The reference's data contract (src/motion_engine/contact_engine.py:38-55): plain numpy data in
and out, a runtime-checkable Protocol, so callers can swap engines."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Protocol, runtime_checkable

import numpy as np


@dataclass
class BodyState:
    """xc=(N,3) centre of mass, Rm=(N,3,3) orientation, vc=(N,3) linear, om=(N,3) angular velocity."""
    xc: np.ndarray
    Rm: np.ndarray
    vc: np.ndarray
    om: np.ndarray


@runtime_checkable
class ContactEngine(Protocol):
    name: str

    def step(self, dt: float, substeps: int = 1) -> None: ...

    def get_state(self) -> BodyState: ...

    def set_kinematic(self, i: int, xc) -> None: ...

    def contact_forces(self) -> np.ndarray: ...
