"""Autotuner for the coloured solvers: colour classes, substeps, iterations and tile shape.

Searches colours in {8, 16, 32}, substeps in {1, 2, 4}, velocity iterations in {8, 16, 32} and bodies
per tile in {128, 192, 224} (position iterations follow as ceil(5/8 * vit), the 16 -> 10 ratio of the
pile workload) in order of
increasing solve cost (substeps * vit * colours), and returns the first configuration whose trial keeps
the maximum penetration at or below the target (1e-3 m by default); among configurations of equal
cost the faster measured one wins. Each trial steps a copy of the caller's scene for ``steps`` steps
from the same initial state, on the GPU, with the tiled solve: the first half is timed (synchronised
wall clock, no read-back inside), the second half gives the maximum penetration, reduced on the
device after every step.

The result's ``as_kwargs()`` are the engine constructor's arguments; the chosen substep count is not
a constructor argument (the protocol takes it in ``step(dt, substeps)``) and is ``result.substeps``.

This component has no counterpart in the reference (its colour count is an output of the JP
colouring, contact_engine_gpu_colored.py:1178-1208, and its iteration counts are constructor
arguments); it is new work on top of the reference solvers.
"""
from __future__ import annotations

import itertools
import math
import time
from dataclasses import dataclass, field

import numpy as np

from .bodies import BodySet

COLOURS = (8, 16, 32)
SUBSTEPS = (1, 2, 4)
ITERS = (8, 16, 32)


@dataclass
class Trial:
    colours: int
    substeps: int
    vit: int
    pit: int
    tile_bodies: int
    max_pen: float
    ms_per_step: float
    ok: bool


@dataclass
class TuneResult:
    best: Trial | None
    trials: list = field(default_factory=list)

    def as_kwargs(self):
        """Constructor arguments of the chosen configuration; step it with ``substeps``."""
        b = self.best
        return dict(n_colors=b.colours, vit=b.vit, pit=b.pit, tile_bodies=b.tile_bodies) if b else {}

    @property
    def substeps(self):
        return self.best.substeps if self.best else None


def _max_penetration(engine) -> float:
    """Deepest contact of the engine's last substep, reduced on the device."""
    from . import _lib
    n = engine.refresh_counters()["n_solved"]
    if n <= 0:
        return 0.0
    return float(engine.buffer_tensor(_lib.BUF_PEN, np.float32, n).max().item())


def run_trial(make_engine, bodies: BodySet, colours, substeps, vit, pit, tile_bodies, steps, dt, warm_state=None):
    import torch
    e = make_engine(n_colors=colours, vit=vit, pit=pit, tile_bodies=tile_bodies)
    e.add_boxes(bodies)
    e._build()
    if warm_state is not None:
        with torch.no_grad():
            for name, arr in zip(("x", "q", "v", "w"), warm_state):
                getattr(e, name).copy_(torch.from_numpy(np.ascontiguousarray(arr)).to(e.x.device))
    half = max(1, steps // 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(half):
        e.step(dt, substeps)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / half
    worst = 0.0
    for _ in range(max(1, steps - half)):
        e.step(dt, substeps)
        worst = max(worst, _max_penetration(e))
    return worst, ms


def autotune(make_engine, bodies: BodySet, target_pen=1e-3, steps=40, dt=1 / 240, warm_state=None,
             colours=COLOURS, substeps=SUBSTEPS, iters=ITERS, tile_bodies=(128, 192, 224), log=None) -> TuneResult:
    """make_engine(**kw) -> a fresh tiled coloured engine (kw: n_colors, vit, pit, tile_bodies)."""
    cands = []
    for c, s, v, tb in itertools.product(colours, substeps, iters, tile_bodies):
        p = max(1, math.ceil(v * 10 / 16))
        cands.append((s * (v + p) * c, c, s, v, p, tb))
    cands.sort()
    res = TuneResult(best=None)
    best_cost = None
    for cost, c, s, v, p, tb in cands:
        if best_cost is not None and cost > best_cost:
            break
        pen, ms = run_trial(make_engine, bodies, c, s, v, p, tb, steps, dt, warm_state)
        t = Trial(c, s, v, p, tb, pen, ms, pen <= target_pen)
        res.trials.append(t)
        if log:
            log(f"[autotune] colours={c} S={s} V={v} P={p} tile={tb}: max pen {pen:.2e} m, {ms:.3f} ms/step "
                f"{'ok' if t.ok else 'over target'}")
        if t.ok and (res.best is None or (cost == best_cost and ms < res.best.ms_per_step)):
            res.best = t
            best_cost = cost
    return res
