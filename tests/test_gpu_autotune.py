"""The autotuner on a small batch of piles, on the GPU: it tries configurations in order of solve cost
(substeps x (vit + pit) x colours), stops at the first cost level whose trial meets the penetration
target, returns the fastest trial of that level, and reports substeps apart from the constructor
arguments; with a target no configuration meets, it returns no configuration and every trial."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _setup():
    pkg = importlib.import_module("3fold-motion-engine_amd")
    at = importlib.import_module("3fold-motion-engine_amd.autotune")
    bodies, _ = pkg.scenes.batched_piles(6, lo=16, hi=96, seed=4)
    mk = lambda **k: pkg.ColorCacheContactEngine(tiled=True, check_every_step=False, **k)
    warm = mk(n_colors=16, vit=16, pit=10)
    warm.add_boxes(bodies)
    for _ in range(60):
        warm.step(1 / 240, 2)
    return pkg, at, bodies, mk, warm.state_numpy()


GRID = dict(colours=(8, 16), substeps=(1, 2), iters=(8, 16), tile_bodies=(128,))


def test_autotuner_meets_a_reachable_target_at_the_lowest_cost():
    _say("autotune: 6 piles, reachable target")
    pkg, at, bodies, mk, st = _setup()
    res = at.autotune(mk, bodies, target_pen=0.2, steps=8, warm_state=st, **GRID)
    assert res.best is not None and res.best.max_pen <= 0.2
    cost = lambda t: t.substeps * (t.vit + t.pit) * t.colours
    tried = [cost(t) for t in res.trials]
    assert tried == sorted(tried), "trials not in order of cost"
    assert cost(res.best) == min(cost(t) for t in res.trials if t.ok)
    assert all(not t.ok for t in res.trials if cost(t) < cost(res.best))
    same_level = [t for t in res.trials if cost(t) == cost(res.best) and t.ok]
    assert res.best.ms_per_step == min(t.ms_per_step for t in same_level)
    kw = res.as_kwargs()
    assert set(kw) == {"n_colors", "vit", "pit", "tile_bodies"} and res.substeps == res.best.substeps
    e = mk(**kw)
    e.add_boxes(bodies)
    e.step(1 / 240, res.substeps)
    assert e.refresh_counters()["overflow"] == 0


def test_autotuner_reports_an_unreachable_target():
    _say("autotune: 6 piles, unreachable target")
    pkg, at, bodies, mk, st = _setup()
    res = at.autotune(mk, bodies, target_pen=1e-9, steps=4, warm_state=st, **GRID)
    assert res.best is None and res.as_kwargs() == {} and res.substeps is None
    assert len(res.trials) == 8 and all(t.max_pen > 1e-9 for t in res.trials)


def test_trial_penetration_is_bit_reproducible():
    _say("autotune: one trial twice")
    pkg, at, bodies, mk, st = _setup()
    a = at.run_trial(mk, bodies, 8, 1, 8, 5, 128, 6, 1 / 240, st)
    b = at.run_trial(mk, bodies, 8, 1, 8, 5, 128, 6, 1 / 240, st)
    assert a[0] == b[0]
