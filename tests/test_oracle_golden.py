"""The CPU oracle against the reference's archived trajectories (golden fixtures extracted by
tests/golden/make_golden.py from reports/position_solve_matrix/inputs/cross_hardware_l4.npz).

The archive holds three engines on two scenes, vit 40, pit 10, dt 1/240, one substep
(probes/cross_position_state_probe.py:19-64): the relaxed Jacobi engine, the pair-chunk engine
(UnfusedPositionPairContactEngine: slip-frame chunks, colours recomputed every step) and the
ColorCache engine (UnfusedPositionColorCacheContactEngine: warm apply, fixed-frame chunks, colour
cache), the two coloured ones with the device manifold reduction to 4 points per body pair.

What these pin: the body build (byte-identical initial arrays, sha256), the bead world points and
bead contact generation (per-step contact counts), the Jacobi passes with int64 fixed-point
accumulation, the bead manifold reduction, the Jones-Plassmann colouring, the colour cache, warm
apply, both chunk sweeps, the integrator, the contact forces, and which kernels round with
contraction: every position, rotation, velocity and force of the 120-step stack matches the archive
bit for bit, and so does every per-step sha256 of the 1000-body lattice's positions, velocities and
angular velocities over its 40 steps. The planted test shows the archive telling the contraction
model apart: the same oracle with the integrator's quaternion update left unfused leaves the
archive within ulps."""
import hashlib
import importlib
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"
PITCH = 0.205


def stack(K):
    return [[0, 0, 0.101 + k * PITCH] for k in range(K)]


def lattice(n, layers=2):
    side = int(np.ceil(np.sqrt(n / layers)))
    out = []
    for lz in range(layers):
        for iy in range(side):
            for ix in range(side):
                out.append([ix * 0.35, iy * 0.35, 0.101 + lz * PITCH])
    return out[:n]


@pytest.fixture(scope="module")
def gold():
    meta = json.loads((GOLD / "reference_trajectories.json").read_text())
    with np.load(GOLD / "reference_trajectories.npz") as d:
        arrays = {k: d[k] for k in d.files}
    return meta, arrays


def _fp(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("scene,count", [("stack8", 8), ("lattice1000", 1000)])
def test_body_build_fingerprints(pkg, gold, scene, count):
    meta, _ = gold
    bodies = importlib.import_module("3fold-motion-engine_amd.bodies")
    cen = stack(count) if scene == "stack8" else lattice(count)
    b = bodies.bead_bodies(cen, [700.0] * count, [0] * count)
    init = meta["records"][f"jacobi__{scene}"]["initial"]
    for name, arr in dict(xc=b.x, q=b.q, v=b.v, w=b.w, invM=b.invM, IbInv=b.IbInv, rest=b.rest).items():
        assert _fp(arr) == init[name]["sha256"], name


ENGINES = ("jacobi", "pair_chunks", "color_cache")
ORACLE_SRC = Path(__file__).resolve().parent.parent / "oracle" / "oracle.c"


def _count(o, engine):
    c = o.counters()
    return c["n_contacts"] if engine == "jacobi" else c["n_generated"]


@pytest.mark.parametrize("engine", ENGINES)
def test_stack8_trajectory_bit_exact(pkg, oracle, gold, engine):
    meta, arr = gold
    bodies = importlib.import_module("3fold-motion-engine_amd.bodies")
    b = bodies.bead_bodies(stack(8), [700.0] * 8, [0] * 8)
    o = oracle.OracleWorld(b, engine, vit=40, pit=10)
    counts = []
    for s in range(120):
        o.step(1 / 240, 1)
        x, q, v, w = o.state()
        counts.append(_count(o, engine))
        np.testing.assert_array_equal(x, arr[f"{engine}__stack8__xc"][s], err_msg=f"step {s}")
        np.testing.assert_array_equal(bodies.quat_to_R_host(q), arr[f"{engine}__stack8__Rm"][s], err_msg=f"step {s}")
        np.testing.assert_array_equal(v, arr[f"{engine}__stack8__vc"][s], err_msg=f"step {s}")
        np.testing.assert_array_equal(w, arr[f"{engine}__stack8__om"][s], err_msg=f"step {s}")
        np.testing.assert_array_equal(o.forces(), arr[f"{engine}__stack8__force"][s], err_msg=f"step {s}")
    assert counts == meta["records"][f"{engine}__stack8"]["contacts"]


@pytest.mark.parametrize("engine", ENGINES)
def test_lattice1000_trajectory_bit_exact(pkg, oracle, gold, engine):
    meta, arr = gold
    bodies = importlib.import_module("3fold-motion-engine_amd.bodies")
    b = bodies.bead_bodies(lattice(1000), [700.0] * 1000, [0] * 1000)
    o = oracle.OracleWorld(b, engine, vit=40, pit=10)
    rec = meta["records"][f"{engine}__lattice1000"]
    steps = list(arr[f"{engine}__lattice1000__steps"])
    for s in range(40):
        o.step(1 / 240, 1)
        x, q, v, w = o.state()
        assert _count(o, engine) == rec["contacts"][s], f"step {s}"
        assert _fp(x) == rec["sha256_per_step"]["xc"][s], f"step {s}"
        assert _fp(v) == rec["sha256_per_step"]["vc"][s], f"step {s}"
        assert _fp(w) == rec["sha256_per_step"]["om"][s], f"step {s}"
        if s in steps:
            k = steps.index(s)
            np.testing.assert_array_equal(x, arr[f"{engine}__lattice1000__xc"][k])
            np.testing.assert_array_equal(v, arr[f"{engine}__lattice1000__vc"][k])
            np.testing.assert_array_equal(w, arr[f"{engine}__lattice1000__om"][k])


def test_archive_rejects_unfused_integrator(pkg, oracle, gold, tmp_path, monkeypatch):
    """The reference's integrator rounds q + (0.5*w*q)*dt with one rounding per component (a
    fused multiply-add); the planted oracle rounds the product first. The archive tells the two
    apart within the first 20 steps of the pair-chunk stack, while the difference stays at
    ulp level (under 1e-6 in every state array)."""
    _, arr = gold
    lib = tmp_path / "liboracle_unfused_integ.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-shared", "-DOR_PLANT_UNFUSED_INTEG",
                    "-o", str(lib), str(ORACLE_SRC), "-lm"], check=True)
    monkeypatch.setattr(oracle, "LIB", lib)
    monkeypatch.setattr(oracle, "_libs", {})
    bodies = importlib.import_module("3fold-motion-engine_amd.bodies")
    o = oracle.OracleWorld(bodies.bead_bodies(stack(8), [700.0] * 8, [0] * 8), "pair_chunks", vit=40, pit=10)
    first = None
    for s in range(20):
        o.step(1 / 240, 1)
        x, q, v, w = o.state()
        same = all(np.array_equal(a, arr[f"pair_chunks__stack8__{k}"][s]) for a, k in ((x, "xc"), (v, "vc"), (w, "om")))
        if not same:
            first = s
            assert np.max(np.abs(x - arr["pair_chunks__stack8__xc"][s])) < 1e-6
            break
    del o  # freed by the planted library while it is still the loaded one
    assert first is not None
