"""The CPU oracle's coloured bead path: the colouring of the reference's archived lattice and
planted mistakes the archive rejects.

The archive (reports/position_solve_matrix/inputs/cross_hardware_l4.npz, fixtures extracted by
tests/golden/make_golden.py) is matched bit for bit by the oracle's pair-chunk and ColorCache
engines over every archived step (test_oracle_golden.py). What this file adds: the lattice's
colouring (every column is a 2-pair chain whose colour order the PCG-hash priority decides), and
that a wrong hash seed or a wrong tie-break for the deepest point of the manifold reduction moves
the archived velocities by orders of magnitude more than any rounding."""
import importlib
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"
ORACLE_SRC = Path(__file__).resolve().parent.parent / "oracle" / "oracle.c"
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


def _world(oracle, centres, engine):
    bodies = importlib.import_module("3fold-motion-engine_amd.bodies")
    n = len(centres)
    b = bodies.bead_bodies(centres, [700.0] * n, [0] * n)
    return oracle.OracleWorld(b, engine, vit=40, pit=10)


def test_lattice1000_colouring(pkg, oracle, gold):
    """At step 8 all 1000 bodies are in contact pairs, two per column, and the columns need two
    colours; the pair count and colour count equal those of the reference's colouring."""
    meta, _ = gold
    o = _world(oracle, lattice(1000), "pair_chunks")
    rec = meta["records"]["pair_chunks__lattice1000"]
    for s in range(9):
        o.step(1 / 240, 1)
        assert o.counters()["n_generated"] == rec["contacts"][s]
    c = o.counters()
    assert c["n_pairs"] == 1000 and c["n_colors"] == 2
    assert c["n_contacts"] <= 4 * c["n_pairs"]


def _planted(tmp_path, flag):
    lib = tmp_path / f"liboracle_{flag.strip('-D').replace('=', '_')}.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-shared", flag, "-o", str(lib),
                    str(ORACLE_SRC), "-lm"], check=True)
    return lib


@pytest.mark.parametrize("flag,step", [("-DOR_JP_SEED=7920", 9), ("-DOR_PLANT_DEEP_LAST", 3)])
def test_archive_rejects_planted_mistakes(pkg, oracle, gold, tmp_path, monkeypatch, flag, step):
    """A wrong seed of the colouring hash (colour order of the lattice's 500 two-pair chains) and a
    wrong tie-break for the deepest point of the reduction both move the lattice velocities by
    more than 1e-3 m/s at the first archived step they reach; the true oracle stays within 1e-6."""
    _, arr = gold
    lib = _planted(tmp_path, flag)
    monkeypatch.setattr(oracle, "LIB", lib)
    monkeypatch.setattr(oracle, "_libs", {})
    o = _world(oracle, lattice(1000), "pair_chunks")
    for _ in range(step + 1):
        o.step(1 / 240, 1)
    steps = list(arr["pair_chunks__lattice1000__steps"])
    _, _, v, _ = o.state()
    del o  # freed by the planted library while it is still the loaded one
    assert np.max(np.abs(v - arr["pair_chunks__lattice1000__vc"][steps.index(step)])) > 1e-3
