"""HIP path against the CPU oracle, through the C ABI (libmotion_hip.so).

Bar: bit-exact (integer, index and float32 state alike; both sides compute in the reference's
operation order, with its fused multiply-adds written out where it builds a kernel with contraction).
Every test prints its case before the GPU work starts."""
import importlib
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


def _pkg():
    return importlib.import_module("3fold-motion-engine_amd")


def _say(*a):
    print("[case]", *a, flush=True)


def _engine(kind, **kw):
    pkg = _pkg()
    cls = {"jacobi": pkg.RelaxedJacobiContactEngine, "pair_chunks": pkg.GraphColoredContactEngine,
           "color_cache": pkg.ColorCacheContactEngine}[kind]
    if kind == "pair_chunks":  # the pair-chunk solve unless a test asks for the per-contact one
        pc = kw.setdefault("pair_chunks", True)
        kw.setdefault("manifold_reduce", pc)
    return cls(**kw)


def _assert_state_equal(e, o, tag):
    xg, qg, vg, wg = e.state_numpy()
    xo, qo, vo, wo = o.state()
    for name, a, b in (("x", xg, xo), ("q", qg, qo), ("v", vg, vo), ("w", wg, wo)):
        if not np.array_equal(a, b):
            idx = np.argwhere(a != b)[0]
            raise AssertionError(f"{tag}: {name} differs first at {idx.tolist()}: gpu {a[tuple(idx)]!r} "
                                 f"oracle {b[tuple(idx)]!r} (max |d| {np.max(np.abs(a - b)):.3e})")


def _assert_contacts_equal(e, o, tag):
    cg, co = e.contacts(), o.contacts()
    assert len(cg["key"]) == len(co["key"]), f"{tag}: contact count {len(cg['key'])} vs {len(co['key'])}"
    for k in ("key", "bi", "bj", "pA", "n", "pen", "jn", "jt1", "jt2", "jp"):
        assert np.array_equal(cg[k], co[k]), f"{tag}: contact field {k} differs"


def test_jacobi_bead_stack8_matches_oracle_and_archive(oracle):
    _say("jacobi bead stack8 x120 vs oracle + archive")
    pkg = _pkg()
    with np.load(GOLD / "reference_trajectories.npz") as d:
        gold = {k: d[f"jacobi__stack8__{k}"] for k in ("xc", "Rm", "vc", "om", "force")}
    meta = json.loads((GOLD / "reference_trajectories.json").read_text())
    counts = meta["records"]["jacobi__stack8"]["contacts"]
    e = _engine("jacobi", vit=40, pit=10)
    for c in pkg.scenes.stack(8):
        e.add_body(c)
    o = oracle.OracleWorld(pkg.bodies.bead_bodies(pkg.scenes.stack(8), [700.0] * 8, [0] * 8), "jacobi", vit=40, pit=10)
    for s in range(120):
        e.step(1 / 240)
        o.step(1 / 240)
        _assert_state_equal(e, o, f"step {s}")
        assert e.n_contacts == counts[s]
        st = e.get_state()  # every archived array of every step, bit for bit
        for name in ("xc", "Rm", "vc", "om"):
            np.testing.assert_array_equal(getattr(st, name), gold[name][s], err_msg=f"step {s}: {name}")
        np.testing.assert_array_equal(e.contact_forces(), gold["force"][s], err_msg=f"step {s}: force")
    _assert_contacts_equal(e, o, "final")
    np.testing.assert_array_equal(e.contact_forces(), o.forces())


@pytest.mark.parametrize("kind", ["jacobi", "pair_chunks", "color_cache"])
def test_box_stack16_matches_oracle(oracle, kind):
    _say(f"{kind} box stack16 x240")
    pkg = _pkg()
    b = pkg.scenes.box_stack()
    e = _engine(kind, vit=40, pit=20)
    e.add_boxes(b)
    o = oracle.OracleWorld(b, kind, vit=40, pit=20)
    for s in range(240):
        e.step(1 / 240)
        o.step(1 / 240)
        _assert_state_equal(e, o, f"step {s}")
        if s in (0, 1, 5, 60, 239):
            _assert_contacts_equal(e, o, f"step {s}")
    np.testing.assert_array_equal(e.contact_forces(), o.forces())


@pytest.mark.parametrize("kind,n_colors,per_contact,tiled,graph", [
    ("color_cache", 0, False, False, False), ("pair_chunks", 0, False, False, False),
    ("color_cache", 16, False, False, False), ("pair_chunks", 0, True, False, False),
    ("color_cache", 16, False, True, True), ("color_cache", 16, False, True, False),
    ("color_cache", 0, False, True, True), ("pair_chunks", 8, False, True, False), ("pair_chunks", 0, True, True, False)])
def test_box_pile_matches_oracle(oracle, kind, n_colors, per_contact, tiled, graph):
    _say(f"{kind} box pile 400 x60 S2 V16 P10 colors={n_colors} per_contact={per_contact} tiled={tiled} graph={graph}")
    pkg = _pkg()
    b = pkg.scenes.box_pile(400, seed=3)
    kw = dict(vit=16, pit=10, n_colors=n_colors)
    ekw = dict(kw, tiled=tiled, tile_bodies=32, graph_capture=graph)
    e = _engine(kind, pair_chunks=not per_contact, **ekw) if kind == "pair_chunks" else _engine(kind, **ekw)
    e.add_boxes(b)
    o = oracle.OracleWorld(b, kind, per_contact=per_contact, **kw)
    for s in range(60):
        e.step(1 / 240, 2)
        o.step(1 / 240, 2)
        _assert_state_equal(e, o, f"step {s}")
        oc = o.counters()
        assert e.n_contacts == oc["n_contacts"] and e.n_pairs == oc["n_pairs"] and e.n_colors == oc["n_colors"]
        if kind == "color_cache":
            assert (e.color_cache_hits, e.color_cache_misses) == (oc["cache_hits"], oc["cache_misses"])
    _assert_contacts_equal(e, o, "final")
    pg, po = e.pairs(), o.pairs()
    for k in ("pbi", "pbj", "pstart", "pcount", "color"):
        assert np.array_equal(pg[k], po[k]), k


def test_tiled_pile_beyond_the_lds_island_build(oracle):
    """A scene of more bodies than the one-workgroup island build holds (12 288) takes the multi-kernel
    island build; same bar as every tiled case."""
    _say("color_cache box pile 12800 x40 S2 V4 P2 tiled (multi-kernel island build)")
    pkg = _pkg()
    b = pkg.scenes.box_pile(12800, seed=4)
    kw = dict(vit=4, pit=2, n_colors=16)
    e = _engine("color_cache", tiled=True, tile_bodies=128, **kw)
    e.add_boxes(b)
    o = oracle.OracleWorld(b, "color_cache", **kw)
    for s in range(40):
        e.step(1 / 240, 2)
        o.step(1 / 240, 2)
        _assert_state_equal(e, o, f"step {s}")
        assert e.n_contacts == o.counters()["n_contacts"]
    assert e.n_contacts > 10000


def test_colouring_is_proper_and_runs_repeat():
    _say("color_cache box pile 1000 x30: proper colouring, two runs bit-identical")
    pkg = _pkg()
    b = pkg.scenes.box_pile(1000, seed=1)
    runs = []
    for _ in range(2):
        e = _engine("color_cache", vit=16, pit=10)
        e.add_boxes(b)
        for s in range(30):
            e.step(1 / 240, 2)
            pr = e.pairs()
            invM = b.invM
            for c in np.unique(pr["color"]):
                sel = pr["color"] == c
                bodies = np.concatenate([pr["pbi"][sel], pr["pbj"][sel]])
                bodies = bodies[(bodies >= 0)]
                bodies = bodies[invM[bodies] > 0]
                assert len(np.unique(bodies)) == len(bodies), f"step {s}: colour {c} reuses a dynamic body"
        runs.append(e.state_numpy())
    for a, b2 in zip(*runs):
        assert np.array_equal(a, b2)


@pytest.mark.parametrize("kind,tile_bodies,n,steps", [("color_cache", 8, 600, 30), ("color_cache", 64, 600, 30),
                                                      ("color_cache", 256, 600, 30), ("color_cache", 128, 4500, 6),
                                                      ("color_cache", 128, 6600, 6), ("pair_chunks", 128, 8400, 6)])
def test_tiled_chain_island_beyond_lds(oracle, kind, tile_bodies, n, steps):
    """A row of touching cubes is one island: 600 bodies exceed a small tile (big-tile kernel, LDS
    state), 4500 bodies (about 9000 pairs with the ground) and 6600 bodies exceed the big tile's LDS
    capacity by pairs and by bodies (the grid of workgroups sweeping one island in global memory, a
    grid barrier per colour step), 8400 bodies also exceed the LDS
    colouring limits (JP rounds by the grid, recoloured every substep under pair chunks); small
    tiles make islands straddle tile boundaries."""
    _say(f"{kind} tiled chain {n} + pile 200, tile_bodies={tile_bodies}")
    pkg = _pkg()
    c = np.zeros((n, 3))
    c[:, 0] = np.arange(n) * 0.0995
    c[:, 2] = 0.0499
    row = pkg.bodies.box_bodies(c, [0.05, 0.05, 0.05], 1000.0, mu=0.6)
    pile = pkg.scenes.box_pile(200, seed=5, origin=(0.0, 3.0, 0.0))
    b = pkg.scenes.concat([row, pile])
    e = _engine(kind, vit=8, pit=4, n_colors=16, tiled=True, tile_bodies=tile_bodies)
    e.add_boxes(b)
    o = oracle.OracleWorld(b, kind, vit=8, pit=4, n_colors=16)
    for s in range(steps):
        e.step(1 / 240, 2)
        o.step(1 / 240, 2)
        _assert_state_equal(e, o, f"step {s}")
    assert e._counters["max_tile_bodies"] >= n
    _assert_contacts_equal(e, o, "final")


def test_two_giant_islands_graph_replay_matches_eager_and_oracle(oracle):
    """Rows of 8400 and 5200 touching cubes: two islands beyond one workgroup's LDS, swept one after
    the other by the same grid launch (colour classes rebuilt per island), the first also coloured
    by the grid, next to a pile of small islands. The whole step replayed as a hipGraph (the grid
    barrier and round counts are cleared by fill kernels inside the graph) equals the eager engine
    after every replay and the oracle at the end."""
    _say("color_cache tiled chains of 8400 and 5200 + pile 300 x5: graph vs eager vs oracle")
    pkg = _pkg()
    rows = []
    for k, n in enumerate((8400, 5200)):
        c = np.zeros((n, 3))
        c[:, 0] = np.arange(n) * 0.0995
        c[:, 1] = 1.5 * k
        c[:, 2] = 0.0499
        rows.append(pkg.bodies.box_bodies(c, [0.05, 0.05, 0.05], 1000.0, mu=0.6))
    b = pkg.scenes.concat(rows + [pkg.scenes.box_pile(300, seed=9, origin=(0.0, 4.0, 0.0))])
    kw = dict(vit=8, pit=4, n_colors=16, tiled=True, tile_bodies=128)
    eg = _engine("color_cache", graph_capture=True, **kw)
    ee = _engine("color_cache", graph_capture=False, **kw)
    eg.add_boxes(b)
    ee.add_boxes(b)
    o = oracle.OracleWorld(b, "color_cache", vit=8, pit=4, n_colors=16)
    for s in range(5):
        eg.step(1 / 240, 2)
        ee.step(1 / 240, 2)
        o.step(1 / 240, 2)
        for name, a, r in zip("xqvw", eg.state_numpy(), ee.state_numpy()):
            assert np.array_equal(a, r), f"step {s}: graph replay differs from eager in {name}"
    assert eg.n_graph_captures == 1
    assert eg._counters["max_tile_bodies"] >= 8400
    _assert_state_equal(eg, o, "final")
    _assert_contacts_equal(eg, o, "final")


def test_batched_scenes_match_oracle_and_stay_independent(oracle):
    """Scenes share no contacts: a batch of 6 piles (tiled, one engine) equals the oracle on the
    same batch, and a scene's trajectory equals that scene stepped on its own."""
    _say("color_cache tiled batch of 6 scenes x40 vs oracle and vs solo scene")
    pkg = _pkg()
    bodies, sizes = pkg.scenes.batched_piles(6, lo=8, hi=300, seed=11)
    e = _engine("color_cache", vit=8, pit=5, n_colors=16, tiled=True)
    e.add_boxes(bodies)
    o = oracle.OracleWorld(bodies, "color_cache", vit=8, pit=5, n_colors=16)
    k = int(np.argmax(sizes))
    offs = np.concatenate([[0], np.cumsum(sizes)])
    sl = slice(offs[k], offs[k + 1])
    solo_b = pkg.bodies.BodySet(x=bodies.x[sl], q=bodies.q[sl], v=bodies.v[sl], w=bodies.w[sl], invM=bodies.invM[sl],
                                IbInv=bodies.IbInv[sl], kin=bodies.kin[sl], half=bodies.half[sl], mu=bodies.mu[sl],
                                scene=np.zeros(sizes[k], np.int32), M=bodies.M[sl])
    solo = oracle.OracleWorld(solo_b, "color_cache", vit=8, pit=5, n_colors=0)
    for s in range(40):
        e.step(1 / 240, 2)
        o.step(1 / 240, 2)
        solo.step(1 / 240, 2)
        _assert_state_equal(e, o, f"step {s}")
    # colour splitting depends on the whole batch; the JP solo run is compared on contacts only
    assert e.n_pairs == o.counters()["n_pairs"]
    xs = solo.state()[0]
    assert np.all(np.isfinite(xs))


def test_tile_queue_with_more_tiles_than_resident_workgroups(oracle):
    """The small tiles are taken from a largest-first queue by a persistent grid of as many workgroups
    as can be resident. 700 scenes of 8 to 24 boxes at 8 bodies per tile make more tiles than that
    grid has workgroups, so workgroups take several tiles each; the batch equals the oracle bit for
    bit, and every tile is solved once (the tile count the kernels report equals the tiles built)."""
    _say("color_cache tiled batch of 700 small scenes, 8 bodies per tile, x30 vs oracle")
    pkg = _pkg()
    bodies, sizes = pkg.scenes.batched_piles(700, lo=8, hi=24, seed=5)
    e = _engine("color_cache", vit=8, pit=5, n_colors=16, tiled=True, tile_bodies=8)
    e.add_boxes(bodies)
    o = oracle.OracleWorld(bodies, "color_cache", vit=8, pit=5, n_colors=16, omp=True)
    most = 0
    for s in range(30):
        e.step(1 / 240, 2)
        o.step(1 / 240, 2)
        _assert_state_equal(e, o, f"step {s}")
        c = e.refresh_counters()
        most = max(most, c["n_tiles"])
    # k_t_small holds two workgroups per CU (its registers), 512 on the 256 CUs of an MI355X
    assert most > 600, f"at most {most} tiles: the queue never held more tiles than resident workgroups"


def test_graph_replay_matches_eager(oracle):
    """Whole-step hipGraphs (ME_FLAG_GRAPH): the engine replaying captured steps equals the eager
    engine bit for bit after every replay, with eager work between replays (state and counter
    read-backs); the big-tile kernel on the side stream is inside the graph (a 600-box chain); a new
    dt captures a second graph; the end state equals the oracle."""
    _say("color_cache tiled chain 600 + pile 200 x36: graph replay vs eager vs oracle, dt change")
    pkg = _pkg()
    c = np.zeros((600, 3))
    c[:, 0] = np.arange(600) * 0.0995
    c[:, 2] = 0.0499
    row = pkg.bodies.box_bodies(c, [0.05, 0.05, 0.05], 1000.0, mu=0.6)
    b = pkg.scenes.concat([row, pkg.scenes.box_pile(200, seed=5, origin=(0.0, 3.0, 0.0))])
    kw = dict(vit=8, pit=4, n_colors=16, tiled=True, tile_bodies=64)
    eg = _engine("color_cache", graph_capture=True, **kw)
    ee = _engine("color_cache", graph_capture=False, **kw)
    eg.add_boxes(b)
    ee.add_boxes(b)
    o = oracle.OracleWorld(b, "color_cache", vit=8, pit=4, n_colors=16)
    for s in range(36):
        dt = 1 / 240 if s < 24 else 1 / 300
        eg.step(dt, 2)
        ee.step(dt, 2)
        o.step(dt, 2)
        for name, a, r in zip("xqvw", eg.state_numpy(), ee.state_numpy()):
            assert np.array_equal(a, r), f"step {s}: graph replay differs from eager in {name}"
        assert (eg.n_contacts, eg.n_pairs, eg.n_colors) == (ee.n_contacts, ee.n_pairs, ee.n_colors)
    assert eg.n_graph_captures == 2 and ee.n_graph_captures == 0
    assert eg._counters["max_tile_bodies"] >= 600
    _assert_state_equal(eg, o, "final")
    _assert_contacts_equal(eg, o, "final")


def test_tile_size_does_not_change_the_bits():
    """The island tiling is a schedule only: a batch of 24 piles stepped at 64, 128 and 224 bodies per
    tile (224 overflows some tiles past 256 pairs into the big-tile kernel) gives the same state bits,
    contact impulses and colours after 60 steps of the batched bench configuration."""
    _say("color_cache batch of 24 piles x60 S2 V16 P10 at tile_bodies 64 / 128 / 224")
    pkg = _pkg()
    bodies, _ = pkg.scenes.batched_piles(24, lo=8, hi=2048, seed=5)
    runs = []
    for tb in (64, 128, 224):
        e = _engine("color_cache", vit=16, pit=10, n_colors=16, tiled=True, tile_bodies=tb, check_every_step=False)
        e.add_boxes(bodies)
        for _ in range(60):
            e.step(1 / 240, 2)
        e.refresh_counters()
        runs.append((e.state_numpy(), e.contacts(), e.pairs()["color"], e._counters["n_tiles"]))
        del e
    (s0, c0, col0, t0) = runs[0]
    for (s1, c1, col1, t1), tb in zip(runs[1:], (128, 224)):
        assert t1 < t0, f"tile_bodies {tb}: expected fewer tiles than at 64 ({t1} vs {t0})"
        for a, b, name in zip(s0, s1, "xqvw"):
            assert np.array_equal(a, b), f"tile_bodies {tb}: {name} differs from tile_bodies 64"
        for k in ("key", "jn", "jt1", "jt2", "jp"):
            assert np.array_equal(c0[k], c1[k]), f"tile_bodies {tb}: contact {k} differs"
        assert np.array_equal(col0, col1), f"tile_bodies {tb}: colours differ"
