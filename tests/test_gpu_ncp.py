"""The batched NCP sensitivity on the HIP path (me_ncp_sensitivity, float64) against the CPU oracle
(oracle/ncp_oracle.py, the restatement of ncp_ref.classify / _active_system / sensitivity).

Bar: labels equal; derivatives within 1e-9 (absolute, relative to the largest entry) of the oracle,
the bound the reference holds its own batched GPU adjoint to against the CPU implicit-function
reference on the all-slip scenes (tests/test_ncp_batch_adjoint.py:21, 120-123); singular systems
flagged and NaN, as the oracle. Every test prints its case before the GPU work starts."""
import importlib
from pathlib import Path

import numpy as np
import pytest

from oracle import ncp_oracle as R

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def _ncp():
    return importlib.import_module("3fold-motion-engine_amd.ncp")


def _say(*a):
    print("[case]", *a, flush=True)


def _close(gpu, ref, tol=1e-9):
    scale = max(float(np.max(np.abs(ref))), 1.0)
    err = float(np.max(np.abs(gpu - ref)))
    assert err <= tol * scale, f"max |gpu - oracle| {err:.3e} (scale {scale:.3e})"


@pytest.mark.parametrize("n_c", [1, 2, 4, 8])
def test_all_slip_scenes_match_oracle(n_c):
    _say(f"ncp all-slip n_c={n_c}")
    ncp = _ncp()
    G, b, mu = R.all_slip_scene(n_c)
    lam, _, labels = R.solve_ncp_pgs(G, b, mu, 2000, 1e-15)
    dm_o, db_o, info_o = R.sensitivity(lam, G, b, mu, labels)
    assert list(ncp.classify(lam, G, b, mu)) == list(labels)
    dm, db, info = ncp.sensitivity(lam, G, b, mu, labels)
    assert not info["singular"] and not info["boundary"]
    _close(dm, dm_o)
    _close(db, db_o)


@pytest.mark.parametrize("seed", [3, 7, 11])
def test_random_scenes_match_oracle(seed):
    _say(f"ncp random scene seed={seed}")
    ncp = _ncp()
    G, b, mu = R.random_scene(4, 2, seed)
    lam, _, labels = R.solve_ncp_pgs(G, b, mu, 5000, 1e-14)
    dm_o, db_o, _ = R.sensitivity(lam, G, b, mu, labels)
    D = np.random.default_rng(seed).standard_normal((12, 3))
    dm, db, info = ncp.sensitivity(lam, G, b, mu, labels, db_dparam=D)
    _close(dm, dm_o)
    _close(db, db_o @ D)


def test_diagonal_slip_closed_form():
    _say("ncp diagonal slip closed form")
    ncp = _ncp()
    G, b, mu = np.eye(3), np.array([-1.0, 2.0, 0.0]), np.array([0.3])
    lam = np.array([1.0, -0.3, 0.0])
    dm, db, info = ncp.sensitivity(lam, G, b, mu, ["slip"])
    np.testing.assert_allclose(dm[:, 0], [0.0, -1.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(db[:, 0], [-1.0, 0.3, 0.0], atol=1e-15)
    assert db[2, 2] == pytest.approx(-0.3 / 2.0, rel=1e-15)


def test_hyperstatic_cube_is_singular_and_nan():
    _say("ncp hyperstatic cube")
    ncp = _ncp()
    z = np.load(GOLD / "ncp_scenes.npz")
    G, b, mu = z["s00_G"], z["s00_b"], z["s00_mu"]
    lam, _, labels = R.solve_ncp_pgs(G, b, mu, 3000, 1e-13)
    dm, db, info = ncp.sensitivity(lam, G, b, mu, labels)
    assert info["singular"] and info["status"] == "REDUNDANT_CONTACTS"
    assert np.isnan(dm).all() and np.isnan(db).all()


def test_batch_of_4096_stance_problems_matches_oracle_and_is_env_independent():
    _say("ncp batch 4096 x 4 contacts, 2 directions")
    ncp = _ncp()
    import torch
    E = 4096
    G, b, mu, lam, labels = ncp.synthetic_problems(E, n_c=4, seed=5)
    D = np.random.default_rng(6).standard_normal((E, 12, 2))
    out = ncp.batch_sensitivity(lam, G, b, mu, directions=D)
    torch.cuda.synchronize()
    assert np.array_equal(out["labels"].cpu().numpy(), labels)
    assert int(out["status"].abs().sum()) == 0
    dmu, dlb = out["dlam_dmu"].cpu().numpy(), out["dlam_db"].cpu().numpy()
    for e in range(0, E, 64):
        labs = [R.LABELS[k] for k in labels[e]]
        dm_o, db_o, info_o = R.sensitivity(lam[e], G[e], b[e], mu[e], labs, db_dparam=D[e])
        assert not info_o["singular"]
        _close(dmu[e], dm_o)
        _close(dlb[e], db_o)
    # an environment alone gives the bytes it gives inside the batch
    for e in (0, 1234, 4095):
        one = ncp.batch_sensitivity(lam[e:e + 1], G[e:e + 1], b[e:e + 1], mu[e:e + 1], directions=D[e:e + 1])
        assert one["dlam_db"].cpu().numpy().tobytes() == dlb[e:e + 1].tobytes()
        assert one["dlam_dmu"].cpu().numpy().tobytes() == dmu[e:e + 1].tobytes()


def test_large_system_uses_global_scratch():
    _say("ncp 40-contact problem (global-memory factorisation)")
    ncp = _ncp()
    G, b, mu, lam, labels = ncp.synthetic_problems(3, n_c=40, nv=160, seed=9)
    out = ncp.batch_sensitivity(lam, G, b, mu)
    dmu, dlb = out["dlam_dmu"].cpu().numpy(), out["dlam_db"].cpu().numpy()
    for e in range(3):
        labs = [R.LABELS[k] for k in labels[e]]
        dm_o, db_o, info_o = R.sensitivity(lam[e], G[e], b[e], mu[e], labs)
        _close(dmu[e], dm_o, 1e-8)
        _close(dlb[e], db_o, 1e-8)


def _corner_boxes():
    """Two boxes tilted about x and y so that one corner each touches the ground, 1 mm deep, with
    small horizontal velocities and high friction: one sticking contact per box, a non-redundant
    contact system."""
    pkg = importlib.import_module("3fold-motion-engine_amd")
    h = np.array([0.1, 0.08, 0.06])
    ax, ay = 0.35, 0.28
    qx = np.array([np.sin(ax / 2), 0, 0, np.cos(ax / 2)])
    qy = np.array([0, np.sin(ay / 2), 0, np.cos(ay / 2)])
    x1, y1, z1, w1 = qy
    x2, y2, z2, w2 = qx
    q = np.array([w1 * x2 + w2 * x1 + y1 * z2 - y2 * z1, w1 * y2 + w2 * y1 + z1 * x2 - z2 * x1,
                  w1 * z2 + w2 * z1 + x1 * y2 - x2 * y1, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2])
    Rm = pkg.bodies.quat_to_R_host(q[None].astype(np.float32))[0].astype(np.float64)
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * h
    zlow = np.sort((corners @ Rm.T)[:, 2])
    assert zlow[1] - zlow[0] > 0.01  # one lowest corner
    cz = -zlow[0] - 1e-3
    return pkg.bodies.box_bodies([[0.0, 0.0, cz], [1.0, 0.0, cz]], h, 500.0, quats=[q, q], mu=2.0,
                                 v=[[0.05, 0.02, 0.0], [-0.03, 0.04, 0.0]], w=[[0.1, 0.0, 0.0], [0.0, -0.2, 0.0]])


def test_contact_step_problem_and_rollout_adjoint_match_the_ift():
    """One ColorCache step of one substep (vit 300, no position sweeps) of two boxes on a corner each:
    the extracted problem (G, b, mu, lam) satisfies the stick branch to float32 rounding, and the
    hand-written rollout adjoint of x_T w.r.t. (v0, w0) equals the composition through the NCP
    sensitivity, dt (I + J^T S^T J M^-1) xbar with S = dlam/db, to 1e-3 relative (float32 solve)."""
    _say("ncp contact-step problem + rollout adjoint vs IFT (2 corner boxes)")
    import torch
    pkg = importlib.import_module("3fold-motion-engine_amd")
    ncp = _ncp()
    bs = _corner_boxes()
    dt = 1.0 / 240
    ro = pkg.adjoint.Rollout(bs, steps=1, dt=dt, substeps=1, vit=300, pit=0, n_colors=0, k=1)
    dev = ro.eng.x.device
    v0 = torch.from_numpy(bs.v).to(dev)
    w0 = torch.from_numpy(bs.w).to(dev)
    mu = torch.from_numpy(bs.mu).to(dev)
    ro.run(v0, w0, mu)
    prob = ncp.contact_problem(ro.eng)
    G, b, mu_c, lam = prob["G"], prob["b"], prob["mu"], prob["lam"]
    assert len(mu_c) == 2
    labels = ncp.classify(lam, G, b, mu_c)
    assert list(labels) == ["stick", "stick"]
    u = G @ lam + b
    assert np.max(np.abs(u)) <= 1e-5 * max(1.0, np.max(np.abs(b)))
    dm, S, info = ncp.sensitivity(lam, G, b, mu_c, labels)
    assert not info["singular"]
    dm_o, S_o, _ = R.sensitivity(lam, G, b, mu_c, list(labels))
    _close(S, S_o)
    xbar = np.random.default_rng(1).standard_normal((2, 3))
    vb, wb, mb = ro.grad(torch.from_numpy(xbar.astype(np.float32)).to(dev))
    J, Minv, bodies = prob["J"], prob["Minv"], prob["bodies"]
    X = np.zeros(J.shape[1])
    for k, body in enumerate(bodies):
        X[6 * k:6 * k + 3] = xbar[body]
    g = dt * (X + J.T @ (S.T @ (J @ (Minv @ X))))
    want = np.concatenate([np.concatenate([g[6 * k:6 * k + 3], g[6 * k + 3:6 * k + 6]]) for k in range(len(bodies))])
    got = np.concatenate([np.concatenate([vb[body].cpu().numpy(), wb[body].cpu().numpy()]) for body in bodies])
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"  rollout adjoint vs IFT composition: max rel err {err:.3e}", flush=True)
    assert err <= 1e-3
