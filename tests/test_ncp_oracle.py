"""The NCP-level sensitivity oracle (oracle/ncp_oracle.py) against what pins it, CPU only.

- The reference's all-slip analytic scenes (tests/test_ncp_batch_adjoint.py:34-53): every contact
  slips, the PGS reaches the natural-residual floor recorded in data/ncp/adjoint_fix_table.csv, and
  the condition number of the assembled active-set Jacobian equals the recorded cond_dFdlam.
- Central finite differences of the PGS solution (the reference's own criterion,
  tests/test_ncp_reference.py:316-350: relative error below 1e-5) on random non-degenerate scenes.
- The diagonal one-contact slip case in closed form: lam = (1, -mu, 0), dlam/dmu = (0, -1, 0).
- The hyperstatic 4-point cube (data/ncp/scenes/00.json): singular active-set Jacobian, derivatives
  withheld as NaN, status REDUNDANT_CONTACTS (tests/test_ncp_reference.py:353-370)."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import ncp_oracle as R

GOLD = Path(__file__).resolve().parent / "golden"


def _rel(a, f):
    return float(np.max(np.abs(a - f)) / max(float(np.max(np.abs(f))), 1e-30))


@pytest.mark.parametrize("n_c", [1, 2, 4, 8])
def test_all_slip_scenes_match_the_reference_table(n_c):
    row = json.loads((GOLD / "ncp_adjoint_table.json").read_text())["all_slip"][f"all_slip_{n_c}"]
    G, b, mu = R.all_slip_scene(n_c)
    lam, hist, labels = R.solve_ncp_pgs(G, b, mu, 2000, 1e-15)
    assert list(labels) == ["slip"] * n_c and row["labels"] == f"slip={n_c}"
    res = R.natural_residual(lam, G, b, mu)
    assert res <= 2.0 * row["natural_residual"] + 1e-16
    dm, db, info = R.sensitivity(lam, G, b, mu, labels)
    assert not info["singular"] and not info["boundary"]
    assert info["cond"] == pytest.approx(row["cond_dFdlam"], rel=1e-9)


@pytest.mark.parametrize("seed", [7])
def test_sensitivity_matches_central_differences(seed):
    G, b, mu = R.random_scene(4, 2, seed)
    lam, hist, labels = R.solve_ncp_pgs(G, b, mu, 5000, 1e-14)
    assert hist[-1] < 1e-13
    dm, db, info = R.sensitivity(lam, G, b, mu, labels)
    assert not info["singular"] and not info["boundary"]
    fm, fb = R.sensitivity_fd(G, b, mu, lam, h_mu=1e-6, h_b=1e-6)
    assert _rel(dm, fm) < 1e-5
    assert _rel(db, fb) < 1e-5


def test_diagonal_slip_closed_form():
    G = np.eye(3)
    b = np.array([-1.0, 2.0, 0.0])
    mu = np.array([0.3])
    lam, _, labels = R.solve_ncp_pgs(G, b, mu, 5000, 1e-14)
    assert list(labels) == ["slip"]
    np.testing.assert_allclose(lam, [1.0, -0.3, 0.0], atol=1e-12)
    dm, db, _ = R.sensitivity(lam, G, b, mu, labels)
    np.testing.assert_allclose(dm[:, 0], [0.0, -1.0, 0.0], atol=1e-12)
    # u_n = lam_n + b_n = 0 -> dlam_n/db_n = -1; lam_t = -mu lam_n e1 -> dlam_t1/db_n = mu;
    # the slip direction turns with b_t2: lam_t2 = -mu lam_n u_t2 / |u_t| with u_t2 = lam_t2 + b_t2,
    # so dlam_t2/db_t2 = -mu lam_n / (|u_t| + mu lam_n), |u_t| = 2 - 0.3 = 1.7
    np.testing.assert_allclose(db[:, 0], [-1.0, 0.3, 0.0], atol=1e-12)
    assert db[2, 2] == pytest.approx(-0.3 / (1.7 + 0.3), rel=1e-12)


def test_hyperstatic_cube_is_flagged_not_hidden():
    z = np.load(GOLD / "ncp_scenes.npz")
    G, b, mu = z["s00_G"], z["s00_b"], z["s00_mu"]
    assert np.linalg.matrix_rank(G, tol=1e-10 * np.abs(G).max()) == 6 < 3 * len(mu)
    lam, hist, labels = R.solve_ncp_pgs(G, b, mu, 3000, 1e-13)
    dm, db, info = R.sensitivity(lam, G, b, mu, labels)
    assert info["singular"] and info["status"] == "REDUNDANT_CONTACTS"
    assert np.isnan(dm).all() and np.isnan(db).all()
