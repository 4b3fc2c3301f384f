"""The oracle's SVD rank guard (oracle/ncp_oracle.py: rank_factor, guarded_response, sensitivity,
contact_diagnostics) held to the contracts of the reference's redundant-contact tests
(tests/test_ncp_redundant_contacts.py) and to finite differences: a box on its four bottom corners is
redundant in every closed label pattern, its nullspace is orthonormal and annihilated by A, the
alternating corner normals are a left-null witness; the minimum-norm response is orthogonal to the
nullspace and incompatible columns stay NaN; the body wrench J^T lam is an invariant observable whose
response equals central differences of the PGS solution, while a spinning slip pattern is not
invariant; the boundary and near-singular guards hold. CPU only."""
import itertools

import numpy as np
import pytest

from oracle import ncp_oracle as R

BOXES = [(0.5, 0.5, 0.5, 1.0), (0.25, 0.75, 0.4, 2.0), (1.0, 0.2, 0.6, 0.5), (0.8, 0.6, 0.3, 3.0)]
PATTERNS = list(itertools.product(("stick", "slip"), repeat=4))


@pytest.mark.parametrize("dims", BOXES)
def test_every_closed_corner_pattern_is_redundant(dims):
    G, _, mu, _, _ = R.corner_box_scene(dims[:3], dims[3])
    for labels in PATTERNS:
        lam, b = R.algebraic_branch(G, mu, labels)
        dm, db, info = R.sensitivity(lam, G, b, mu, labels)
        assert info["status"] == "REDUNDANT_CONTACTS" and info["nullity"] >= 1
        assert np.isnan(dm).all() and np.isnan(db).all()
        A = R.active_system(lam, G, b, mu, labels)[0]
        N = info["nullspace"]
        assert np.linalg.norm(A @ N) < 1e-10 * np.linalg.norm(A)
        assert np.allclose(N.T @ N, np.eye(N.shape[1]), atol=1e-12)
        witness = np.zeros(12)
        witness[::3] = [1.0, -1.0, -1.0, 1.0]
        assert np.linalg.norm(witness @ A) < 1e-12 * np.linalg.norm(A)


def test_minimum_norm_is_opt_in_orthogonal_to_the_nullspace_and_guarded():
    G, _, mu, J, _ = R.corner_box_scene()
    labels = ["slip"] * 4
    lam, b = R.algebraic_branch(G, mu, labels)
    dm0, _, _ = R.sensitivity(lam, G, b, mu, labels)
    assert np.isnan(dm0).all()
    dm, db, info = R.sensitivity(lam, G, b, mu, labels, selection="minimum_norm", observable=J.T)
    assert np.isfinite(dm[:, info["mu_compatible"]]).all()
    assert np.isnan(db[:, ~info["b_compatible"]]).all()
    assert not info["b_compatible"][::3].any()
    assert np.linalg.norm(info["nullspace"].T @ dm[:, info["mu_compatible"]]) < 1e-12


def test_spinning_slip_wrench_is_not_invariant():
    G, _, mu, J, _ = R.corner_box_scene()
    lam, b = R.algebraic_branch(G, mu, ["slip"] * 4, spin=True)
    _, _, info = R.sensitivity(lam, G, b, mu, ["slip"] * 4, observable=J.T)
    assert info["singular"] and not info["observable_invariant"].all()
    assert np.isnan(info["observable_dmu"][~info["observable_invariant"]]).all()


def test_wrench_response_of_a_sliding_box_equals_central_differences():
    G, b, mu, J, _ = R.corner_box_scene(moving=True)
    lam, hist, labels = R.solve_ncp_pgs(G, b, mu, 20000, 1e-15)
    assert list(labels) == ["slip"] * 4
    _, _, info = R.sensitivity(lam, G, b, mu, labels, observable=J.T, db_dparam=J)
    assert info["singular"] and info["observable_invariant"].all() and info["b_compatible"].all()
    assert info["observable_b_valid"].all()
    h = 1e-6
    for j in (0, 2):
        lp = R.solve_ncp_pgs(G, b + h * J[:, j], mu, 20000, 1e-15, lam)[0]
        lm = R.solve_ncp_pgs(G, b - h * J[:, j], mu, 20000, 1e-15, lam)[0]
        fd = J.T @ ((lp - lm) / (2 * h))
        assert np.allclose(fd, info["observable_db"][:, j], atol=1e-6, rtol=1e-6)


def test_near_singular_and_zero_matrices_are_guarded():
    A = np.diag([1.0, 1e-14, 2.0])
    fac = R.rank_factor(A)
    X, chk = R.guarded_response(A, np.eye(3), fac, selection="minimum_norm", observable=np.eye(3))
    assert fac[3]["singular"]
    assert not chk["compatible"][1] and np.isnan(X[:, 1]).all()
    assert not chk["observable_invariant"][1]
    assert R.rank_factor(np.zeros((3, 3)))[3]["nullity"] == 3


def test_boundary_disables_even_minimum_norm_derivatives():
    lam = np.array([1.0, -0.3, 0.0])
    G = np.eye(3)
    b = -lam
    dm, db, info = R.sensitivity(lam, G, b, [0.3], ["slip"], selection="minimum_norm", observable=np.eye(3))
    assert info["boundary"]
    assert np.isnan(dm).all() and np.isnan(db).all()
    assert not info["observable_mu_valid"].any()


def test_contact_diagnostics_of_a_box_resting_in_stick():
    # four sticking corners of one body: A = G = J M^-1 J^T, whose nullspace is that of J^T, so the
    # wrench J^T lam is invariant though the per-corner split is not unique
    G, _, mu, J, _ = R.corner_box_scene(moving=False)
    lam = np.zeros(12)
    lam[::3] = 9.81 * 0.01 / 4
    b = -G @ lam
    info = R.contact_diagnostics(lam, G, b, mu, ["stick"] * 4, observable=J.T)
    assert info["singular"] and info["rank"] == 6 and not info["boundary"]
    assert info["observable_invariant"].all()
    assert np.linalg.norm(J.T @ info["nullspace"]) < 1e-10
