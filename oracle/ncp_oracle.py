"""CPU restatement of the reference's NCP-level sensitivity: TEST INFRASTRUCTURE ONLY.

Imported by tests/ and the benchmark's CPU-baseline leg; never by the product. float64 numpy,
following src/motion_engine/ncp/ncp_ref.py of antonbj3/3fold-motion-engine:

  proj_cone           ncp_ref.py:158-183   projection onto the Coulomb cone K_mu
  natural_residual    ncp_ref.py:224-234   ||lam - proj_K(lam - rho (u + Gamma(u)))||_inf
  classify            ncp_ref.py:251-266   open / stick / slip labels at a solved lam
  solve_ncp_pgs       ncp_ref.py:340-450   projected Gauss-Seidel with exact one-contact solves
                                             (take-off / stick / slip root on the angle), the FD solver
  active_system       ncp_ref.py:786-842   fixed-branch residual Jacobian A, B_mu, B_b
  rank_factor         contact_rank.py:13-30 SVD rank at rcond, nullspaces
  guarded_response    contact_rank.py:33-70 A X = -B, NaN for ambiguous columns unless a
                                             minimum-norm selection is asked for; observable rows
  sensitivity         ncp_ref.py:907-959   (dlam/dmu, dlam/db, info)
  contact_diagnostics ncp_ref.py:845-866   rank, nullspaces, observable invariance
  sensitivity_fd      ncp_ref.py:962-993   central differences of the PGS solution

Pins (tests/test_ncp_oracle.py): the reference's all-slip analytic scenes
(tests/test_ncp_batch_adjoint.py:34-53) reproduce the labels and the condition numbers of the
active-set Jacobian recorded in data/ncp/adjoint_fix_table.csv; central finite differences of the
PGS solution (the reference's own criterion, tests/test_ncp_reference.py:316-350); the diagonal
slip case in closed form; the hyperstatic 4-point cube of data/ncp/scenes/00.json flagged singular
with NaN derivatives; the redundant-corner contracts of tests/test_ncp_redundant_contacts.py
(REDUNDANT_CONTACTS with an orthonormal nullspace and the alternating-normal left-null witness in all
64 closed corner instances, the minimum-norm response orthogonal to the nullspace, observable rows
invariant for the wrench and not for a spinning slip, the boundary and near-singular guards).

corner_box_scene / algebraic_branch are the fixtures of those contracts."""
from __future__ import annotations

import numpy as np

LABELS = ("open", "stick", "slip")


def proj_cone(x, mu):
    """Projection of one contact impulse x = (n, t1, t2) onto {|t| <= mu n}."""
    x = np.asarray(x, np.float64)
    xn, xt = x[0], x[1:3]
    nt = float(np.linalg.norm(xt))
    if nt <= mu * xn:
        return x.copy()
    if mu * nt <= -xn:
        return np.zeros(3)
    a = (mu * nt + xn) / (1.0 + mu * mu)
    return np.concatenate([[a], (mu * a / nt) * xt])


def natural_residual(lam, G, b, mu):
    """De Saxce natural-map residual with rho_c = 1 / ||G_cc||_2."""
    lam = np.asarray(lam, np.float64)
    n_c = len(mu)
    u = (G @ lam + b).reshape(n_c, 3)
    L = lam.reshape(n_c, 3)
    out = 0.0
    for c in range(n_c):
        s = np.linalg.norm(G[3 * c:3 * c + 3, 3 * c:3 * c + 3], ord=2)
        rho = 1.0 / s if s > 0 else 1.0
        w = u[c] + np.array([mu[c] * np.linalg.norm(u[c, 1:3]), 0.0, 0.0])
        out = max(out, float(np.max(np.abs(L[c] - proj_cone(L[c] - rho * w, mu[c])))))
    return out


def classify(lam, G, b, mu, tol_lam=1e-10, tol_rel=1e-6):
    lam = np.asarray(lam, np.float64)
    n_c = len(mu)
    L = lam.reshape(n_c, 3)
    scale = max(float(np.max(np.abs(L))) if n_c else 0.0, 1e-30)
    out = []
    for c in range(n_c):
        ln, lt = L[c, 0], float(np.linalg.norm(L[c, 1:3]))
        if ln <= tol_lam + tol_rel * scale:
            out.append("open")
        elif lt >= mu[c] * ln * (1.0 - tol_rel) - tol_lam:
            out.append("slip")
        else:
            out.append("stick")
    return np.array(out)


_THETA_GRID = 64
_BISECT_IT = 60


def _rho_c(Gcc):
    s = np.linalg.norm(Gcc, ord=2)
    return 1.0 / s if s > 0 else 1.0


def _local_residual(l3, A, w, mu, rho):
    u = A @ l3 + w
    g = np.array([mu * np.hypot(u[1], u[2]), 0.0, 0.0])
    return float(np.max(np.abs(l3 - proj_cone(l3 - rho * (u + g), mu))))


def _slip_at(P, mu, th):
    """Slip branch at angle th: lam = lam_n (1, -mu d), d = (cos th, sin th), lam_n from u_n = 0;
    returns (cross(d, u_t), d . u_t, lam_n, lam_t0, lam_t1) or None (ncp_ref.py:313-337)."""
    a_nn, a_nt0, a_nt1, a_tn0, a_tn1, A00, A01, A10, A11, w0, w1, w2 = P
    cd, sd = np.cos(th), np.sin(th)
    den = a_nn - mu * (a_nt0 * cd + a_nt1 * sd)
    if abs(den) < 1e-300:
        return None
    ln = -w0 / den
    lt0, lt1 = -mu * ln * cd, -mu * ln * sd
    ut0 = ln * a_tn0 + A00 * lt0 + A01 * lt1 + w1
    ut1 = ln * a_tn1 + A10 * lt0 + A11 * lt1 + w2
    return (cd * ut1 - sd * ut0, cd * ut0 + sd * ut1, ln, lt0, lt1)


def _local_solve(A, w, mu, lam_init, rho, tol=1e-14):
    """Exact one-contact NCP solve with the others frozen (ncp_ref.py:340-421, de Saxce form):
    take-off, stick, then slip by a sign change of cross(d, u_t) over a 64-point angle grid refined
    by 60 bisections; the candidate with the smallest local residual below tol wins, else a local
    projected fixed point."""
    if w[0] >= 0.0:
        return np.zeros(3)
    cands = []
    try:
        st = -np.linalg.solve(A, w)
        if st[0] >= 0.0 and np.hypot(st[1], st[2]) <= mu * st[0]:
            return st
        if np.all(np.isfinite(st)):
            cands.append(st)
    except np.linalg.LinAlgError:
        pass
    if mu > 0.0:
        P = (A[0, 0], A[0, 1], A[0, 2], A[1, 0], A[2, 0], A[1, 1], A[1, 2], A[2, 1], A[2, 2], w[0], w[1], w[2])
        step = 2.0 * np.pi / _THETA_GRID
        prev = _slip_at(P, mu, 0.0)
        for i in range(1, _THETA_GRID + 1):
            th = i * step
            cur = _slip_at(P, mu, th)
            if prev is not None and cur is not None and prev[0] * cur[0] <= 0.0:
                lo, hi, flo = th - step, th, prev[0]
                for _ in range(_BISECT_IT):
                    mid = 0.5 * (lo + hi)
                    r = _slip_at(P, mu, mid)
                    if r is None:
                        break
                    if flo * r[0] <= 0.0:
                        hi = mid
                    else:
                        lo, flo = mid, r[0]
                r = _slip_at(P, mu, 0.5 * (lo + hi))
                if r is not None and r[2] > 0.0 and r[1] > 0.0:
                    cands.append(np.array([r[2], r[3], r[4]]))
            prev = cur
    best, best_r = None, np.inf
    for c in cands:
        if np.all(np.isfinite(c)):
            r = _local_residual(c, A, w, mu, rho)
            if r < best_r:
                best, best_r = c, r
    if best is not None and best_r <= tol:
        return best
    l = np.array(lam_init, np.float64)
    for _ in range(200):
        u = A @ l + w
        g = np.array([mu * np.hypot(u[1], u[2]), 0.0, 0.0])
        l2 = proj_cone(l - rho * (u + g), mu)
        done = np.max(np.abs(l2 - l)) < 1e-16
        l = l2
        if done:
            break
    for c in (l, np.array(lam_init, np.float64)):
        if np.all(np.isfinite(c)):
            r = _local_residual(c, A, w, mu, rho)
            if r < best_r:
                best, best_r = c, r
    return best if best is not None else np.zeros(3)


def solve_ncp_pgs(G, b, mu, iters=2000, tol=1e-12, warm=None):
    """Projected Gauss-Seidel with exact local solves and the natural-map stopping test
    (ncp_ref.py:424-450). Returns (lam, residual history, labels)."""
    G = np.asarray(G, np.float64)
    b = np.asarray(b, np.float64)
    mu = np.asarray(mu, np.float64)
    n_c = len(mu)
    lam = np.zeros(3 * n_c) if warm is None else np.array(warm, np.float64).copy()
    blocks = [G[3 * c:3 * c + 3, 3 * c:3 * c + 3] for c in range(n_c)]
    rho = [_rho_c(B) for B in blocks]
    hist = []
    for _ in range(int(iters)):
        for c in range(n_c):
            sl = slice(3 * c, 3 * c + 3)
            w = b[sl] + G[sl, :] @ lam - blocks[c] @ lam[sl]
            lam[sl] = _local_solve(blocks[c], w, mu[c], lam[sl], rho[c])
        hist.append(natural_residual(lam, G, b, mu))
        if hist[-1] < tol:
            break
    return lam, np.array(hist), classify(lam, G, b, mu)


def active_system(lam, G, b, mu, labels):
    """open: lam_c = 0; stick: u_c = 0; slip: u_n = 0 and lam_t + mu lam_n u_t/|u_t| = 0, with
    u = G lam + b. Returns (A, B_mu, B_b, boundary) for A dlam = -(B_mu dmu + B_b db)."""
    lam = np.asarray(lam, np.float64)
    n_c = len(mu)
    n = 3 * n_c
    u = G @ lam + b
    A = np.zeros((n, n))
    Bmu = np.zeros((n, n_c))
    Bb = np.zeros((n, n))
    boundary = []
    for c in range(n_c):
        i = 3 * c
        lab = labels[c]
        if lab == "open":
            A[i:i + 3, i:i + 3] = np.eye(3)
            if abs(u[i]) < 1e-9:
                boundary.append((c, "open_with_zero_normal_velocity"))
        elif lab == "stick":
            A[i:i + 3, :] = G[i:i + 3, :]
            Bb[i:i + 3, i:i + 3] = np.eye(3)
        else:
            A[i, :] = G[i, :]
            Bb[i, i] = 1.0
            s = u[i + 1:i + 3]
            ns = float(np.linalg.norm(s))
            if ns < 1e-14:
                boundary.append((c, "slip_with_zero_slip_speed"))
                ns = 1e-14
            sh = s / ns
            P = (np.eye(2) - np.outer(sh, sh)) / ns
            A[i + 1:i + 3, :] = mu[c] * lam[i] * (P @ G[i + 1:i + 3, :])
            A[i + 1:i + 3, i] += mu[c] * sh
            A[i + 1, i + 1] += 1.0
            A[i + 2, i + 2] += 1.0
            Bb[i + 1:i + 3, i + 1:i + 3] = mu[c] * lam[i] * P
            Bmu[i + 1:i + 3, c] = lam[i] * sh
        ln, lt = lam[i], float(np.linalg.norm(lam[i + 1:i + 3]))
        scale = max(abs(ln), 1e-12)
        if lab == "stick" and abs(lt - mu[c] * ln) < 1e-7 * mu[c] * scale:
            boundary.append((c, "stick_on_cone_boundary"))
        if lab != "open" and ln < 1e-9:
            boundary.append((c, "near_zero_normal_impulse"))
    return A, Bmu, Bb, boundary


def rank_factor(A, rcond=1e-12):
    U, s, Vt = np.linalg.svd(A, full_matrices=True)
    cutoff = rcond * s[0] if s.size else 0.0
    rank = int(np.sum(s > cutoff))
    cond = float(s[0] / s[-1]) if s.size and s[-1] > 0 else (float("inf") if s.size else 0.0)
    info = {"status": "REDUNDANT_CONTACTS" if rank < A.shape[1] else "UNIQUE_LINEAR_SYSTEM",
            "rank": rank, "nullity": A.shape[1] - rank, "nullspace": Vt[rank:].T.copy(),
            "left_nullspace": U[:, rank:].copy(), "singular_values": s.copy(), "rank_cutoff": cutoff,
            "cond": cond, "singular": rank < A.shape[1]}
    return U, s, Vt, info


def guarded_response(A, B, factors, selection=None, observable=None, tol=1e-10):
    """contact_rank.py:33-70: X = -A^+ B with NaN for ambiguous (singular A, no selection),
    incompatible and non-finite columns; observable rows Q reported where Q N = 0."""
    U, s, Vt, info = factors
    B = np.asarray(B, np.float64)
    rank = info["rank"]
    bscale = np.max(np.abs(B), axis=0, initial=0.0)
    sB = B / np.where(bscale > 0, bscale, 1.0)
    rel = np.linalg.norm(info["left_nullspace"].T @ sB, axis=0)
    compatible = rel <= tol * np.linalg.norm(sB, axis=0)
    if info["singular"]:
        projected = -(Vt[:rank].T @ ((U[:, :rank].T @ B) / s[:rank, None]))
        X = projected.copy() if selection == "minimum_norm" else np.full_like(projected, np.nan)
        X[:, ~compatible] = np.nan
    else:
        projected = -np.linalg.solve(A, B)
        X = projected.copy()
    finite = np.isfinite(projected).all(axis=0)
    X[:, ~finite] = np.nan
    out = {"response_finite": finite, "compatible": compatible, "compatibility_residual": rel * bscale}
    if observable is not None:
        Q = np.atleast_2d(np.asarray(observable, np.float64))
        qscale = np.max(np.abs(Q), axis=1, initial=0.0)
        sQ = Q / np.where(qscale > 0, qscale, 1.0)[:, None]
        leak = np.linalg.norm(sQ @ info["nullspace"], axis=1)
        invariant = leak <= tol * np.linalg.norm(sQ, axis=1)
        Y = Q @ projected
        valid = invariant[:, None] & compatible[None, :] & np.isfinite(Y)
        Y[~valid] = np.nan
        out.update(observable=Y, observable_valid=valid, observable_invariant=invariant,
                   observable_nullspace_residual=leak * qscale)
    return X, out


def sensitivity(lam, G, b, mu, labels, rcond=1e-12, *, selection=None, observable=None, db_dparam=None):
    """(dlam/dmu, dlam/db, info); db_dparam maps parameters to b (dlam/db @ db_dparam)."""
    A, Bmu, Bb, boundary = active_system(lam, G, b, mu, labels)
    if db_dparam is not None:
        Bb = Bb @ np.asarray(db_dparam, np.float64)
    factors = rank_factor(A, rcond)
    dm, im = guarded_response(A, Bmu, factors, selection, observable)
    db, ib = guarded_response(A, Bb, factors, selection, observable)
    info = factors[3]
    info.update(boundary=boundary, labels=list(labels), mu_compatible=im["compatible"],
                b_compatible=ib["compatible"], mu_response_finite=im["response_finite"],
                b_response_finite=ib["response_finite"], mu_compatibility_residual=im["compatibility_residual"],
                b_compatibility_residual=ib["compatibility_residual"])
    if observable is not None:
        info.update(observable_dmu=im["observable"], observable_db=ib["observable"],
                    observable_mu_valid=im["observable_valid"], observable_b_valid=ib["observable_valid"],
                    observable_invariant=im["observable_invariant"],
                    observable_nullspace_residual=im["observable_nullspace_residual"])
    if boundary:
        dm[:] = np.nan
        db[:] = np.nan
        if observable is not None:
            info["observable_dmu"][:] = np.nan
            info["observable_db"][:] = np.nan
            info["observable_mu_valid"][:] = False
            info["observable_b_valid"][:] = False
    return dm, db, info


def contact_diagnostics(lam, G, b, mu, labels, rcond=1e-12, observable=None):
    """ncp_ref.py:845-866: rank, nullspaces and observable invariance at a fixed branch."""
    A, _, _, boundary = active_system(lam, G, b, mu, labels)
    factors = rank_factor(A, rcond)
    info = factors[3]
    info.update(boundary=boundary, labels=list(labels))
    if observable is not None:
        _, chk = guarded_response(A, np.zeros((A.shape[0], 0)), factors, observable=observable)
        info.update(observable_invariant=chk["observable_invariant"],
                    observable_nullspace_residual=chk["observable_nullspace_residual"])
    return info


def sensitivity_fd(G, b, mu, lam_star, h_mu=1e-6, h_b=1e-6, iters=20000, tol=1e-14):
    G = np.asarray(G, np.float64)
    b = np.asarray(b, np.float64)
    mu = np.asarray(mu, np.float64)
    n_c = len(mu)
    n = 3 * n_c
    dmu = np.zeros((n, n_c))
    db = np.zeros((n, n))
    for c in range(n_c):
        mp, mm = mu.copy(), mu.copy()
        mp[c] += h_mu
        mm[c] -= h_mu
        dmu[:, c] = (solve_ncp_pgs(G, b, mp, iters, tol, lam_star)[0]
                     - solve_ncp_pgs(G, b, mm, iters, tol, lam_star)[0]) / (2 * h_mu)
    for j in range(n):
        bp, bm = b.copy(), b.copy()
        bp[j] += h_b
        bm[j] -= h_b
        db[:, j] = (solve_ncp_pgs(G, bp, mu, iters, tol, lam_star)[0]
                    - solve_ncp_pgs(G, bm, mu, iters, tol, lam_star)[0]) / (2 * h_b)
    return dmu, db


def corner_box_scene(half=(0.5, 0.5, 0.5), mass=1.0, angle=0.0, moving=True, mu=0.3, dt=0.01):
    """One solid box standing on the ground on its four bottom corners, tilted by `angle` about y, the
    fixture of the reference's redundant-contact tests (tests/test_ncp_redundant_contacts.py:15-28):
    contact frame n = the box's z axis, t1 = n x a (a = x, or y when |n_x| >= 0.9), t2 = n x t1 (the
    reference's _tangents, ncp/ncp_ref.py:1057-1062), J_c = [D, -D [r]x] per corner (r from the
    centre), Minv = diag(1/m, I^-1 in world), v_free = the box's +x velocity (when moving) + dt g.
    Returns (G, b, mu, J, Minv)."""
    h = np.asarray(half, np.float64)
    c, s_ = np.cos(angle), np.sin(angle)
    rot = np.array([[c, 0.0, s_], [0.0, 1.0, 0.0], [-s_, 0.0, c]])
    centre = rot @ np.array([0.0, 0.0, h[2]])
    nrm = rot[:, 2]
    a = np.array([1.0, 0.0, 0.0]) if abs(nrm[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t1 = np.cross(nrm, a)
    t1 /= np.linalg.norm(t1)
    D = np.stack([nrm, t1, np.cross(nrm, t1)])
    J = np.zeros((12, 6))
    for k, (sx, sy) in enumerate(((-1, -1), (-1, 1), (1, -1), (1, 1))):
        r = rot @ np.array([sx * h[0], sy * h[1], 0.0]) - centre
        rx = np.array([[0.0, -r[2], r[1]], [r[2], 0.0, -r[0]], [-r[1], r[0], 0.0]])
        J[3 * k:3 * k + 3, :3] = D
        J[3 * k:3 * k + 3, 3:] = -D @ rx
    Ib = mass / 3.0 * np.array([h[1] ** 2 + h[2] ** 2, h[0] ** 2 + h[2] ** 2, h[0] ** 2 + h[1] ** 2])
    Minv = np.zeros((6, 6))
    Minv[:3, :3] = np.eye(3) / mass
    Minv[3:, 3:] = rot @ np.diag(1.0 / Ib) @ rot.T
    vf = np.zeros(6)
    vf[:3] = (rot @ np.array([1.0, 0.0, 0.0]) if moving else 0.0) + dt * np.array([0.0, 0.0, -9.81])
    G = J @ Minv @ J.T
    return G, J @ vf, np.full(4, mu), J, Minv


def algebraic_branch(G, mu, labels, spin=False):
    """A residual fixture with the given labels (the reference's algebraic_branch,
    tests/test_ncp_redundant_contacts.py:31-42): unit normal impulses, slip impulses on the cone
    against unit slip velocities (diagonal per corner with spin), open contacts separating; b = u -
    G lam. Returns (lam, b)."""
    n = 3 * len(mu)
    lam, u = np.zeros(n), np.zeros(n)
    for c, label in enumerate(labels):
        i = 3 * c
        lam[i] = 0.0 if label == "open" else 1.0
        if label == "open":
            u[i] = 0.5
        if label == "slip":
            sh = (np.array([1.0, 0.0]) if not spin
                  else np.array([(c % 2) * 2 - 1.0, (c // 2) * 2 - 1.0]) / np.sqrt(2.0))
            lam[i + 1:i + 3] = -mu[c] * sh
            u[i + 1:i + 3] = sh
    return lam, u - G @ lam


def all_slip_scene(n_c):
    """SPD, block-tridiagonal Delassus with a rhs that puts every contact in slip (contact c couples
    body c to body c+1, the last to the ground; tests/test_ncp_batch_adjoint.py:34-53)."""
    n = 3 * n_c
    J = np.zeros((n, n))
    for c in range(n_c):
        J[3 * c:3 * c + 3, 3 * c:3 * c + 3] = np.eye(3)
        if c + 1 < n_c:
            J[3 * c:3 * c + 3, 3 * (c + 1):3 * (c + 1) + 3] = -np.eye(3)
    Minv = np.diag(np.repeat(1.0 / (1.0 + 0.5 * np.arange(n_c)), 3))
    G = J @ Minv @ J.T
    b = np.zeros(n)
    for c in range(n_c):
        b[3 * c:3 * c + 3] = [-1.0 - 0.1 * c, 0.9 + 0.2 * c, -0.7 + 0.1 * c]
    return G, b, np.full(n_c, 0.3)


def random_scene(n_c, n_b, seed, mu=0.4):
    """Random bodies (6 dof each, SPD inverse inertia) and contacts (normal, tangent frame, levers)
    between random body pairs or the ground, and a free velocity that presses every contact in
    with a tangential drive: a generic, non-degenerate contact problem."""
    rng = np.random.default_rng(seed)
    nv = 6 * n_b
    Minv = np.zeros((nv, nv))
    for k in range(n_b):
        Minv[6 * k:6 * k + 3, 6 * k:6 * k + 3] = np.eye(3) / rng.uniform(0.5, 2.0)
        a = rng.standard_normal((3, 3))
        Minv[6 * k + 3:6 * k + 6, 6 * k + 3:6 * k + 6] = np.linalg.inv(a @ a.T + np.eye(3))
    J = np.zeros((3 * n_c, nv))
    for c in range(n_c):
        nrm = rng.standard_normal(3)
        nrm /= np.linalg.norm(nrm)
        t1 = np.cross(nrm, rng.standard_normal(3))
        t1 /= np.linalg.norm(t1)
        D = np.stack([nrm, t1, np.cross(nrm, t1)])
        a = int(rng.integers(0, n_b))
        b = int(rng.integers(-1, n_b))
        ra = rng.standard_normal(3) * 0.2
        J[3 * c:3 * c + 3, 6 * a:6 * a + 3] += D
        J[3 * c:3 * c + 3, 6 * a + 3:6 * a + 6] += np.cross(ra[None, :], D)
        if b >= 0 and b != a:
            rb = rng.standard_normal(3) * 0.2
            J[3 * c:3 * c + 3, 6 * b:6 * b + 3] -= D
            J[3 * c:3 * c + 3, 6 * b + 3:6 * b + 6] -= np.cross(rb[None, :], D)
    G = J @ Minv @ J.T
    v = rng.standard_normal(nv) * 0.5
    bb = J @ v
    bb[0::3] = -np.abs(bb[0::3]) - 0.2
    return G, bb, np.full(n_c, mu)
