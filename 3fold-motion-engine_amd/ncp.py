"""NCP-level sensitivity of contact impulses on the GPU (float64): the semantics of the reference's
``ncp_ref.classify`` / ``_active_system`` / ``sensitivity`` (ncp/ncp_ref.py:251, 786, 907) and its
batched GPU form ``BatchProxADMMGPU.build_active`` + ``grad_affine`` (ncp/ncp_batch_gpu.py:1160-1187).

At a solved impulse lam of the contact problem u = G lam + b, lam in the Coulomb cones, every
contact is labelled open / stick / slip and the fixed-branch residual
    open: lam_c = 0;  stick: u_c = 0;  slip: u_n = 0, lam_t + mu lam_n u_t / |u_t| = 0
is differentiated by the implicit function theorem: A dlam = -(B_mu dmu + B_b db).

* ``batch_sensitivity`` solves it for many environments in one launch, per environment by LU with
  partial pivoting in LDS (``me_ncp_sensitivity``); singular systems (redundant contacts) and
  branch-boundary cases give NaN derivatives.
* ``sensitivity`` and ``contact_diagnostics`` (ncp_ref.py:845-866, 907-959) take the reference's SVD
  rank guard (contact_rank.py:13-70) on the GPU (``me_ncp_svd``: one-sided Jacobi SVD of A and of
  A^T): rank at rcond, right and left nullspaces, compatibility of each parameter column, the
  minimum-norm linear response on request (``selection='minimum_norm'``) and observable rows Q whose
  responses Q dlam are reported only where Q N = 0; with the same info dictionary as the reference.

``contact_problem(engine)`` builds (G, b, mu, lam) of the last substep of a ColorCache engine, so the
sensitivity can be taken at the contact step's own solution.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

LABELS = ("open", "stick", "slip")
_BOUNDARY = ((16, "open_with_zero_normal_velocity"), (32, "slip_with_zero_slip_speed"),
             (64, "stick_on_cone_boundary"), (128, "near_zero_normal_impulse"))
SINGULAR, BOUNDARY, NONFINITE = 1, 2, 4


def _torch():
    import torch
    return torch


def _dev(a, device, shape):
    torch = _torch()
    t = torch.as_tensor(np.asarray(a, np.float64) if not isinstance(a, torch.Tensor) else a)
    return t.to(device=device, dtype=torch.float64).reshape(shape).contiguous()


def batch_sensitivity(lam, G, b, mu, directions=None, labels=None, tol_lam=1e-10, tol_rel=1e-6, rcond=1e-12,
                      device="cuda:0"):
    """E problems of n contacts: G (E,3n,3n), b (E,3n), mu (E,n), lam (E,3n); directions (E,3n,m) or
    (3n,m) (None: the identity, m = 3n); labels (E,n) of 0/1/2 or None (classify).
    Returns a dict of torch tensors: dlam_dmu (E,3n,n), dlam_db (E,3n,m), labels (E,n) (0 open,
    1 stick, 2 slip), boundary (E,n) reason bits, status (E,) (1 singular, 2 boundary, 4 non-finite)."""
    torch = _torch()
    Gt = torch.as_tensor(G) if not isinstance(G, torch.Tensor) else G
    E = int(Gt.shape[0]) if Gt.dim() == 3 else 1
    n3 = int(Gt.shape[-1])
    if n3 % 3:
        raise ValueError("G must be (E, 3n, 3n)")
    n = n3 // 3
    Gd = _dev(G, device, (E, n3, n3))
    bd = _dev(b, device, (E, n3))
    md = _dev(mu, device, (E, n))
    ld = _dev(lam, device, (E, n3))
    if directions is None:
        m, Dd = n3, None
    elif np.shape(directions)[-1] == 0:  # labels only
        m, Dd = 0, torch.zeros(1, dtype=torch.float64, device=device)
    else:
        D = _dev(directions, device, (-1, n3, np.shape(directions)[-1]))
        m = int(D.shape[-1])
        Dd = (D.expand(E, n3, m) if D.shape[0] == 1 else D).contiguous()
        if Dd.shape[0] != E:
            raise ValueError("directions must be (3n, m) or (E, 3n, m)")
    labs_in = None
    if labels is not None:
        la = np.asarray(labels)
        if la.dtype.kind in "US":
            la = np.vectorize(LABELS.index)(la)
        labs_in = torch.as_tensor(la.astype(np.int32)).to(device).reshape(E, n).contiguous()
    dmu = torch.empty((E, n3, n), dtype=torch.float64, device=device)
    dlb = torch.empty((E, n3, max(m, 1)), dtype=torch.float64, device=device)
    lab = torch.empty((E, n), dtype=torch.int32, device=device)
    st = torch.empty(E, dtype=torch.int32, device=device)
    L = _lib.lib()
    nb = C.c_size_t()
    _lib.check(L.me_ncp_workspace_size(E, n, m, C.byref(nb)), "me_ncp_workspace_size")
    ws = torch.empty(max(int(nb.value), 8), dtype=torch.uint8, device=device)
    stream = torch.cuda.current_stream(torch.device(device)).cuda_stream
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    _lib.check(L.me_ncp_sensitivity(E, n, m, p(Gd), p(bd), p(md), p(ld), p(Dd), p(labs_in), float(tol_lam),
                                    float(tol_rel), float(rcond), p(dmu), p(dlb), p(lab), p(st), p(ws),
                                    int(nb.value), C.c_void_p(stream)), "me_ncp_sensitivity")
    return dict(dlam_dmu=dmu, dlam_db=dlb[:, :, :m], labels=lab & 3, boundary=lab & ~3, status=st)


def classify(lam, G, b, mu, tol_lam=1e-10, tol_rel=1e-6, device="cuda:0"):
    """Open / stick / slip labels at a solved lam (ncp_ref.classify), computed on the GPU."""
    n3 = len(np.asarray(b).reshape(-1))
    out = batch_sensitivity(lam, G, b, mu, directions=np.zeros((n3, 0)), tol_lam=tol_lam, tol_rel=tol_rel,
                            device=device)
    return np.array([LABELS[int(k)] for k in out["labels"][0].cpu().numpy()])


def svd_response(lam, G, b, mu, directions=None, labels=None, observable=None, minimum_norm=False, rcond=1e-12,
                 tol=1e-10, tol_lam=1e-10, tol_rel=1e-6, device="cuda:0"):
    """The SVD-guarded response of E problems of n contacts (``me_ncp_svd``). G (E,3n,3n), b (E,3n),
    mu (E,n), lam (E,3n); directions (3n,m) or (E,3n,m), None for the identity, or a (3n,0) array for
    no b columns; labels (E,n) of 0/1/2 or label strings, None to classify; observable (K,3n) shared
    or (E,K,3n). Returns a dict of numpy arrays: X (E,3n,n+m), S, V, U, rank, compat, finite, resid,
    Y, valid, invariant, leak, labels (label bits), boundary (reason bits), status."""
    torch = _torch()
    Gt = torch.as_tensor(G) if not isinstance(G, torch.Tensor) else G
    E = int(Gt.shape[0]) if Gt.dim() == 3 else 1
    n3 = int(Gt.shape[-1])
    if n3 % 3:
        raise ValueError("G must be (E, 3n, 3n)")
    n = n3 // 3
    dev = torch.device(device)
    Gd, bd, md, ld = _dev(G, dev, (E, n3, n3)), _dev(b, dev, (E, n3)), _dev(mu, dev, (E, n)), _dev(lam, dev, (E, n3))
    if directions is None:
        m, Dd = n3, None
    elif np.shape(directions)[-1] == 0:
        m, Dd = 0, None
    else:
        D = _dev(directions, dev, (-1, n3, np.shape(directions)[-1]))
        m = int(D.shape[-1])
        Dd = (D.expand(E, n3, m) if D.shape[0] == 1 else D).contiguous()
        if Dd.shape[0] != E:
            raise ValueError("directions must be (3n, m) or (E, 3n, m)")
    labs_in = None
    if labels is not None:
        la = np.asarray(labels)
        if la.dtype.kind in "US":
            la = np.vectorize(LABELS.index)(la)
        labs_in = torch.as_tensor(la.astype(np.int32)).to(dev).reshape(E, n).contiguous()
    K, Qd, q_shared = 0, None, 0
    if observable is not None:
        Qa = np.asarray(observable, np.float64)
        if Qa.ndim == 1:
            Qa = Qa[None, :]
        if Qa.shape[-1] != n3:
            raise ValueError("observable must have 3 n_c columns")
        q_shared = int(Qa.ndim == 2)
        K = int(Qa.shape[-2])
        Qd = _dev(Qa, dev, (-1, K, n3))
    R = n + m
    f64, i32 = dict(dtype=torch.float64, device=dev), dict(dtype=torch.int32, device=dev)
    out = dict(X=torch.empty((E, n3, R), **f64), S=torch.empty((E, n3), **f64), V=torch.empty((E, n3, n3), **f64),
               U=torch.empty((E, n3, n3), **f64), rank=torch.empty(E, **i32), compat=torch.empty((E, R), **i32),
               finite=torch.empty((E, R), **i32), resid=torch.empty((E, R), **f64),
               Y=torch.empty((E, max(K, 1), R), **f64), valid=torch.empty((E, max(K, 1), R), **i32),
               invariant=torch.empty((E, max(K, 1)), **i32), leak=torch.empty((E, max(K, 1)), **f64),
               labels=torch.empty((E, n), **i32), status=torch.empty(E, **i32))
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    o = _lib.NcpSvdOut(*[p(out[k]) for k in ("X", "S", "V", "U", "rank", "compat", "finite", "resid", "Y", "valid",
                                             "invariant", "leak", "labels", "status")])
    L = _lib.lib()
    nb = C.c_size_t()
    _lib.check(L.me_ncp_svd_workspace_size(E, n, m, C.byref(nb)), "me_ncp_svd_workspace_size")
    ws = torch.empty(max(int(nb.value), 8), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    _lib.check(L.me_ncp_svd(E, n, m, K, p(Gd), p(bd), p(md), p(ld), p(Dd), p(labs_in), p(Qd), q_shared,
                            float(tol_lam), float(tol_rel), float(rcond), float(tol), int(bool(minimum_norm)),
                            C.byref(o), p(ws), int(nb.value), C.c_void_p(stream)), "me_ncp_svd")
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["boundary"] = res["labels"] & ~3
    res["labels"] = res["labels"] & 3
    if K == 0:
        for k in ("Y", "valid", "invariant", "leak"):
            res[k] = res[k][:, :0]
    return res


def _check_inputs(lam, G, b, mu, labels):
    mu = np.asarray(mu, np.float64)
    n = len(mu)
    if np.shape(G) != (3 * n, 3 * n) or np.shape(b) != (3 * n,) or np.shape(lam) != (3 * n,) or len(labels) != n:
        raise ValueError("contact arrays and labels have incompatible shapes")
    if any(lab not in LABELS for lab in labels):
        raise ValueError("labels must be open, stick or slip")
    for a in (lam, G, b, mu):
        if not np.all(np.isfinite(np.asarray(a, np.float64))):
            raise ValueError("contact inputs must be finite")
    return mu, n


def _rank_info(r, n3, rcond, labels):
    """rank_factor's info (contact_rank.py:13-30) and the branch fields of ncp_ref.py:845-866."""
    rank = int(r["rank"][0])
    s = r["S"][0]
    V, U = r["V"][0], r["U"][0]
    cond = float(s[0] / s[-1]) if s.size and s[-1] > 0 else (float("inf") if s.size else 0.0)
    bnd = r["boundary"][0]
    boundary = [(c, why) for c in range(len(labels)) for bit, why in _BOUNDARY if bnd[c] & bit]
    singular = rank < n3
    info = {"status": "REDUNDANT_CONTACTS" if singular else "UNIQUE_LINEAR_SYSTEM", "rank": rank,
            "nullity": n3 - rank, "nullspace": V[:, rank:].copy(), "left_nullspace": U[:, rank:].copy(),
            "singular_values": s.copy(), "rank_cutoff": rcond * s[0] if s.size else 0.0, "rcond": rcond,
            "cond": cond, "singular": singular, "scope": "NUMERICAL_FIXED_BRANCH_AT_SUPPLIED_IMPULSE",
            "selection": None}
    info.update(boundary=boundary, labels=list(labels), residual_differentiable=not boundary,
                response_scope="NONSMOOTH_BRANCH" if boundary else
                ("LINEARIZED_RESPONSE_ONLY" if singular else "IMPLICIT_RESPONSE_UNDER_VALID_BRANCH_ASSUMPTIONS"))
    return info, boundary


def sensitivity(lam, G, b, mu, labels, rcond=1e-12, *, selection=None, observable=None, db_dparam=None,
                device="cuda:0"):
    """(dlam/dmu, dlam/db, info) of one problem, as ncp_ref.sensitivity (ncp_ref.py:907-959): the
    SVD rank guard on the GPU, NaN for ambiguous derivatives unless selection='minimum_norm', NaN
    for incompatible columns and on a branch boundary in either mode; observable=Q (K, 3n) reports
    the responses of Q lam for rows with Q N = 0 (info['observable_dmu'], ['observable_db'], ...);
    db_dparam maps parameters to b (the b columns are dlam/db @ db_dparam)."""
    if selection not in (None, "minimum_norm"):
        raise ValueError("selection must be None or 'minimum_norm'")
    if not 0.0 < rcond < 1.0:
        raise ValueError("rcond must be between zero and one")
    mu, n = _check_inputs(lam, G, b, mu, labels)
    n3 = 3 * n
    if db_dparam is not None:
        P = np.asarray(db_dparam, dtype=np.float64)
        if P.ndim != 2 or P.shape[0] != n3 or not np.isfinite(P).all():
            raise ValueError("db_dparam must be a finite matrix with 3*n_c rows")
    else:
        P = None
    if observable is not None:
        Q = np.atleast_2d(np.asarray(observable, dtype=np.float64))
        if Q.shape[1] != n3 or not np.isfinite(Q).all():
            raise ValueError("observable must be finite with A.shape[1] columns")
    else:
        Q = None
    r = svd_response(lam, G, b, mu, directions=P, labels=list(labels), observable=Q,
                     minimum_norm=selection == "minimum_norm", rcond=rcond, device=device)
    info, boundary = _rank_info(r, n3, rcond, labels)
    X = r["X"][0]
    dm, db = X[:, :n].copy(), X[:, n:].copy()
    comp, fin, res = r["compat"][0].astype(bool), r["finite"][0].astype(bool), r["resid"][0]
    info.update(selection="MINIMUM_NORM_LINEAR_RESPONSE" if selection else None,
                mu_compatible=comp[:n], b_compatible=comp[n:], mu_response_finite=fin[:n],
                b_response_finite=fin[n:], mu_compatibility_residual=res[:n], b_compatibility_residual=res[n:])
    if Q is not None:
        Y, valid = r["Y"][0], r["valid"][0].astype(bool)
        info.update(observable_dmu=Y[:, :n].copy(), observable_db=Y[:, n:].copy(),
                    observable_mu_valid=valid[:, :n].copy(), observable_b_valid=valid[:, n:].copy(),
                    observable_invariant=r["invariant"][0].astype(bool), observable_nullspace_residual=r["leak"][0])
    if boundary:
        dm[:] = np.nan
        db[:] = np.nan
    return dm, db, info


def contact_diagnostics(lam, G, b, mu, labels, rcond=1e-12, observable=None, device="cuda:0"):
    """Numerical local rank, nullspaces and optional invariant observable rows at a fixed branch
    (ncp_ref.contact_diagnostics, ncp_ref.py:845-866), on the GPU."""
    if not 0.0 < rcond < 1.0:
        raise ValueError("rcond must be between zero and one")
    mu, n = _check_inputs(lam, G, b, mu, labels)
    Q = None
    if observable is not None:
        Q = np.atleast_2d(np.asarray(observable, dtype=np.float64))
        if Q.shape[1] != 3 * n or not np.isfinite(Q).all():
            raise ValueError("observable must be finite with A.shape[1] columns")
    r = svd_response(lam, G, b, mu, directions=np.zeros((3 * n, 0)), labels=list(labels), observable=Q,
                     rcond=rcond, device=device)
    info, _ = _rank_info(r, 3 * n, rcond, labels)
    if Q is not None:
        info.update(observable_invariant=r["invariant"][0].astype(bool), observable_nullspace_residual=r["leak"][0])
    return info


def synthetic_problems(E, n_c=4, nv=18, seed=0, p_open=0.1, p_slip=0.4, mu_range=(0.3, 0.9)):
    """E contact problems with exactly known solutions (numpy, float64), the shape of the reference's
    batched stance scenes (ncp/ncp_batch_gpu.py: n_c = 4 point contacts on nv = 18 generalised
    coordinates): G = J M^-1 J^T with a random J (3n_c, nv) and SPD M^-1; every contact drawn open,
    stick or slip; lam and u = G lam + b built to satisfy its branch exactly (stick: u = 0 with lam
    inside the cone; slip: u_n = 0, lam_t = -mu lam_n u_t/|u_t|; open: lam = 0, u_n > 0), then
    b = u - G lam. Returns (G, b, mu, lam, labels)."""
    rng = np.random.default_rng(seed)
    n3 = 3 * n_c
    J = rng.standard_normal((E, n3, nv))
    A = rng.standard_normal((E, nv, nv))
    Minv = A @ A.transpose(0, 2, 1) / nv + 0.1 * np.eye(nv)
    G = J @ Minv @ J.transpose(0, 2, 1)
    mu = rng.uniform(*mu_range, (E, n_c))
    r = rng.random((E, n_c))
    labels = np.where(r < p_open, 0, np.where(r < p_open + p_slip, 2, 1))
    ln = rng.uniform(0.5, 2.0, (E, n_c))
    th = rng.uniform(0.0, 2.0 * np.pi, (E, n_c))
    d = np.stack([np.cos(th), np.sin(th)], axis=-1)
    lam = np.zeros((E, n_c, 3))
    u = np.zeros((E, n_c, 3))
    stick, slip, opn = labels == 1, labels == 2, labels == 0
    lam[..., 0] = np.where(opn, 0.0, ln)
    rad = np.where(stick, rng.uniform(0.0, 0.8, (E, n_c)), 1.0) * mu * ln
    lam[..., 1:] = np.where(opn[..., None], 0.0, rad[..., None] * d)
    u[..., 1:] = np.where(slip[..., None], -rng.uniform(0.1, 1.0, (E, n_c))[..., None] * d,
                          np.where(opn[..., None], rng.standard_normal((E, n_c, 2)), 0.0))
    u[..., 0] = np.where(opn, rng.uniform(0.1, 1.0, (E, n_c)), 0.0)
    lam = lam.reshape(E, n3)
    b = u.reshape(E, n3) - np.einsum("eij,ej->ei", G, lam)
    return G, b, mu, lam, labels


def _fixed_t1(nrm):
    a = np.array([0.0, 1.0, 0.0]) if abs(nrm[0]) >= 0.9 else np.array([1.0, 0.0, 0.0])
    t = a - np.dot(a, nrm) * nrm
    return t / np.linalg.norm(t)


def contact_problem(engine):
    """(G, b, mu, lam) of the last substep of a ColorCache engine (one step of one substep, or the
    last substep of a step), in the solver's contact frame (n, t1, t2 with the normal-derived t1),
    float64: J maps the generalised velocities (v, w) of the dynamic bodies to contact-frame
    relative velocities at the contact points, M^-1 holds their inverse masses and world inverse
    inertias of the substep, b = J V_free with the velocities after gravity and before the solve,
    lam the solve's accumulated impulses. Returns dict(G, b, mu, lam, J, Minv, bodies)."""
    if engine._solver != _lib.SOLVER_COLOR_CACHE or engine.P:
        raise ValueError("contact_problem needs a ColorCache engine on box bodies (its friction frame is fixed)")
    ct = engine.contacts()
    N = engine.N
    x = engine.buffer(_lib.BUF_XSNAP, np.float32, (N, 3)).astype(np.float64)
    vf = engine.buffer(_lib.BUF_VSNAP, np.float32, (N, 3)).astype(np.float64)
    wf = engine.buffer(_lib.BUF_WSNAP, np.float32, (N, 3)).astype(np.float64)
    invIw = engine.buffer(_lib.BUF_INVIW, np.float32, (N, 3, 3)).astype(np.float64)
    invM = engine.invM.cpu().numpy().astype(np.float64)
    Cn = len(ct["key"])
    ends = np.concatenate([ct["bi"], ct["bj"]])
    bodies = np.unique(ends[(ends >= 0)])
    bodies = bodies[invM[bodies] > 0]
    col = {int(k): i for i, k in enumerate(bodies)}
    K = len(bodies)
    J = np.zeros((3 * Cn, 6 * K))
    for c in range(Cn):
        nrm = ct["n"][c].astype(np.float64)
        t1 = _fixed_t1(nrm)
        D = np.stack([nrm, t1, np.cross(nrm, t1)])
        for body, p, sign in ((int(ct["bi"][c]), ct["pA"][c], 1.0), (int(ct["bj"][c]), ct["pB"][c], -1.0)):
            if body in col:
                k = col[body]
                r = p.astype(np.float64) - x[body]
                J[3 * c:3 * c + 3, 6 * k:6 * k + 3] += sign * D
                J[3 * c:3 * c + 3, 6 * k + 3:6 * k + 6] += sign * np.cross(r[None, :], D)
    Minv = np.zeros((6 * K, 6 * K))
    V = np.zeros(6 * K)
    for k, body in enumerate(bodies):
        Minv[6 * k:6 * k + 3, 6 * k:6 * k + 3] = invM[body] * np.eye(3)
        Minv[6 * k + 3:6 * k + 6, 6 * k + 3:6 * k + 6] = invIw[body]
        V[6 * k:6 * k + 3], V[6 * k + 3:6 * k + 6] = vf[body], wf[body]
    lam = np.stack([ct["jn"], ct["jt1"], ct["jt2"]], axis=1).astype(np.float64).reshape(-1)
    return dict(G=J @ Minv @ J.T, b=J @ V, mu=ct["mu"].astype(np.float64), lam=lam, J=J, Minv=Minv,
                bodies=bodies)
