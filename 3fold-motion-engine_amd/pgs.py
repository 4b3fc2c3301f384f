"""Exact-cone NCP projected Gauss-Seidel on MI355X: the reference's NCPSolverGPU and its host-side
worlds (src/motion_engine/ncp/ncp_gpu.py).

The solver (`NCPSolver`, alias `NCPSolverGPU`) runs on libmotion_hip.so (me_pgs_* in
include/motion_hip.h); everything else here is host-side numpy, as it is in the reference: the
A/B contract (`Scene`, `BlockScene`), the TGS soft-step parameters, the box world with its SAT
contacts (`BoxWorld`), island sleeping, the N-body lattice fixture and the test scenes.
`ExactConeContactEngine` puts a box world behind the ContactEngine protocol.

Solve semantics (ncp_gpu.py:964-1053): upload, set-up (rho, warm start, v = v_free + M^-1 J^T
lambda), the colouring (cached while pairs, feature ids and masses are unchanged), `iters` sweeps,
then in soft-step mode with relax_iters > 0 the velocity snapshot `v_bias` and `relax_iters`
unbiased sweeps, then the natural-map residual of the last pass.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .protocol import BodyState

_S = 1.0e10        # int64 fixed-point scale of the per-body totals (ncp_gpu.py:98)
_BETA = 0.2        # Baumgarte coefficient (ncp_gpu.py:99)
_SLOP = 1e-4       # penetration allowance (ncp_gpu.py:100)
_G = 9.81
_NCOLMAX = 1024

# ── TGS soft step (soft_step.py:23-64, restated as ncp_gpu.py:105-183 does) ──
CONTACT_HERTZ = 30.0
CONTACT_DAMPING_RATIO = 10.0
CONTACT_SPEED = 3.0


@dataclass(frozen=True)
class Softness:
    bias_rate: float
    mass_scale: float
    impulse_scale: float


def make_soft(hertz: float, zeta: float, h: float) -> Softness:
    """Discrete implicit spring of stiffness `hertz`, damping ratio `zeta`, step h (soft_step.py:35-44);
    mass_scale + impulse_scale = 1, and hertz <= 0 is the rigid short-circuit (all zero)."""
    if hertz <= 0.0:
        return Softness(0.0, 0.0, 0.0)
    w = 2.0 * np.pi * hertz
    d = 2.0 * zeta + h * w
    k = h * w * d
    inv = 1.0 / (1.0 + k)
    return Softness(bias_rate=w / d, mass_scale=k * inv, impulse_scale=inv)


def contact_hertz(inv_h: float, world_hertz: float = CONTACT_HERTZ) -> float:
    """Contact stiffness capped at a quarter of the substep Nyquist rate (soft_step.py:47-49)."""
    return min(world_hertz, 0.125 * inv_h)


def bias_velocity(separation: float, soft: Softness, use_bias: bool, inv_h: float,
                  contact_speed: float = CONTACT_SPEED):
    """(velocity_bias, mass_scale, impulse_scale) of one contact (soft_step.py:52-64): a gap is closed
    speculatively in one step; a penetration is pushed out softly, clamped at -contact_speed; the
    relax pass (use_bias False) carries no bias."""
    if separation > 0.0:
        return separation * inv_h, 1.0, 0.0
    if not use_bias:
        return 0.0, 1.0, 0.0
    push = max(soft.mass_scale * soft.bias_rate * separation, -contact_speed)
    return push, soft.mass_scale, soft.impulse_scale


@dataclass
class SoftStep:
    """Soft-step configuration: `relax_frac` of each substep's iterations go to the unbiased relax pass
    (ncp_gpu.py:161-183)."""
    hertz: float = CONTACT_HERTZ
    zeta: float = CONTACT_DAMPING_RATIO
    contact_speed: float = CONTACT_SPEED
    relax_frac: float = 0.25

    def params(self, h: float) -> np.ndarray:
        """(bias_rate, mass_scale, impulse_scale, inv_h, contact_speed) for a substep of length h."""
        inv_h = 1.0 / h
        s = make_soft(contact_hertz(inv_h, self.hertz), self.zeta, h)
        return np.array([s.bias_rate, s.mass_scale, s.impulse_scale, inv_h, self.contact_speed], np.float64)

    def split(self, iters: int):
        """(biased, relax) iterations out of a budget of `iters`: at least one relax iteration when
        iters > 1, never all of them."""
        iters = max(1, int(iters))
        if iters == 1:
            return 1, 0
        relax = min(iters - 1, max(1, int(round(self.relax_frac * iters))))
        return iters - relax, relax


# ── the A/B contract (ncp_gpu.py:185-300) ──

@dataclass
class BlockScene:
    """Per-contact 3x6 Jacobian blocks and per-body 6x6 inverse masses; u_c = Ja_c v_a + Jb_c v_b +
    bias_c (Jb carries the minus of the relative velocity); body -1 is static."""
    Ja: np.ndarray
    Jb: np.ndarray
    Minv: np.ndarray
    v_free: np.ndarray
    mu: np.ndarray
    body_pairs: np.ndarray
    dt: float
    bias: np.ndarray
    feature: np.ndarray
    sep: Optional[np.ndarray] = None

    @property
    def n_c(self) -> int:
        return int(self.Ja.shape[0])

    @property
    def n_b(self) -> int:
        return int(self.Minv.shape[0])

    def _J(self) -> np.ndarray:
        J = np.zeros((3 * self.n_c, 6 * self.n_b))
        for c, (a, b) in enumerate(np.asarray(self.body_pairs).reshape(-1, 2)):
            if a >= 0:
                J[3 * c:3 * c + 3, 6 * a:6 * a + 6] += self.Ja[c]
            if b >= 0:
                J[3 * c:3 * c + 3, 6 * b:6 * b + 6] += self.Jb[c]
        return J

    def delassus(self) -> np.ndarray:
        """G = J M^-1 J^T, dense float64 (host-side comparisons only)."""
        J = self._J()
        M = np.zeros((6 * self.n_b, 6 * self.n_b))
        for i in range(self.n_b):
            M[6 * i:6 * i + 6, 6 * i:6 * i + 6] = self.Minv[i]
        return J @ M @ J.T

    def rhs(self) -> np.ndarray:
        """b = J v_free + bias, float64."""
        out = np.asarray(self.bias, np.float64).reshape(-1, 3).copy()
        for c, (a, b) in enumerate(np.asarray(self.body_pairs).reshape(-1, 2)):
            if a >= 0:
                out[c] += self.Ja[c] @ self.v_free[a]
            if b >= 0:
                out[c] += self.Jb[c] @ self.v_free[b]
        return out.reshape(-1)


@dataclass
class Scene:
    """The dense contract: J (3n_c x 6n_b), Minv (dense or its diagonal), v_free, mu, body_pairs, dt,
    optional bias (3n_c), feature ids and separations."""
    J: np.ndarray
    Minv: np.ndarray
    v_free: np.ndarray
    mu: np.ndarray
    body_pairs: np.ndarray
    dt: float
    bias: Optional[np.ndarray] = None
    feature: Optional[np.ndarray] = None
    sep: Optional[np.ndarray] = None

    def blocks(self) -> BlockScene:
        J = np.asarray(self.J, np.float64)
        n_c, n_v = J.shape[0] // 3, J.shape[1]
        if n_v % 6:
            raise ValueError(f"Scene.blocks: n_v={n_v} is not a multiple of 6 (6-DOF rigid bodies)")
        n_b = n_v // 6
        Mi = np.asarray(self.Minv, np.float64)
        if Mi.ndim == 1:
            Mi = np.diag(Mi)
        pairs = np.asarray(self.body_pairs, np.int32).reshape(n_c, 2)
        Ja, Jb = np.zeros((n_c, 3, 6)), np.zeros((n_c, 3, 6))
        for c, (a, b) in enumerate(pairs):
            if a >= 0:
                Ja[c] = J[3 * c:3 * c + 3, 6 * a:6 * a + 6]
            if b >= 0:
                Jb[c] = J[3 * c:3 * c + 3, 6 * b:6 * b + 6]
        Minv = np.stack([Mi[6 * i:6 * i + 6, 6 * i:6 * i + 6] for i in range(n_b)]) if n_b else np.zeros((0, 6, 6))
        return BlockScene(Ja=Ja, Jb=Jb, Minv=Minv, v_free=np.asarray(self.v_free, np.float64).reshape(n_b, 6),
                          mu=np.asarray(self.mu, np.float64).reshape(n_c), body_pairs=pairs, dt=float(self.dt),
                          bias=(np.zeros((n_c, 3)) if self.bias is None
                                else np.asarray(self.bias, np.float64).reshape(n_c, 3)),
                          feature=(np.arange(n_c, dtype=np.int32) if self.feature is None
                                   else np.asarray(self.feature, np.int32)),
                          sep=None if self.sep is None else np.asarray(self.sep, np.float64).reshape(n_c))


@dataclass
class SolveResult:
    lam: np.ndarray              # (n_c, 3) impulses in the contact frame (n, t1, t2)
    v: np.ndarray                # (n_b, 6) velocity after the contact impulses
    residual: float              # ||lambda - proj_K(lambda - rho (u + Gamma(u)))||_inf
    iters: int
    n_colors: int
    cache_hit: bool
    ms: dict = field(default_factory=dict)
    history: Optional[dict] = None
    v_bias: Optional[np.ndarray] = None   # velocity after the biased pass (soft step integrates with it)
    force_selection: str = "COLORED_PGS_ORDER_WARM_START"
    contact_model: str = "NCP_RIGID"
    bias_applied: bool = True


# ── the solver ──

def _torch():
    import torch
    return torch


class NCPSolver:
    """Persistent exact-cone PGS solver on one MI355X (NCPSolverGPU, ncp_gpu.py:783-1069). Capacity is
    fixed at construction; the colouring lives across `solve` calls (colour cache).

    `graph_capture` is accepted for the reference's signature: every block of iterations here is one
    kernel launch (one workgroup per tile, all colours and iterations inside), so there is no launch
    sequence to capture."""

    def __init__(self, n_bodies: int, max_contacts: int, dtype: str = "float32", device: Optional[str] = None,
                 desaxce: bool = True, graph_capture: bool = True, color_cache: bool = True):
        torch = _torch()
        if dtype not in ("float32", "float64"):
            raise ValueError(f"dtype must be 'float32' or 'float64', got {dtype!r}")
        if not torch.cuda.is_available():
            raise _lib.MotionError("NCPSolver needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.dtype = dtype
        self.np_t = np.float64 if dtype == "float64" else np.float32
        self.t_t = torch.float64 if dtype == "float64" else torch.float32
        self.rb = 8 if dtype == "float64" else 4
        self.dev = torch.device(device or "cuda:0")
        self.NB, self.MC = int(n_bodies), int(max_contacts)
        self.desaxce = 1 if desaxce else 0
        self.use_cache = bool(color_cache)
        self.want_graph = bool(graph_capture)
        L = _lib.lib()
        n = C.c_size_t()
        _lib.check(L.me_pgs_workspace_size(self.NB, self.MC, self.rb, C.byref(n)), "me_pgs_workspace_size")
        self._ws_raw = torch.empty(int(n.value) + 256, dtype=torch.uint8, device=self.dev)
        off = (-self._ws_raw.data_ptr()) % 256
        self._ws = self._ws_raw[off:off + int(n.value)]
        flags = (_lib.PGS_DESAXCE if desaxce else 0) | (_lib.PGS_COLOR_CACHE if color_cache else 0)
        h = C.c_void_p()
        _lib.check(L.me_pgs_create(self.NB, self.MC, self.rb, flags, C.c_void_p(self._ws.data_ptr()), n.value,
                                   C.byref(h)), "me_pgs_create")
        self._h = h
        self._dev_arrays = {}
        self._scene = _lib.PgsScene()
        self.C = 0
        self.n_colors = 0
        self.cache_hits = 0
        self.cache_misses = 0
        self.prof = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _torch().cuda.synchronize(self.dev)
                _lib.lib().me_pgs_destroy(h)
            except Exception:
                pass
            self._h = None

    def _stream(self):
        return C.c_void_p(_torch().cuda.current_stream(self.dev).cuda_stream)

    def _sync(self):
        _torch().cuda.synchronize(self.dev)

    def _put(self, name, arr, dtype):
        t = _torch().from_numpy(np.ascontiguousarray(arr, dtype)).to(self.dev)
        self._dev_arrays[name] = t
        return C.c_void_p(t.data_ptr()) if t.numel() else None

    def upload(self, bs: BlockScene):
        C_, NB = bs.n_c, bs.n_b
        if C_ > self.MC:
            raise ValueError(f"NCPSolver: {C_} contacts exceeds capacity {self.MC}")
        if NB != self.NB:
            raise ValueError(f"NCPSolver: scene has {NB} bodies, solver built for {self.NB}")
        t = self.np_t
        sc = self._scene
        sc.Ja = self._put("Ja", np.asarray(bs.Ja).reshape(C_, 18), t)
        sc.Jb = self._put("Jb", np.asarray(bs.Jb).reshape(C_, 18), t)
        sc.mu = self._put("mu", np.asarray(bs.mu).reshape(C_), t)
        sc.bias = self._put("bias", np.asarray(bs.bias).reshape(C_, 3), t)
        sep = np.zeros(C_) if bs.sep is None else np.asarray(bs.sep, np.float64).reshape(C_)
        sc.sep = self._put("sep", sep, t)
        sc.body_pairs = self._put("pab", np.asarray(bs.body_pairs).reshape(C_, 2), np.int32)
        sc.feature = self._put("feat", np.asarray(bs.feature).reshape(C_), np.int32)
        sc.Minv = self._put("Minv", np.asarray(bs.Minv).reshape(NB, 36), t)
        sc.v_free = self._put("vfree", np.asarray(bs.v_free).reshape(NB, 6), t)
        self.C = C_

    def _buf(self, which, dtype, count=None):
        ptr, n = C.c_void_p(), C.c_int64()
        _lib.check(_lib.lib().me_pgs_buffer(self._h, which, C.byref(ptr), C.byref(n)), "me_pgs_buffer")
        item = np.dtype(dtype).itemsize
        off = (ptr.value or 0) - self._ws.data_ptr()
        cnt = int(n.value) if count is None else int(count)
        tdt = {np.float32: _torch().float32, np.float64: _torch().float64, np.int32: _torch().int32}[dtype]
        return self._ws[off:off + cnt * item].view(tdt)

    def _info(self):
        out = (C.c_int64 * len(_lib.PGS_INFO))()
        _lib.check(_lib.lib().me_pgs_info(self._h, out), "me_pgs_info")
        return dict(zip(_lib.PGS_INFO, [int(x) for x in out]))

    def _iterate(self, n, use_bias):
        if n > 0:
            _lib.check(_lib.lib().me_pgs_iterate(self._h, int(n), int(use_bias), self._stream()), "me_pgs_iterate")

    def _residual(self, use_bias) -> float:
        r = C.c_double()
        _lib.check(_lib.lib().me_pgs_residual(self._h, int(use_bias), C.byref(r), self._stream()),
                   "me_pgs_residual")
        return float(r.value)

    def lam_device(self):
        return self._buf(_lib.PGS_LAM, self.np_t).view(-1, 3)

    def v_device(self):
        return self._buf(_lib.PGS_V, self.np_t).view(-1, 6)

    def solve(self, bs: BlockScene, iters: int = 200, warm: Optional[np.ndarray] = None,
              measure_residual: bool = True, history_every: int = 0, soft: Optional[np.ndarray] = None,
              relax_iters: int = 0) -> SolveResult:
        """`soft`: SoftStep.params(h) (None = the Baumgarte path of `bs.bias`); with relax_iters > 0,
        `iters` biased sweeps, the snapshot v_bias, then `relax_iters` unbiased sweeps continuing from
        the same lambda and accumulator."""
        self.prof = {}
        t0 = time.perf_counter()
        self.upload(bs)
        self._sync()
        self.prof["upload"] = (time.perf_counter() - t0) * 1e3
        C_ = self.C
        if C_ == 0:
            return SolveResult(np.zeros((0, 3)), np.asarray(bs.v_free, np.float64).copy(), 0.0, 0, 0, False,
                               dict(self.prof), v_bias=np.asarray(bs.v_free, np.float64).copy())
        t0 = time.perf_counter()
        wptr = None
        if warm is not None:
            wptr = self._put("warm", np.asarray(warm).reshape(C_, 3), self.np_t)
        sp = None
        if soft is not None:
            self._soft = (C.c_double * 5)(*[float(x) for x in np.asarray(soft, np.float64).reshape(5)])
            sp = self._soft
        hit = C.c_int32()
        _lib.check(_lib.lib().me_pgs_prepare(self._h, C.byref(self._scene), C_, wptr, sp, C.byref(hit),
                                             self._stream()), "me_pgs_prepare")
        self._sync()
        self.prof["setup"] = (time.perf_counter() - t0) * 1e3
        info = self._info()
        if info["overflow"]:
            raise RuntimeError(f"NCPSolver: more than {_NCOLMAX} colours")
        self.n_colors, self.cache_hits, self.cache_misses = info["n_colors"], info["cache_hits"], info["cache_misses"]
        t0 = time.perf_counter()
        hist = None
        if history_every > 0:
            hist = {"iter": [0], "res": [self._residual(1)], "lam": [self._lam_host()]}
            done = 0
            while done < iters:
                n = min(history_every, iters - done)
                self._iterate(n, 1)
                done += n
                hist["iter"].append(done)
                hist["res"].append(self._residual(1))
                hist["lam"].append(self._lam_host())
        else:
            self._iterate(iters, 1)
        v_bias = None
        if relax_iters > 0:
            _lib.check(_lib.lib().me_pgs_snapshot(self._h, self._stream()), "me_pgs_snapshot")
            v_bias = self._buf(_lib.PGS_VBIAS, self.np_t).cpu().numpy().reshape(-1, 6).astype(np.float64)
            self._iterate(relax_iters, 0)
        self._sync()
        self.prof["iter"] = (time.perf_counter() - t0) * 1e3
        r = 0.0
        if measure_residual:
            t0 = time.perf_counter()
            r = self._residual(0 if relax_iters > 0 else 1)
            self.prof["resid"] = (time.perf_counter() - t0) * 1e3
        return SolveResult(lam=self._lam_host(), v=self.v_device().cpu().numpy().astype(np.float64),
                           residual=r, iters=iters + int(relax_iters), n_colors=self.n_colors, cache_hit=bool(hit.value),
                           ms=dict(self.prof), history=hist, v_bias=v_bias,
                           contact_model="SOFT_STEP" if soft is not None else ("NCP_RIGID" if self.desaxce else "CCP"),
                           bias_applied=relax_iters <= 0)

    def _lam_host(self):
        return self.lam_device()[:self.C].cpu().numpy().astype(np.float64)

    def coloring(self):
        """(colour per contact, colour-grouped order, per-colour offset, per-colour count)."""
        col = self._buf(_lib.PGS_COLOR, np.int32).cpu().numpy()[:self.C].copy()
        order = np.argsort(col, kind="stable").astype(np.int32)
        count = np.bincount(col, minlength=self.n_colors).astype(np.int32)[:self.n_colors]
        start = np.concatenate([[0], np.cumsum(count)[:-1]]).astype(np.int32)
        return col, order, start, count

    def raw_bytes(self):
        """Raw device bytes of lambda and v (bit-identity checks)."""
        return (self.lam_device()[:self.C].cpu().numpy().tobytes(), self.v_device().cpu().numpy().tobytes())

    def contact_velocity(self):
        """u = G lambda + b of the last residual evaluation."""
        return self._buf(_lib.PGS_U, self.np_t).cpu().numpy().reshape(-1, 3)[:self.C].astype(np.float64)

    def rho(self):
        return self._buf(_lib.PGS_RHO, self.np_t).cpu().numpy()[:self.C].astype(np.float64)


NCPSolverGPU = NCPSolver


def solve_scene_gpu(bs: BlockScene, iters: int = 200, dtype: str = "float32", **kw) -> SolveResult:
    s = NCPSolver(bs.n_b, max(bs.n_c, 1), dtype=dtype, **kw)
    return s.solve(bs, iters=iters)


# ── rigid boxes, SAT contacts (ncp_gpu.py:1083-1234) ──

def _quat_to_R(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _qmul(a, b):
    """Hamilton product of (x, y, z, w) quaternions."""
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw,
                     aw * bw - ax * bx - ay * by - az * bz])


@dataclass
class Box:
    half: np.ndarray
    mass: float
    x: np.ndarray
    q: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, 0.0, 1.0]))
    v: np.ndarray = field(default_factory=lambda: np.zeros(3))
    w: np.ndarray = field(default_factory=lambda: np.zeros(3))
    kinematic: bool = False   # zero inverse mass, no gravity, not integrated (set_kinematic)

    @property
    def Ib(self):
        e = (2.0 * self.half) ** 2
        return self.mass / 12.0 * np.diag([e[1] + e[2], e[0] + e[2], e[0] + e[1]])


# corner k = (sx, sy, sz) with x slowest: k = 4 (sx > 0) + 2 (sy > 0) + (sz > 0)
_CORNER = np.array([[sx, sy, sz] for sx in (-1.0, 1.0) for sy in (-1.0, 1.0) for sz in (-1.0, 1.0)])


def _corners(b: Box):
    return b.x + (_CORNER * b.half) @ _quat_to_R(b.q).T


def _face_corners(half, x, R, axis, sgn):
    """Corners of the face (axis, sgn) in cyclic order, its outward normal and centre."""
    j, k = (axis + 1) % 3, (axis + 2) % 3
    n = sgn * R[:, axis]
    c = x + half[axis] * n
    u, v = half[j] * R[:, j], half[k] * R[:, k]
    return np.array([c + u + v, c - u + v, c - u - v, c + u - v]), n, c


def _clip(poly, c, ax, ha):
    """Sutherland-Hodgman clip of polygon `poly` to the slab |(p - c) . ax| <= ha (+side, then -side)."""
    for s in (1.0, -1.0):
        kept = []
        m = len(poly)
        for i in range(m):
            p, q = poly[i], poly[(i + 1) % m]
            dp = s * np.dot(p - c, ax) - ha
            dq = s * np.dot(q - c, ax) - ha
            if dp <= 0:
                kept.append(p)
            if dp * dq < 0:
                kept.append(p + (q - p) * (dp / (dp - dq)))
        poly = kept
        if not poly:
            return []
    return poly


def box_plane_contacts(bi, b: Box, margin=2e-3):
    """Corner-plane contacts against the ground z = 0 (normal +z, from the ground to the box):
    (body, -1, point, normal, gap, corner id)."""
    return [(bi, -1, p.copy(), np.array([0.0, 0.0, 1.0]), float(p[2]), k)
            for k, p in enumerate(_corners(b)) if p[2] < margin]


def box_box_contacts(ia, A: Box, ib, B: Box, margin=2e-3):
    """SAT over the 6 face and 9 edge axes, then reference/incident face clipping, at most 4 points.
    The normal points from B to A. A face axis of B must beat A's best by 1e-4 and an edge axis the
    best face axis by 1e-3, so near-parallel boxes keep their face manifold."""
    Ra, Rb = _quat_to_R(A.q), _quat_to_R(B.q)
    d = B.x - A.x

    def separation(ax):
        ra = float(np.sum(np.abs(Ra.T @ ax) * A.half))
        rb = float(np.sum(np.abs(Rb.T @ ax) * B.half))
        proj = float(np.dot(d, ax))
        return abs(proj) - (ra + rb), (ax if proj >= 0 else -ax)

    face = None   # (separation, owner, axis index, axis oriented from A to B)
    for owner, R in (("A", Ra), ("B", Rb)):
        for i in range(3):
            sp, ax = separation(R[:, i])
            if sp > margin:
                return []
            need = 1e-4 if (face is not None and owner == "B" and face[1] == "A") else 1e-12
            if face is None or sp > face[0] + need:
                face = (sp, owner, i, ax)
    edge = None
    for i in range(3):
        for j in range(3):
            cr = np.cross(Ra[:, i], Rb[:, j])
            nn = np.linalg.norm(cr)
            if nn <= 1e-6:
                continue
            sp, ax = separation(cr / nn)
            if sp > margin:
                return []
            if edge is None or sp > edge[0] + 1e-12:
                edge = (sp, "E", 3 * i + j, ax)
    if edge is not None and edge[0] > face[0] + 1e-3:
        sep, _, idx, nAB = edge
        ca, cb = _corners(A), _corners(B)
        pa, pb = ca[int(np.argmax(ca @ nAB))], cb[int(np.argmin(cb @ nAB))]
        return [(ia, ib, 0.5 * (pa + pb), -nAB, float(sep), 100 + idx)]
    sep, owner, idx, nAB = face
    if owner == "A":
        rh, rx, rR, ih, ix, iR = A.half, A.x, Ra, B.half, B.x, Rb
        rsgn = 1.0 if np.dot(Ra[:, idx], nAB) >= 0 else -1.0
        nBA = -nAB
    else:
        rh, rx, rR, ih, ix, iR = B.half, B.x, Rb, A.half, A.x, Ra
        rsgn = 1.0 if np.dot(Rb[:, idx], -nAB) >= 0 else -1.0
        nBA = nAB
    _, rn, rc = _face_corners(rh, rx, rR, idx, rsgn)
    # incident face: the face of the other box most anti-parallel to the reference normal
    dots = [float(np.dot(iR[:, a] * sg, rn)) for a in range(3) for sg in (1.0, -1.0)]
    m = int(np.argmin(dots))
    poly, _, _ = _face_corners(ih, ix, iR, m // 2, 1.0 if m % 2 == 0 else -1.0)
    j, k = (idx + 1) % 3, (idx + 2) % 3
    poly = _clip(list(poly), rc, rR[:, j], rh[j])
    if not poly:
        return []
    poly = _clip(poly, rc, rR[:, k], rh[k])
    if not poly:
        return []
    pts = [(float(np.dot(p - rc, rn)), p) for p in poly]
    pts = [gp for gp in pts if gp[0] < margin]
    if not pts:
        return []
    if len(pts) > 4:
        pts = sorted(pts, key=lambda gp: gp[0])[:4]
    base = (2 * idx + (0 if rsgn > 0 else 1)) * 8
    return [(ia, ib, p.copy(), nBA.copy(), g, base + n) for n, (g, p) in enumerate(pts)]


def _basis(n):
    """Tangents (t1, t2) of normal n: t1 from the x axis (the y axis when n is near x)."""
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t1 = a - np.dot(a, n) * n
    t1 /= np.linalg.norm(t1)
    return t1, np.cross(n, t1)


def contact_labels(lam, mu, tol_lam=1e-10, tol_rel=1e-6) -> np.ndarray:
    """Active-set label per contact: 0 open, 1 stick, 2 slip (ncp_gpu.py:1425-1437)."""
    L = np.asarray(lam, np.float64).reshape(-1, 3)
    mu = np.asarray(mu, np.float64).reshape(-1)
    scale = max(float(np.max(np.abs(L))) if L.size else 0.0, 1e-30)
    ln, lt = L[:, 0], np.linalg.norm(L[:, 1:3], axis=1)
    out = np.ones(len(L), np.int8)
    out[lt >= mu * ln * (1.0 - tol_rel) - tol_lam] = 2
    out[ln <= tol_lam + tol_rel * scale] = 0
    return out


# ── islands and sleeping (ncp_gpu.py:1440-1551) ──

@dataclass
class SleepConfig:
    enabled: bool = True
    v_sleep: float = 1e-3
    w_sleep: float = 1e-3
    k_steps: int = 8


def connected_islands(n_b: int, pairs: np.ndarray) -> np.ndarray:
    """Island label per body = the smallest body index of its component over dynamic-dynamic pairs
    (a static endpoint -1 never joins); independent of the pair order."""
    lab = np.arange(n_b, dtype=np.int64)
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    m = (pairs[:, 0] >= 0) & (pairs[:, 1] >= 0)
    a, b = pairs[m, 0], pairs[m, 1]
    if len(a) == 0:
        return lab
    while True:
        prev = lab
        lo = np.minimum(lab[a], lab[b])
        lab = lab.copy()
        np.minimum.at(lab, a, lo)
        np.minimum.at(lab, b, lo)
        for _ in range(64):          # pointer jumping (lab[i] <= i)
            nl = lab[lab]
            if np.array_equal(nl, lab):
                break
            lab = nl
        if np.array_equal(lab, prev):
            return lab


class Sleeper:
    """An island sleeps once every body has stayed below (v_sleep, w_sleep) for k_steps steps; a new
    contact pair, an external force or motion above the thresholds wakes it (ncp_gpu.py:1480-1551)."""

    def __init__(self, n_b: int, cfg: SleepConfig):
        self.cfg = cfg
        self.n_b = int(n_b)
        self.still = np.zeros(n_b, np.int64)
        self.asleep = np.zeros(n_b, bool)
        self.labels = np.arange(n_b, dtype=np.int64)
        self.prev_keys = None
        self._keys_raw = None

    def begin(self, pairs, v, w, forced=None) -> np.ndarray:
        cfg = self.cfg
        pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
        keys = pairs[:, 0] * (self.n_b + 1) + pairs[:, 1] + 1 if len(pairs) else np.zeros(0, np.int64)
        same = self._keys_raw is not None and np.array_equal(keys, self._keys_raw)
        if not same:
            if self.prev_keys is not None and len(keys):
                fresh = ~np.isin(keys, self.prev_keys)
                if fresh.any():
                    bd = np.concatenate([pairs[fresh, 0], pairs[fresh, 1]])
                    self.still[bd[bd >= 0]] = 0
            self.prev_keys = np.unique(keys)
            self._keys_raw = keys.copy()
        if forced is not None:
            self.still[np.asarray(forced, bool)] = 0
        moving = (np.linalg.norm(v, axis=1) > cfg.v_sleep) | (np.linalg.norm(w, axis=1) > cfg.w_sleep)
        self.still[moving] = 0
        if not same:
            self.labels = connected_islands(self.n_b, pairs)
        isl = np.full(self.n_b, np.iinfo(np.int64).max, np.int64)
        np.minimum.at(isl, self.labels, self.still)
        self.asleep = isl[self.labels] >= cfg.k_steps
        return self.asleep

    def end(self, v, w):
        cfg = self.cfg
        quiet = (np.linalg.norm(v, axis=1) <= cfg.v_sleep) & (np.linalg.norm(w, axis=1) <= cfg.w_sleep)
        awake = ~self.asleep
        self.still[awake & quiet] += 1
        self.still[awake & ~quiet] = 0

    def n_islands_asleep(self) -> int:
        return len(np.unique(self.labels[self.asleep])) if self.asleep.any() else 0


# ── the box world (ncp_gpu.py:1237-1414) ──

class BoxWorld:
    """Free boxes on a static ground plane z = 0: host-side SAT contacts, velocity-level step through the
    exact-cone solver on the GPU. `substeps` splits each step into S sub-steps of dt/S with `iters`
    iterations each; `soft` switches the solve to the TGS soft step (warm started, biased pass, position
    update with the biased velocity, relax pass)."""

    def __init__(self, dt=1.0 / 240.0, mu=0.5, gravity=True, dtype="float32", solver_kw=None, iters=200,
                 margin=2e-3, desaxce=True, sleep: Optional[SleepConfig] = None, substeps=1, warm_start=None,
                 soft: Optional[SoftStep] = None):
        self.bodies: list = []
        self.soft = soft
        self.warm_start = (soft is not None) if warm_start is None else bool(warm_start)
        self._warm = {}
        self.dt = float(dt)
        self.substeps = int(substeps)
        self.sleep_cfg = sleep
        self.sleeper: Optional[Sleeper] = None
        self.n_asleep = 0
        self.mu = float(mu)
        self.gravity = gravity
        self.iters = int(iters)
        self.margin = float(margin)
        self.dtype = dtype
        self.desaxce = desaxce
        self.solver_kw = solver_kw or {}
        self._solver = None
        self.last = None
        self.n_contacts = 0
        self.n_colors = 0

    def _make_solver(self, n_bodies, capacity):
        return NCPSolver(n_bodies, capacity, dtype=self.dtype, desaxce=self.desaxce, **self.solver_kw)

    def add(self, half, mass, x, v=None):
        self.bodies.append(Box(np.asarray(half, float), float(mass), np.asarray(x, float),
                               v=np.zeros(3) if v is None else np.asarray(v, float)))
        return len(self.bodies) - 1

    def contacts(self):
        out = []
        for i, b in enumerate(self.bodies):
            out += box_plane_contacts(i, b, self.margin)
        n = len(self.bodies)
        for i in range(n):
            for j in range(i + 1, n):
                out += box_box_contacts(i, self.bodies[i], j, self.bodies[j], self.margin)
        return out

    def _Minv(self):
        Mi = np.zeros((len(self.bodies), 6, 6))
        for i, b in enumerate(self.bodies):
            if b.kinematic:
                continue
            R = _quat_to_R(b.q)
            Mi[i, :3, :3] = np.eye(3) / b.mass
            Mi[i, 3:, 3:] = R @ np.linalg.inv(b.Ib) @ R.T
        return Mi

    def build_scene(self, f_ext=None, dt=None, contacts=None, asleep=None) -> BlockScene:
        n = len(self.bodies)
        dt = self.dt if dt is None else float(dt)
        Mi = self._Minv()
        vf = np.zeros((n, 6))
        for i, b in enumerate(self.bodies):
            f = np.array([0.0, 0.0, -_G]) * b.mass if self.gravity else np.zeros(3)
            if f_ext is not None:
                f = f + np.asarray(f_ext(i), float)
            vf[i, :3] = b.v + dt * Mi[i, :3, :3] @ f
            vf[i, 3:] = b.w
            if (asleep is not None and asleep[i]) or b.kinematic:
                vf[i] = 0.0
        cs = self.contacts() if contacts is None else contacts
        nc = len(cs)
        Ja, Jb = np.zeros((nc, 3, 6)), np.zeros((nc, 3, 6))
        pairs, bias = np.zeros((nc, 2), np.int32), np.zeros((nc, 3))
        feat, sep = np.zeros(nc, np.int32), np.zeros(nc)
        for c, (ia, ib, p, nrm, g, fid) in enumerate(cs):
            t1, t2 = _basis(nrm)
            D = np.stack([nrm, t1, t2])
            ra = p - self.bodies[ia].x
            Ja[c, :, :3] = D
            Ja[c, :, 3:] = np.cross(ra[None, :], D)
            if ib >= 0:
                rb = p - self.bodies[ib].x
                Jb[c, :, :3] = -D
                Jb[c, :, 3:] = -np.cross(rb[None, :], D)
            pairs[c] = (ia, ib)
            sep[c] = g
            if self.soft is None:   # a gap may close within the step; a penetration is pushed out
                bias[c, 0] = g / dt if g > 0 else -_BETA * max(-g - _SLOP, 0.0) / dt
            feat[c] = ia * 4096 + (ib + 1) * 256 + fid
        return BlockScene(Ja=Ja, Jb=Jb, Minv=Mi, v_free=vf, mu=np.full(nc, self.mu), body_pairs=pairs, dt=dt,
                          bias=bias, feature=feat, sep=sep)

    def step(self, f_ext=None, history_every=0):
        r = None
        for _ in range(max(1, self.substeps)):
            r = self._substep(f_ext, self.dt / max(1, self.substeps), history_every)
        return r

    def _substep(self, f_ext, dt, history_every=0):
        n = len(self.bodies)
        cs = self.contacts()
        asleep = None
        if self.sleep_cfg is not None and self.sleep_cfg.enabled:
            if self.sleeper is None:
                self.sleeper = Sleeper(n, self.sleep_cfg)
            pairs_all = np.array([[c[0], c[1]] for c in cs], np.int64).reshape(-1, 2)
            vel = np.array([b.v for b in self.bodies]).reshape(n, 3)
            omg = np.array([b.w for b in self.bodies]).reshape(n, 3)
            forced = np.zeros(n, bool)
            if f_ext is not None:
                forced = np.array([np.any(np.asarray(f_ext(i), float) != 0.0) for i in range(n)])
            asleep = self.sleeper.begin(pairs_all, vel, omg, forced)
            for i in range(n):
                if asleep[i]:
                    self.bodies[i].v = np.zeros(3)
                    self.bodies[i].w = np.zeros(3)
            cs = [c for c in cs if not ((c[0] >= 0 and asleep[c[0]]) or (c[1] >= 0 and asleep[c[1]]))]
            self.n_asleep = int(asleep.sum())
        bs = self.build_scene(f_ext, dt=dt, contacts=cs, asleep=asleep)
        if self._solver is None:
            self._solver = self._make_solver(n, max(4096, len(cs)))
        if bs.n_c:
            warm = None
            if self.warm_start and self._warm:
                warm = np.array([self._warm.get(int(f), (0.0, 0.0, 0.0)) for f in bs.feature])
            if self.soft is None:
                r = self._solver.solve(bs, iters=self.iters, warm=warm, history_every=history_every)
            else:
                nb_it, nr_it = self.soft.split(self.iters)
                r = self._solver.solve(bs, iters=nb_it, warm=warm, history_every=history_every,
                                       soft=self.soft.params(dt), relax_iters=nr_it)
                self.last_split = (nb_it, nr_it)
            if self.warm_start:
                self._warm = {int(f): tuple(l) for f, l in zip(bs.feature, r.lam)}
            v = r.v
            self.n_colors = r.n_colors
        else:
            v = bs.v_free
            self.n_colors = 0
            r = SolveResult(np.zeros((0, 3)), v, 0.0, 0, 0, False, v_bias=v)
        self.n_contacts = bs.n_c
        self.last = r
        # positions move with the biased pass's velocity; the body keeps the relaxed one
        vp = v if r.v_bias is None else r.v_bias
        for i, b in enumerate(self.bodies):
            if (asleep is not None and asleep[i]) or b.kinematic:
                continue
            b.v = v[i, :3].copy()
            b.w = v[i, 3:].copy()
            b.x = b.x + vp[i, :3] * dt
            qq = b.q + 0.5 * _qmul(np.array([vp[i, 3], vp[i, 4], vp[i, 5], 0.0]), b.q) * dt
            b.q = qq / np.linalg.norm(qq)
        if self.sleeper is not None:
            self.sleeper.end(np.array([b.v for b in self.bodies]).reshape(n, 3),
                             np.array([b.w for b in self.bodies]).reshape(n, 3))
        return r

    def penetration(self):
        return max([0.0] + [-g for (_, _, _, _, g, _) in self.contacts()])

    def kinetic_energy(self):
        return float(sum(0.5 * b.mass * b.v @ b.v + 0.5 * b.w @ (b.Ib @ b.w) for b in self.bodies))


# ── the two-layer N-body lattice (ncp_gpu.py:1554-1713) ──

def _quat_to_R_batch(q):
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((len(q), 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


class LatticeWorld:
    """Two-layer box lattice (lower layer on the ground, an upper box on each of the first n_up lower
    boxes): horizontal spacing 0.35, lower centres at 0.0999, the upper layer 0.1999 above. The pair set
    is structural, each pair a 4-point corner manifold (n_c = 4 N); contact normals are world +z and a
    stacked pair's plane passes through the lower box's top face centre."""

    def __init__(self, n=10000, half=0.1, spacing=0.35, z0=0.0999, dz=0.1999, mass=1.0, dt=1.0 / 240.0, mu=0.5,
                 iters=40, dtype="float32", sleep=None, solver_kw=None, points_per_pair=4):
        g = int(np.ceil(np.sqrt(n / 2.0)))
        self.n_low = min(g * g, n)
        self.n_up = n - self.n_low
        if self.n_up > self.n_low:
            raise ValueError("lattice: more upper than lower bodies")
        self.n, self.half, self.dt, self.mu, self.iters, self.dtype = int(n), float(half), float(dt), float(mu), \
            int(iters), dtype
        self.mass = np.full(n, float(mass))
        ix, iy = np.divmod(np.arange(self.n_low), g)
        x = np.zeros((n, 3))
        x[:self.n_low, 0] = (ix - (g - 1) / 2.0) * spacing
        x[:self.n_low, 1] = (iy - (g - 1) / 2.0) * spacing
        x[:self.n_low, 2] = z0
        x[self.n_low:, :2] = x[:self.n_up, :2]
        x[self.n_low:, 2] = z0 + dz
        self.x = x
        self.v, self.w = np.zeros((n, 3)), np.zeros((n, 3))
        self.q = np.tile(np.array([0.0, 0.0, 0.0, 1.0]), (n, 1))
        h = float(half)
        self.Ib = np.full(3, mass / 12.0 * ((2 * h) ** 2 + (2 * h) ** 2))
        off = (np.array([[-h, -h, -h], [h, -h, -h], [-h, h, -h], [h, h, -h]]) if points_per_pair == 4
               else np.array([[0.0, 0.0, -h]]))
        K = len(off)
        ca = np.concatenate([np.repeat(np.arange(self.n_low), K), np.repeat(np.arange(self.n_low, n), K)])
        cb = np.concatenate([np.full(self.n_low * K, -1), np.repeat(np.arange(self.n_up), K)])
        self.ca, self.cb = ca.astype(np.int64), cb.astype(np.int64)
        self.off = np.tile(off, (self.n_low + self.n_up, 1))
        self.feat = (np.arange(len(ca)) % K).astype(np.int32) + 8 * (self.cb + 1).astype(np.int32)
        self.pairs = np.stack([self.ca, self.cb], axis=1)
        self.sleep_cfg = sleep
        self.sleeper = None if sleep is None or not sleep.enabled else Sleeper(n, sleep)
        self.solver = NCPSolver(n, len(ca), dtype=dtype, **(solver_kw or {}))
        self.n_contacts = self.n_colors = self.n_asleep = 0
        self.last = None

    def _gaps_points(self, idx=None):
        ca = self.ca if idx is None else self.ca[idx]
        cb = self.cb if idx is None else self.cb[idx]
        off = self.off if idx is None else self.off[idx]
        R = _quat_to_R_batch(self.q)
        pa = self.x[ca] + np.einsum("nij,nj->ni", R[ca], off)
        gap = np.empty(len(ca))
        gnd = cb < 0
        gap[gnd] = pa[gnd, 2]
        up = ~gnd
        gap[up] = pa[up, 2] - (self.x[cb[up], 2] + R[cb[up], 2, 2] * self.half)
        return gap, pa

    def build_scene(self, f_ext=None, keep=None, asleep=None) -> BlockScene:
        n, dt = self.n, self.dt
        idx = np.arange(len(self.ca)) if keep is None else np.flatnonzero(keep)
        gap, pa = self._gaps_points(idx)
        nc = len(idx)
        Minv = np.zeros((n, 6, 6))
        Minv[:, 0, 0] = Minv[:, 1, 1] = Minv[:, 2, 2] = 1.0 / self.mass
        R = _quat_to_R_batch(self.q)
        Minv[:, 3:, 3:] = R @ (np.diag(1.0 / self.Ib)[None] @ np.transpose(R, (0, 2, 1)))
        f = np.zeros((n, 3))
        f[:, 2] = -_G * self.mass
        if f_ext is not None:
            f = f + f_ext
        vf = np.zeros((n, 6))
        vf[:, :3] = self.v + dt * f / self.mass[:, None]
        vf[:, 3:] = self.w
        if asleep is not None:
            vf[asleep] = 0.0
        D = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
        a, b = self.ca[idx], self.cb[idx]
        ra = pa - self.x[a]
        Ja, Jb = np.zeros((nc, 3, 6)), np.zeros((nc, 3, 6))
        Ja[:, :, :3] = D
        Ja[:, :, 3:] = np.cross(ra[:, None, :], D[None, :, :])
        hasb = b >= 0
        rb = np.zeros((nc, 3))
        rb[hasb] = pa[hasb] - self.x[b[hasb]]
        Jb[hasb, :, :3] = -D
        Jb[hasb, :, 3:] = -np.cross(rb[hasb][:, None, :], D[None, :, :])
        bias = np.zeros((nc, 3))
        bias[:, 0] = np.where(gap > 0, gap / dt, -_BETA * np.maximum(-gap - _SLOP, 0.0) / dt)
        return BlockScene(Ja=Ja, Jb=Jb, Minv=Minv, v_free=vf, mu=np.full(nc, self.mu),
                          body_pairs=np.stack([a, b], axis=1).astype(np.int32), dt=dt, bias=bias,
                          feature=self.feat[idx])

    def step(self, f_ext=None):
        n = self.n
        asleep = keep = None
        if self.sleeper is not None:
            forced = None if f_ext is None else np.any(np.asarray(f_ext) != 0.0, axis=1)
            asleep = self.sleeper.begin(self.pairs, self.v, self.w, forced)
            self.v[asleep] = 0.0
            self.w[asleep] = 0.0
            keep = ~(asleep[self.ca] | np.where(self.cb >= 0, asleep[np.maximum(self.cb, 0)], False))
            self.n_asleep = int(asleep.sum())
        bs = self.build_scene(f_ext, keep=keep, asleep=asleep)
        self.n_contacts = bs.n_c
        if bs.n_c:
            r = self.solver.solve(bs, iters=self.iters, measure_residual=False)
            v6 = r.v
            self.n_colors = r.n_colors
        else:
            v6 = bs.v_free
            self.n_colors = 0
            r = SolveResult(np.zeros((0, 3)), v6, 0.0, 0, 0, False)
        self.last = r
        mv = np.ones(n, bool) if asleep is None else ~asleep
        self.v[mv] = v6[mv, :3]
        self.w[mv] = v6[mv, 3:]
        self.x[mv] = self.x[mv] + self.v[mv] * self.dt
        w4 = np.zeros((n, 4))
        w4[:, :3] = self.w
        q = self.q
        qd = np.stack([w4[:, 3] * q[:, 0] + w4[:, 0] * q[:, 3] + w4[:, 1] * q[:, 2] - w4[:, 2] * q[:, 1],
                       w4[:, 3] * q[:, 1] - w4[:, 0] * q[:, 2] + w4[:, 1] * q[:, 3] + w4[:, 2] * q[:, 0],
                       w4[:, 3] * q[:, 2] + w4[:, 0] * q[:, 1] - w4[:, 1] * q[:, 0] + w4[:, 2] * q[:, 3],
                       w4[:, 3] * q[:, 3] - w4[:, 0] * q[:, 0] - w4[:, 1] * q[:, 1] - w4[:, 2] * q[:, 2]], axis=1)
        qq = q.copy()
        qq[mv] = q[mv] + 0.5 * qd[mv] * self.dt
        self.q = qq / np.linalg.norm(qq, axis=1, keepdims=True)
        if self.sleeper is not None:
            self.sleeper.end(self.v, self.w)
        return r


def lattice_pushed_forces(w: LatticeWorld, frac=0.1, fx=40.0):
    """A horizontal force on every 1/frac-th stacked island (its lower and upper box together)."""
    f = np.zeros((w.n, 3))
    j = np.arange(0, w.n_up, int(round(1.0 / frac)))
    f[j, 0] = fx
    f[w.n_low + j, 0] = fx
    return f


# ── scenes (ncp_gpu.py:1719-1854) ──

def scene_pushed_cube(dtype="float32", steps=500, iters=200, solver_kw=None, world_cls=None):
    """0.20 m, 15 kg cube, mu 0.5, dt 4 ms, a horizontal 147.15 N push alternating x / y every 0.5 s,
    ramped over 0.3 s."""
    dt, F, per, ramp = 0.004, 147.15, 0.5, 0.3
    w = (world_cls or BoxWorld)(dt=dt, mu=0.5, dtype=dtype, iters=iters, solver_kw=solver_kw)
    w.add([0.1, 0.1, 0.1], 15.0, [0.0, 0.0, 0.1])
    traj, pen, res = [], [], []
    for s in range(steps):
        t = s * dt
        k = int(t // per)
        amp = F * min(1.0, (t - k * per) / ramp)
        f = np.array([amp, 0.0, 0.0]) if k % 2 == 0 else np.array([0.0, amp, 0.0])
        r = w.step(f_ext=lambda i, f=f: f)
        traj.append(w.bodies[0].x.copy())
        pen.append(w.penetration())
        res.append(r.residual)
    return w, np.array(traj), np.array(pen), np.array(res)


def scene_sliding_box(dtype="float32", steps=200, iters=200, desaxce=True, v0=1.0, mu=0.3, dt=1.0 / 240.0,
                      substeps=1, soft=None, world_cls=None):
    """A 0.2 m, 1 kg box sliding at v0 on mu: the normal gap and the speed per step."""
    w = (world_cls or BoxWorld)(dt=dt, mu=mu, dtype=dtype, iters=iters, desaxce=desaxce, substeps=substeps,
                                soft=soft)
    w.add([0.1, 0.1, 0.1], 1.0, [0.0, 0.0, 0.1], v=[v0, 0.0, 0.0])
    gap, vx = [], []
    for _ in range(steps):
        w.step()
        gap.append(w.bodies[0].x[2] - 0.1)
        vx.append(w.bodies[0].v[0])
    return w, np.array(gap), np.array(vx)


def sliding_distance(dtype="float32", iters=400, desaxce=True, v0=1.0, mu=0.3, dt=1.0 / 240.0, max_steps=4000,
                     substeps=1, soft=None, world_cls=None):
    """Distance slid until the box stops (first step with vx <= 0) and the closed form v0^2 / (2 mu g)."""
    w = (world_cls or BoxWorld)(dt=dt, mu=mu, dtype=dtype, iters=iters, desaxce=desaxce, substeps=substeps,
                                soft=soft)
    w.add([0.1, 0.1, 0.1], 1.0, [0.0, 0.0, 0.1], v=[v0, 0.0, 0.0])
    x0 = w.bodies[0].x[0]
    n = 0
    for n in range(1, max_steps + 1):
        w.step()
        if w.bodies[0].v[0] <= 0.0:
            break
    return float(w.bodies[0].x[0] - x0), v0 * v0 / (2.0 * mu * _G), n


def scene_massratio(ratio, dtype="float32", steps=400, iters=200, substeps=1, soft=None, world_cls=None):
    """Three stacked 0.2 m boxes, the top one `ratio` times heavier: penetration per step."""
    H = 0.20
    w = (world_cls or BoxWorld)(dt=1.0 / 240.0, mu=0.5, dtype=dtype, iters=iters, substeps=substeps, soft=soft)
    for k in range(3):
        w.add([H / 2, H / 2, H / 2], ratio if k == 2 else 1.0, [0.0, 0.0, (k + 0.5) * H])
    pen = []
    for _ in range(steps):
        w.step()
        pen.append(w.penetration())
    return w, np.array(pen)


def scene_tower(K=8, dtype="float32", steps=400, iters=200, sleep=None, substeps=1, soft=None, world_cls=None):
    """A K-box tower; per-step penetration, and the jitter floor (max speed / spin) in
    w.vmax_history / w.wmax_history, the sleeping count in w.asleep_history."""
    H = 0.20
    w = (world_cls or BoxWorld)(dt=1.0 / 240.0, mu=0.5, dtype=dtype, iters=iters, sleep=sleep,
                                substeps=substeps, soft=soft)
    for k in range(K):
        w.add([H / 2, H / 2, H / 2], 1.0, [0.0, 0.0, (k + 0.5) * H])
    pen, asleep, vmax, wmax = [], [], [], []
    for _ in range(steps):
        w.step()
        pen.append(w.penetration())
        asleep.append(w.n_asleep)
        vmax.append(max(float(np.linalg.norm(b.v)) for b in w.bodies))
        wmax.append(max(float(np.linalg.norm(b.w)) for b in w.bodies))
    w.asleep_history, w.vmax_history, w.wmax_history = np.array(asleep), np.array(vmax), np.array(wmax)
    return w, np.array(pen)


def synthetic_scene(n_c, n_b=512, seed=0, mu=0.5):
    """n_c random contacts between n_b free bodies, built directly as blocks."""
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n_b, n_c)
    b = (a + 1 + rng.integers(0, n_b - 1, n_c)) % n_b
    nrm = rng.normal(size=(n_c, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    Ja, Jb = np.zeros((n_c, 3, 6)), np.zeros((n_c, 3, 6))
    for c in range(n_c):
        t1, t2 = _basis(nrm[c])
        D = np.stack([nrm[c], t1, t2])
        ra, rb = rng.uniform(-0.1, 0.1, 3), rng.uniform(-0.1, 0.1, 3)
        Ja[c, :, :3] = D
        Ja[c, :, 3:] = np.cross(ra[None, :], D)
        Jb[c, :, :3] = -D
        Jb[c, :, 3:] = -np.cross(rb[None, :], D)
    Ib = np.linalg.inv(1.0 / 12.0 * np.diag([0.08, 0.08, 0.08]))
    Mi = np.tile(np.block([[np.eye(3), np.zeros((3, 3))], [np.zeros((3, 3)), Ib]]), (n_b, 1, 1))
    vf = rng.normal(scale=0.2, size=(n_b, 6))
    return BlockScene(Ja=Ja, Jb=Jb, Minv=Mi, v_free=vf, mu=np.full(n_c, mu),
                      body_pairs=np.stack([a, b], axis=1).astype(np.int32), dt=1.0 / 240.0, bias=np.zeros((n_c, 3)),
                      feature=np.arange(n_c, dtype=np.int32))


def random_wellposed_scene(n_c=8, n_b=8, seed=1, mu=0.4):
    """3 n_c <= 6 n_b with full-rank J: the Delassus is SPD and the NCP solution unique."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n_b)
    a = perm[np.arange(n_c) % n_b]
    b = perm[(np.arange(n_c) + 3) % n_b]
    Ja = rng.normal(size=(n_c, 3, 6))
    Jb = rng.normal(size=(n_c, 3, 6))
    Mi = np.tile(np.eye(6), (n_b, 1, 1)) * rng.uniform(0.5, 2.0, (n_b, 1, 1))
    vf = rng.normal(scale=0.5, size=(n_b, 6))
    bias = np.zeros((n_c, 3))
    bias[:, 0] = -rng.uniform(0.1, 1.0, n_c)
    return BlockScene(Ja=Ja, Jb=Jb, Minv=Mi, v_free=vf, mu=np.full(n_c, mu),
                      body_pairs=np.stack([a, b], axis=1).astype(np.int32), dt=1.0 / 240.0, bias=bias,
                      feature=np.arange(n_c, dtype=np.int32))


# ── the fourth ContactEngine member ──

class ExactConeContactEngine:
    """A box world stepped by the exact-cone PGS behind the ContactEngine protocol
    (contact_engine.py:49-55): step(dt, substeps), get_state, set_kinematic, contact_forces.
    Kinematic bodies enter with zero inverse mass (static to the solver) and keep the pose
    `set_kinematic` gives them."""
    name = "exact_cone_pgs"

    def __init__(self, mu=0.5, iters=200, dtype="float32", desaxce=True, soft: Optional[SoftStep] = None,
                 margin=2e-3, solver_kw=None):
        self.world = BoxWorld(mu=mu, iters=iters, dtype=dtype, desaxce=desaxce, soft=soft, margin=margin,
                              solver_kw=solver_kw)
        self._kin = set()
        self._force = np.zeros((0, 3))

    def add_box(self, center, half, mass, v=None):
        return self.world.add(half, mass, center, v=v)

    def step(self, dt: float, substeps: int = 1) -> None:
        w = self.world
        w.dt, w.substeps = float(dt), int(substeps)
        v_pre = np.array([b.v for b in w.bodies]).reshape(-1, 3)
        w.step()
        m = np.array([b.mass for b in w.bodies])
        g = np.array([0.0, 0.0, -_G]) if w.gravity else np.zeros(3)
        v_post = np.array([b.v for b in w.bodies]).reshape(-1, 3)
        # contact force = M (v_post - v_pre - g dt) / dt, kinematic bodies excluded
        self._force = m[:, None] * (v_post - v_pre - g[None, :] * dt) / dt
        for i in self._kin:
            self._force[i] = 0.0

    def get_state(self) -> BodyState:
        bs = self.world.bodies
        return BodyState(xc=np.array([b.x for b in bs]).reshape(-1, 3),
                         Rm=np.array([_quat_to_R(b.q) for b in bs]).reshape(-1, 3, 3),
                         vc=np.array([b.v for b in bs]).reshape(-1, 3),
                         om=np.array([b.w for b in bs]).reshape(-1, 3))

    def set_kinematic(self, i: int, xc) -> None:
        b = self.world.bodies[i]
        b.x = np.asarray(xc, float).copy()
        b.v, b.w = np.zeros(3), np.zeros(3)
        b.kinematic = True
        self._kin.add(i)

    def contact_forces(self) -> np.ndarray:
        return self._force.copy()
