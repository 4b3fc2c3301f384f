"""Exact-cone NCP projected Gauss-Seidel on MI355X, behind the reference module's interface
(src/motion_engine/ncp/ncp_gpu.py): the solver `NCPSolver` (alias `NCPSolverGPU`) on libmotion_hip.so
(me_pgs_* in include/motion_hip.h), the A/B contract it solves (`Scene`, `BlockScene`,
`SolveResult`), the TGS soft-step parameters, and -- from `boxworld` -- the box world, island
sleeping, the N-body lattice and the test scenes, so that a caller changes only the import.

Solve semantics (ncp_gpu.py:964-1053): upload, set-up (rho, warm start, v = v_free + M^-1 J^T
lambda), the colouring (cached while pairs, feature ids and masses are unchanged), `iters` sweeps,
then in soft-step mode with relax_iters > 0 the velocity snapshot `v_bias` and `relax_iters`
unbiased sweeps, then the natural-map residual of the last pass.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib

_NCOLMAX = 1024

# ── TGS soft step: the contact as a discrete implicit spring (soft_step.py:23-64) ──
CONTACT_HERTZ = 30.0
CONTACT_DAMPING_RATIO = 10.0
CONTACT_SPEED = 3.0


@dataclass(frozen=True)
class Softness:
    bias_rate: float
    mass_scale: float
    impulse_scale: float


def make_soft(hertz: float, zeta: float, h: float) -> Softness:
    """Coefficients of a spring of frequency `hertz` and damping ratio `zeta` integrated implicitly
    over h: with omega = 2 pi hertz and k = h omega (2 zeta + h omega), the bias rate is
    omega / (2 zeta + h omega), the mass scale k / (1 + k) and the impulse scale 1 / (1 + k) (they
    sum to 1). A non-positive frequency means a rigid contact: all three are zero."""
    if not hertz > 0.0:
        return Softness(0.0, 0.0, 0.0)
    omega = 2.0 * math.pi * hertz
    denom = 2.0 * zeta + h * omega
    k = h * omega * denom
    return Softness(bias_rate=omega / denom, mass_scale=k / (1.0 + k), impulse_scale=1.0 / (1.0 + k))


def contact_hertz(inv_h: float, world_hertz: float = CONTACT_HERTZ) -> float:
    """The contact frequency, at most an eighth of the substep rate (a quarter of Nyquist)."""
    return min(world_hertz, inv_h / 8.0)


def bias_velocity(separation: float, soft: Softness, use_bias: bool, inv_h: float,
                  contact_speed: float = CONTACT_SPEED):
    """(velocity bias, mass scale, impulse scale) of one contact: an open gap closes within the
    substep (rigid scales); a penetration is pushed out at the soft rate, no faster than
    contact_speed, on the biased pass only."""
    if separation > 0.0:
        return separation * inv_h, 1.0, 0.0
    if use_bias:
        return max(soft.mass_scale * soft.bias_rate * separation, -contact_speed), soft.mass_scale, \
            soft.impulse_scale
    return 0.0, 1.0, 0.0


@dataclass
class SoftStep:
    """A soft-step configuration; `relax_frac` of each substep's iterations form the unbiased relax
    pass (ncp_gpu.py:161-183)."""
    hertz: float = CONTACT_HERTZ
    zeta: float = CONTACT_DAMPING_RATIO
    contact_speed: float = CONTACT_SPEED
    relax_frac: float = 0.25

    def params(self, h: float) -> np.ndarray:
        """The solver's five numbers for a substep of length h: bias rate, mass scale, impulse scale,
        1/h, contact speed."""
        s = make_soft(contact_hertz(1.0 / h, self.hertz), self.zeta, h)
        return np.array([s.bias_rate, s.mass_scale, s.impulse_scale, 1.0 / h, self.contact_speed], np.float64)

    def split(self, iters: int):
        """(biased, relax) iterations of a budget: at least one of each when the budget allows two."""
        n = max(1, int(iters))
        relax = 0 if n == 1 else min(n - 1, max(1, int(round(self.relax_frac * n))))
        return n - relax, relax


# ── the A/B contract (ncp_gpu.py:185-300) ──

@dataclass
class BlockScene:
    """One 3x6 Jacobian block per contact side and one 6x6 inverse mass per body: the contact velocity
    is u_c = Ja_c v_a + Jb_c v_b + bias_c (Jb carries the minus of the relative velocity); body -1 is
    static."""
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
        return int(np.shape(self.Ja)[0])

    @property
    def n_b(self) -> int:
        return int(np.shape(self.Minv)[0])

    def _sides(self):
        pairs = np.asarray(self.body_pairs, np.int64).reshape(-1, 2)
        return [(np.flatnonzero(pairs[:, s] >= 0), pairs[:, s], blk) for s, blk in ((0, self.Ja), (1, self.Jb))]

    def _J(self) -> np.ndarray:
        """The dense 3 n_c x 6 n_b Jacobian."""
        J = np.zeros((self.n_c, 3, self.n_b, 6))
        for rows, body, blk in self._sides():
            np.add.at(J, (rows, slice(None), body[rows]), np.asarray(blk)[rows])
        return J.reshape(3 * self.n_c, 6 * self.n_b)

    def delassus(self) -> np.ndarray:
        """G = J M^-1 J^T, dense float64 (host-side comparisons only)."""
        J = self._J()
        M = np.zeros((self.n_b, 6, self.n_b, 6))
        M[np.arange(self.n_b), :, np.arange(self.n_b), :] = np.asarray(self.Minv)
        return J @ M.reshape(6 * self.n_b, 6 * self.n_b) @ J.T

    def rhs(self) -> np.ndarray:
        """b = J v_free + bias, float64."""
        out = np.asarray(self.bias, np.float64).reshape(-1, 3).copy()
        vf = np.asarray(self.v_free, np.float64)
        for rows, body, blk in self._sides():
            out[rows] += np.matmul(np.asarray(blk)[rows], vf[body[rows]][:, :, None])[..., 0]
        return out.reshape(-1)


@dataclass
class Scene:
    """The dense contract: J (3 n_c x 6 n_b), Minv (dense or its diagonal), v_free, mu, body_pairs, dt,
    optional bias (3 n_c), feature ids and separations."""
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
        Jr = J.reshape(n_c, 3, n_b, 6)

        def side(s):
            return np.where((pairs[:, s] >= 0)[:, None, None],
                            Jr[np.arange(n_c), :, np.maximum(pairs[:, s], 0), :], 0.0)

        Minv = Mi.reshape(n_b, 6, n_b, 6)[np.arange(n_b), :, np.arange(n_b), :]
        return BlockScene(Ja=side(0), Jb=side(1), Minv=Minv,
                          v_free=np.asarray(self.v_free, np.float64).reshape(n_b, 6),
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
    contact_info: dict = field(default_factory=lambda: {"status": "UNASSESSED_CONTACT_UNIQUENESS"})

    def diagnose_contacts(self, scene, rcond=1e-12, max_contacts=256):
        """Opt-in bounded diagnostic of the solved impulses (SolveResult.diagnose_contacts,
        ncp_gpu.py:311-347): at the labels of this result, the numerical rank and nullspaces of the
        fixed-branch residual Jacobian and whether the body wrenches J^T lam are invariant observables
        (ncp.contact_diagnostics on the GPU, observable = J^T). `scene` is the scene solved for this
        result (a BlockScene or a dense Scene). A numerical fixed-branch statement, not a global
        uniqueness proof; soft-step and relaxed-cone results stay unassessed. The GPU diagnostic holds
        at most 256 contacts: larger scenes, or more than max_contacts, are UNASSESSED_SIZE_LIMIT."""
        bs = scene.blocks() if hasattr(scene, "blocks") else scene
        limit = min(int(max_contacts), 256)
        if self.contact_model != "NCP_RIGID":
            self.contact_info = {"status": "UNASSESSED_CONTACT_MODEL", "scope": self.contact_model}
        elif bs.n_c > limit:
            self.contact_info = {"status": "UNASSESSED_SIZE_LIMIT", "max_contacts": limit}
        else:
            if np.shape(self.lam) != (bs.n_c, 3):
                raise ValueError("scene and result contact counts do not agree")
            from . import ncp
            J = bs._J()
            G, b = bs.delassus(), bs.rhs()
            if not self.bias_applied:
                b = b - np.asarray(bs.bias, np.float64).reshape(-1)
            lam = np.asarray(self.lam, np.float64).reshape(-1)
            mu = np.asarray(bs.mu, np.float64)
            labels = ncp.classify(lam, G, b, mu)
            self.contact_info = ncp.contact_diagnostics(lam, G, b, mu, list(labels), rcond=rcond, observable=J.T)
        self.contact_info["force_selection"] = self.force_selection
        return self.contact_info


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


# the host-side worlds and scenes of the reference module, from this package's own box world
from .boxworld import (Box, BoxWorld, ExactConeContactEngine, LatticeWorld, SleepConfig, Sleeper,  # noqa: E402
                       connected_islands, contact_labels, lattice_pushed_forces, random_wellposed_scene,
                       scene_massratio, scene_pushed_cube, scene_sliding_box, scene_tower, sliding_distance,
                       synthetic_scene)
