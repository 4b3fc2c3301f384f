"""ctypes wrapper of the exact-cone PGS oracle (pgs_oracle.c). TEST INFRASTRUCTURE ONLY: imported by
tests/, the smoke check and the benchmark's CPU-baseline leg; never by the product.

`solve` is one NCPSolverGPU.solve (src/motion_engine/ncp/ncp_gpu.py:964-1053) from a fresh solver,
restated on the CPU in float32 or float64. `OracleNCPSolver` gives it the solver's call shape so the
host-side worlds can run on it."""
from __future__ import annotations

import ctypes as C
import types
from pathlib import Path

import numpy as np

from . import oracle as _orc

HERE = Path(__file__).resolve().parent
LIBS = {"float32": HERE / "_build" / "libpgs32.so", "float64": HERE / "_build" / "libpgs64.so"}
_libs = {}


def lib(dtype="float32"):
    if dtype not in _libs:
        path = LIBS[dtype]
        if not path.exists():
            _orc.build()
        L = C.CDLL(str(path))
        sfx = "f64" if dtype == "float64" else "f32"
        real = C.c_double if dtype == "float64" else C.c_float
        vp = C.c_void_p
        f = getattr(L, f"pgs_solve_{sfx}")
        f.argtypes = [C.c_int, C.c_int] + [vp] * 11 + [C.c_int, C.c_int, C.c_int] + [vp] * 8
        f.restype = C.c_int
        g = getattr(L, f"pgs_eigmax_{sfx}")
        g.argtypes = [vp]
        g.restype = real
        for n in ("acos", "cos"):
            h = getattr(L, f"pgs_{n}_{sfx}")
            h.argtypes = [real]
            h.restype = real
        p = getattr(L, f"pgs_proj_{sfx}")
        p.argtypes = [vp, real]
        _libs[dtype] = (L, sfx)
    return _libs[dtype]


def _np(dtype):
    return np.float64 if dtype == "float64" else np.float32


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def solve(bs, iters=200, warm=None, soft=None, relax_iters=0, desaxce=True, dtype="float32"):
    """One solve of BlockScene `bs` (attributes Ja, Jb, Minv, v_free, mu, body_pairs, bias, feature,
    sep). Returns lam (n_c,3), v (n_b,6), v_bias, rho, col, n_colors, residual, res, u, all in the
    solver's precision."""
    L, sfx = lib(dtype)
    t = _np(dtype)
    n_c, n_b = int(bs.Ja.shape[0]), int(bs.Minv.shape[0])
    arr = lambda x, shape: np.ascontiguousarray(np.asarray(x, t).reshape(shape))
    Ja, Jb = arr(bs.Ja, (n_c, 3, 6)), arr(bs.Jb, (n_c, 3, 6))
    Minv, vf = arr(bs.Minv, (n_b, 6, 6)), arr(bs.v_free, (n_b, 6))
    mu, bias = arr(bs.mu, (n_c,)), arr(bs.bias, (n_c, 3))
    sep = arr(np.zeros(n_c) if getattr(bs, "sep", None) is None else bs.sep, (n_c,))
    pab = np.ascontiguousarray(np.asarray(bs.body_pairs, np.int32).reshape(n_c, 2))
    feat = np.ascontiguousarray(np.asarray(bs.feature, np.int32).reshape(n_c))
    w = None if warm is None else arr(warm, (n_c, 3))
    sp = None if soft is None else np.ascontiguousarray(np.asarray(soft, np.float64).reshape(5))
    lam = np.zeros((n_c, 3), t)
    v = np.zeros((n_b, 6), t)
    vb = np.zeros((n_b, 6), t)
    rho = np.zeros(n_c, t)
    col = np.zeros(n_c, np.int32)
    res = np.zeros(n_c, t)
    u = np.zeros((n_c, 3), t)
    r = C.c_double(0.0)
    nc = getattr(L, f"pgs_solve_{sfx}")(n_b, n_c, _ptr(Ja), _ptr(Jb), _ptr(Minv), _ptr(vf), _ptr(mu), _ptr(bias),
                                        _ptr(sep), _ptr(pab), _ptr(feat), _ptr(w), _ptr(sp), int(bool(desaxce)),
                                        int(iters), int(relax_iters), _ptr(lam), _ptr(v), _ptr(vb), _ptr(rho),
                                        _ptr(col), _ptr(res), _ptr(u), C.byref(r))
    return types.SimpleNamespace(lam=lam, v=v, v_bias=vb, rho=rho, col=col, n_colors=int(nc),
                                 residual=float(r.value), res=res, u=u)


def eigmax(A, dtype="float64"):
    L, sfx = lib(dtype)
    a = np.ascontiguousarray(np.asarray(A, _np(dtype)).reshape(9))
    return float(getattr(L, f"pgs_eigmax_{sfx}")(_ptr(a)))


def acos(x, dtype="float64"):
    L, sfx = lib(dtype)
    return float(getattr(L, f"pgs_acos_{sfx}")(float(x)))


def cos(x, dtype="float64"):
    L, sfx = lib(dtype)
    return float(getattr(L, f"pgs_cos_{sfx}")(float(x)))


def proj(z, mu, dtype="float64"):
    L, sfx = lib(dtype)
    a = np.ascontiguousarray(np.asarray(z, _np(dtype)).reshape(3)).copy()
    getattr(L, f"pgs_proj_{sfx}")(_ptr(a), float(mu))
    return a


class OracleNCPSolver:
    """The solver's call shape over `solve` (stateless: it colours every call, which a colour-cache
    hit does not change)."""

    def __init__(self, n_bodies, max_contacts, dtype="float32", desaxce=True, **_kw):
        self.NB, self.MC, self.dtype, self.desaxce = int(n_bodies), int(max_contacts), dtype, bool(desaxce)
        self.n_colors = 0

    def solve(self, bs, iters=200, warm=None, measure_residual=True, soft=None, relax_iters=0, **_kw):
        n_c = int(bs.Ja.shape[0])
        if n_c > self.MC:
            raise ValueError(f"{n_c} contacts exceeds capacity {self.MC}")
        r = solve(bs, iters=iters, warm=warm, soft=soft, relax_iters=relax_iters, desaxce=self.desaxce,
                  dtype=self.dtype)
        self.n_colors = r.n_colors
        return types.SimpleNamespace(lam=r.lam.astype(np.float64), v=r.v.astype(np.float64),
                                     v_bias=r.v_bias.astype(np.float64) if relax_iters > 0 else None,
                                     residual=r.residual if measure_residual else 0.0, iters=iters + relax_iters,
                                     n_colors=r.n_colors, cache_hit=False)
