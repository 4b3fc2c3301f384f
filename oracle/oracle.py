"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke
check and the benchmark's CPU-baseline leg; never by the product)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
LIB64 = HERE / "_build" / "liboracle64.so"  # the same restatement in float64 (FD reference of the adjoint)
LIBMP = HERE / "_build" / "liboracle_omp.so"  # float32 with OpenMP over a colour's pairs (timing only)

JACOBI, PAIR_CHUNKS, COLOR_CACHE = 0, 1, 2
SOLVERS = {"jacobi": JACOBI, "pair_chunks": PAIR_CHUNKS, "color_cache": COLOR_CACHE}


class Params(C.Structure):
    _fields_ = [("solver", C.c_int), ("vit", C.c_int), ("pit", C.c_int), ("relax", C.c_float),
                ("t2d", C.c_float), ("warm_start", C.c_int), ("R", C.c_float),
                ("margin", C.c_float), ("cap", C.c_int), ("n_colors", C.c_int),
                ("per_contact", C.c_int), ("manifold", C.c_int)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


_libs = {}


def lib(fp64=False, omp=False):
    key = (bool(fp64), bool(omp) and not fp64)
    if key not in _libs:
        path = LIB64 if fp64 else (LIBMP if omp else LIB)
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        vp = C.c_void_p
        real = C.c_double if fp64 else C.c_float
        L.or_create.restype = vp
        L.or_create.argtypes = [C.c_int, C.c_int, C.POINTER(Params)]
        L.or_destroy.argtypes = [vp]
        L.or_set_bodies.argtypes = [vp] + [vp] * 10
        L.or_set_rest.argtypes = [vp, vp]
        L.or_set_mass.argtypes = [vp, vp]
        L.or_set_velocity.argtypes = [vp, vp, vp]
        L.or_step.argtypes = [vp, C.c_double, C.c_int]
        L.or_step.restype = C.c_int
        L.or_get_state.argtypes = [vp] * 5
        L.or_get_aux.argtypes = [vp] * 4
        L.or_contact_forces.argtypes = [vp, vp]
        L.or_counters.argtypes = [vp, vp]
        L.or_get_contacts.argtypes = [vp] * 12
        L.or_get_contacts.restype = C.c_int
        L.or_get_pairs.argtypes = [vp] * 6
        L.or_get_pairs.restype = C.c_int
        L.or_quat_to_matrix.argtypes = [vp, vp]
        L.or_quat_rotate.argtypes = [vp, vp, vp]
        L.or_integrate.argtypes = [C.c_int] + [vp] * 8 + [real]
        L.or_box_box.argtypes = [vp] * 6 + [real] + [vp] * 4
        L.or_box_box.restype = C.c_int
        L.or_box_ground.argtypes = [vp] * 3 + [real] + [vp] * 3
        L.or_box_ground.restype = C.c_int
        L.or_pcg.argtypes = [C.c_uint32]
        L.or_pcg.restype = C.c_uint32
        L.or_jp_key.argtypes = [C.c_int]
        L.or_jp_key.restype = C.c_int64
        L.or_time_steps.argtypes = [vp, C.c_double, C.c_int, C.c_int]
        L.or_time_steps.restype = C.c_double
        _libs[key] = L
    return _libs[key]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a, shape):
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(shape))


class OracleWorld:
    """One simulated world on the CPU restatement (fp64=True: the float64 build)."""

    def __init__(self, bodies, solver="jacobi", vit=40, pit=20, relax=0.2, t2d=1.0,
                 warm_start=True, R=0.05, margin=0.0, cap=None, n_colors=0, per_contact=False, fp64=False,
                 manifold=None, omp=False):
        self.fp64 = bool(fp64)
        self.omp = bool(omp) and not self.fp64
        self.dtype = np.float64 if fp64 else np.float32
        L = lib(self.fp64, self.omp)
        _f32 = self._arr
        N = bodies.n
        P = int(len(bodies.rest))
        if cap is None:
            cap = max(1024, 4 * N * 8) if P == 0 else max(1024, N * P * 8)
        if manifold is None:  # the coloured bead engines reduce to 4 points per pair
            manifold = 4 if (P > 0 and solver != "jacobi" and not per_contact) else 0
        prm = Params(SOLVERS[solver], vit, pit, relax, t2d, int(warm_start), R, margin, cap,
                     n_colors, int(per_contact), int(manifold))
        self._prm = prm
        self.N, self.P, self.cap = N, P, cap
        self.h = L.or_create(N, P, C.byref(prm))
        self._keep = dict(
            x=_f32(bodies.x, (N, 3)), q=_f32(bodies.q, (N, 4)), v=_f32(bodies.v, (N, 3)),
            w=_f32(bodies.w, (N, 3)), invM=_f32(bodies.invM, (N,)),
            IbInv=_f32(bodies.IbInv, (N, 3, 3)),
            kin=np.ascontiguousarray(bodies.kin, np.int32), half=_f32(bodies.half, (N, 3)),
            mu=_f32(bodies.mu, (N,)), scene=np.ascontiguousarray(bodies.scene, np.int32))
        k = self._keep
        L.or_set_bodies(self.h, _p(k["x"]), _p(k["q"]), _p(k["v"]), _p(k["w"]), _p(k["invM"]),
                        _p(k["IbInv"]), _p(k["kin"]), _p(k["half"]), _p(k["mu"]), _p(k["scene"]))
        M = np.ascontiguousarray(bodies.M, np.float64)
        L.or_set_mass(self.h, _p(M))
        if P:
            rest = _f32(bodies.rest, (P, 3))
            L.or_set_rest(self.h, _p(rest))

    def _arr(self, a, shape):
        return np.ascontiguousarray(np.asarray(a, self.dtype).reshape(shape))

    def __del__(self):
        try:
            lib(self.fp64, self.omp).or_destroy(self.h)
        except Exception:
            pass

    def step(self, dt, substeps=1):
        rc = lib(self.fp64, self.omp).or_step(self.h, float(dt), int(substeps))
        if rc == 4:
            raise RuntimeError("oracle: contact capacity exceeded")
        if rc:
            raise RuntimeError(f"oracle: step failed ({rc})")

    def time_steps(self, dt, substeps, steps):
        return lib(self.fp64, self.omp).or_time_steps(self.h, float(dt), int(substeps), int(steps))

    def state(self):
        N = self.N
        x = np.zeros((N, 3), self.dtype); q = np.zeros((N, 4), self.dtype)
        v = np.zeros((N, 3), self.dtype); w = np.zeros((N, 3), self.dtype)
        lib(self.fp64, self.omp).or_get_state(self.h, _p(x), _p(q), _p(v), _p(w))
        return x, q, v, w

    def aux(self):
        N = self.N
        pv = np.zeros((N, 3), self.dtype); po = np.zeros((N, 3), self.dtype)
        I = np.zeros((N, 3, 3), self.dtype)
        lib(self.fp64, self.omp).or_get_aux(self.h, _p(pv), _p(po), _p(I))
        return pv, po, I

    def forces(self):
        F = np.zeros((self.N, 3), np.float64)
        lib(self.fp64, self.omp).or_contact_forces(self.h, _p(F))
        return F

    def counters(self):
        out = np.zeros(7, np.int64)
        lib(self.fp64, self.omp).or_counters(self.h, _p(out))
        return dict(n_contacts=int(out[0]), n_pairs=int(out[1]), n_colors=int(out[2]),
                    cache_hits=int(out[3]), cache_misses=int(out[4]), overflow=int(out[5]),
                    n_generated=int(out[6]))

    def contacts(self):
        C_ = self.counters()["n_contacts"]
        key = np.zeros(C_, np.int64); bi = np.zeros(C_, np.int32); bj = np.zeros(C_, np.int32)
        pA = np.zeros((C_, 3), self.dtype); pB = np.zeros((C_, 3), self.dtype)
        n = np.zeros((C_, 3), self.dtype); pen = np.zeros(C_, self.dtype)
        jn = np.zeros(C_, self.dtype); jt1 = np.zeros(C_, self.dtype)
        jt2 = np.zeros(C_, self.dtype); jp = np.zeros(C_, self.dtype)
        lib(self.fp64, self.omp).or_get_contacts(self.h, _p(key), _p(bi), _p(bj), _p(pA), _p(pB), _p(n), _p(pen),
                              _p(jn), _p(jt1), _p(jt2), _p(jp))
        return dict(key=key, bi=bi, bj=bj, pA=pA, pB=pB, n=n, pen=pen, jn=jn, jt1=jt1, jt2=jt2,
                    jp=jp)

    def pairs(self):
        P = self.counters()["n_pairs"]
        a = np.zeros(P, np.int32); b = np.zeros(P, np.int32); s = np.zeros(P, np.int32)
        c = np.zeros(P, np.int32); col = np.zeros(P, np.int32)
        lib(self.fp64, self.omp).or_get_pairs(self.h, _p(a), _p(b), _p(s), _p(c), _p(col))
        return dict(pbi=a, pbj=b, pstart=s, pcount=c, color=col)


def box_box(xa, qa, ha, xb, qb, hb, margin=0.0):
    p = np.zeros((4, 3), np.float32); n = np.zeros((4, 3), np.float32)
    pen = np.zeros(4, np.float32); code = np.zeros(4, np.int32)
    args = [_f32(a, (k,)) for a, k in ((xa, 3), (qa, 4), (ha, 3), (xb, 3), (qb, 4), (hb, 3))]
    k = lib().or_box_box(*[_p(a) for a in args], float(margin), _p(p), _p(n), _p(pen), _p(code))
    return p[:k], n[:k], pen[:k], code[:k]


def box_ground(x, q, h, margin=0.0):
    p = np.zeros((4, 3), np.float32); pen = np.zeros(4, np.float32); code = np.zeros(4, np.int32)
    args = [_f32(a, (k,)) for a, k in ((x, 3), (q, 4), (h, 3))]
    k = lib().or_box_ground(*[_p(a) for a in args], float(margin), _p(p), _p(pen), _p(code))
    return p[:k], pen[:k], code[:k]


def quat_to_matrix(q):
    qq = _f32(q, (4,)); R = np.zeros((3, 3), np.float32)
    lib().or_quat_to_matrix(_p(qq), _p(R))
    return R
