"""ctypes binding of libmotion_hip.so (the C ABI in include/motion_hip.h).

The library must be built in-tree (``build.py``); there is no fallback: a missing library, or a
process without a ROCm GPU, raises instead of running anything else."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmotion_hip.so"

ME_OK, ME_EARG, ME_ERUNTIME, ME_ETHREAD, ME_ECAPACITY, ME_ENONFINITE = 0, 1, 2, 3, 4, 5
SOLVER_JACOBI, SOLVER_PAIR_CHUNKS, SOLVER_COLOR_CACHE = 0, 1, 2
FLAG_TILED, FLAG_PER_CONTACT, FLAG_GRAPH, FLAG_FAST_MATH = 1, 2, 4, 8

(BUF_KEY, BUF_BI, BUF_BJ, BUF_PA, BUF_PB, BUF_N, BUF_PEN, BUF_JN, BUF_JT1, BUF_JT2, BUF_JP,
 BUF_PBI, BUF_PBJ, BUF_PSTART, BUF_PCOUNT, BUF_COLOR, BUF_VSNAP, BUF_PV, BUF_PO, BUF_INVIW,
 BUF_ISLAND, BUF_DEBUG, BUF_BIG_WORK, BUF_MU, BUF_XSNAP, BUF_WSNAP, BUF_TILE_WORK) = range(27)

COUNTERS = ("n_contacts", "n_pairs", "n_colors", "cache_hits", "cache_misses", "overflow",
            "n_candidates", "n_islands", "n_tiles", "max_tile_bodies", "graph_captures", "n_solved")
OVF_CONTACTS, OVF_PAIRS, OVF_CAND, OVF_COLORS, OVF_TILE, OVF_TAPE, OVF_GIANT, OVF_ADJ = 1, 2, 4, 8, 16, 32, 64, 128
OVF_SCAN = 256


class Config(C.Structure):
    _fields_ = [("solver", C.c_int32), ("n_bodies", C.c_int32), ("n_beads", C.c_int32),
                ("vit", C.c_int32), ("pit", C.c_int32), ("relax", C.c_float), ("t2d", C.c_float),
                ("warm_start", C.c_int32), ("bead_radius", C.c_float), ("margin", C.c_float),
                ("max_pairs", C.c_int32), ("max_contacts", C.c_int32), ("n_colors", C.c_int32),
                ("tile_bodies", C.c_int32), ("flags", C.c_int32)]


class Bodies(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("x", "q", "v", "w", "inv_mass", "inv_inertia", "half",
                                           "mu", "kinematic", "scene", "rest")]


class PgsScene(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("Ja", "Jb", "Minv", "v_free", "mu", "bias", "sep", "body_pairs",
                                           "feature")]


PGS_DESAXCE, PGS_COLOR_CACHE = 1, 2
PGS_LAM, PGS_V, PGS_VBIAS, PGS_RHO, PGS_RES, PGS_U, PGS_COLOR = range(7)
PGS_INFO = ("n_colors", "cache_hits", "cache_misses", "n_islands", "n_tiles", "overflow")


class NcpSvdOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("X", "S", "V", "U", "rank", "compat", "finite", "resid", "Y", "valid",
                                           "invariant", "leak", "labels", "status")]


class MotionError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise MotionError(f"{LIB_PATH} is not built: run build.py (hipcc --offload-arch=gfx950)")
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        L.me_workspace_size.argtypes = [C.POINTER(Config), C.POINTER(C.c_size_t)]
        L.me_create.argtypes = [C.POINTER(Config), C.POINTER(Bodies), vp, C.c_size_t, vp, C.POINTER(vp)]
        L.me_destroy.argtypes = [vp]
        L.me_set_bodies.argtypes = [vp, C.POINTER(Bodies)]
        L.me_step.argtypes = [vp, C.c_double, C.c_int, vp]
        L.me_buffer.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(C.c_int64)]
        L.me_read_counters.argtypes = [vp, C.POINTER(C.c_int64), vp]
        L.me_profile.argtypes = [vp, C.c_int]
        L.me_profile_read.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.me_tape_size.argtypes = [vp, C.POINTER(C.c_size_t)]
        L.me_adopt.argtypes = [vp]
        L.me_state_size.argtypes = [vp, C.POINTER(C.c_size_t)]
        L.me_state_copy.argtypes = [vp, vp, vp, vp]
        L.me_ncp_workspace_size.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_size_t)]
        L.me_ncp_sensitivity.argtypes = [C.c_int32, C.c_int32, C.c_int32] + [vp] * 6 + [C.c_double] * 3 + \
            [vp] * 5 + [C.c_size_t, vp]
        L.me_ncp_sensitivity_noalloc.argtypes = [C.c_int32, C.c_int32, C.c_int32] + [vp] * 6 + [C.c_double] * 3 + \
            [vp] * 4 + [vp]
        L.me_ncp_svd_workspace_size.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_size_t)]
        L.me_ncp_svd.argtypes = [C.c_int32] * 4 + [vp] * 7 + [C.c_int32] + [C.c_double] * 4 + \
            [C.c_int32, C.POINTER(NcpSvdOut), vp, C.c_size_t, vp]
        L.me_pgs_workspace_size.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_size_t)]
        L.me_pgs_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp, C.c_size_t, C.POINTER(vp)]
        L.me_pgs_destroy.argtypes = [vp]
        L.me_pgs_prepare.argtypes = [vp, C.POINTER(PgsScene), C.c_int32, vp, vp, C.POINTER(C.c_int32), vp]
        L.me_pgs_iterate.argtypes = [vp, C.c_int32, C.c_int32, vp]
        L.me_pgs_snapshot.argtypes = [vp, vp]
        L.me_pgs_residual.argtypes = [vp, C.c_int32, C.POINTER(C.c_double), vp]
        L.me_pgs_buffer.argtypes = [vp, C.c_int32, C.POINTER(vp), C.POINTER(C.c_int64)]
        L.me_pgs_info.argtypes = [vp, vp]
        L.me_step_taped.argtypes = [vp, C.c_double, vp, C.c_size_t, vp]
        L.me_adjoint_substep.argtypes = [vp, C.c_double, vp, C.c_size_t] + [vp] * 7 + [vp]
        L.me_last_error.restype = C.c_char_p
        for f in ("me_workspace_size", "me_create", "me_destroy", "me_set_bodies", "me_step",
                  "me_buffer", "me_read_counters", "me_profile", "me_profile_read"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc == ME_OK:
        return
    msg = lib().me_last_error().decode(errors="replace")
    if rc == ME_EARG:
        raise ValueError(f"{what}: {msg}")
    raise MotionError(f"{what}: {msg} (code {rc})")


EXPORTED = ("me_workspace_size", "me_create", "me_destroy", "me_set_bodies", "me_step",
            "me_buffer", "me_read_counters", "me_profile", "me_profile_read", "me_last_error",
            "me_tape_size", "me_step_taped", "me_adjoint_substep", "me_adopt", "me_state_size", "me_state_copy",
            "me_ncp_workspace_size", "me_ncp_sensitivity", "me_pgs_workspace_size", "me_pgs_create",
            "me_pgs_destroy", "me_pgs_prepare", "me_pgs_iterate", "me_pgs_snapshot", "me_pgs_residual",
            "me_pgs_buffer", "me_pgs_info", "me_ncp_sensitivity_noalloc", "me_ncp_svd_workspace_size",
            "me_ncp_svd")
K_SWEEP_VEL, K_SWEEP_POS, K_NARROW, K_TILED, K_JAC_VEL, K_TILE_BIG, K_TILE_SMALL, K_TILE_GIANT, K_ISLANDS, \
    K_COLOR = range(10)
