"""The adjoint's hand-written reverses against central differences of their own forwards, in double
precision on the CPU (tests/native/adjoint_check.hip over csrc/adjoint_math.hpp and
csrc/adjoint_geom.hpp, the code the GPU adjoint kernels instantiate in float).

Each check draws a random point away from the update's kinks (clamps, cone projection, face and
axis choices), a random output adjoint u, and compares the reverse's input adjoint with
(u . f(x + h e_i) - u . f(x - h e_i)) / 2h for every input coordinate i."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
SRC = HERE / "native" / "adjoint_check.hip"
LIB = HERE / "native" / "libadjoint_check.so"


def _lib():
    if not LIB.exists() or LIB.stat().st_mtime < max(SRC.stat().st_mtime, *(p.stat().st_mtime for p in (
            HERE.parent / "3fold-motion-engine_amd" / "csrc").glob("adjoint_*.hpp"))):
        cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
               "-ffp-contract=off", "-o", str(LIB), str(SRC)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            pytest.skip("hipcc unavailable for the adjoint check library: " + r.stderr[-300:])
    L = C.CDLL(str(LIB))
    dp = C.POINTER(C.c_double)
    for f in ("adc_contact_fwd", "adc_prep_fwd", "adc_integ_fwd", "adc_invI_fwd"):
        getattr(L, f).argtypes = [dp, dp]
    for f in ("adc_contact_rev", "adc_prep_rev", "adc_integ_rev", "adc_invI_rev"):
        getattr(L, f).argtypes = [dp, dp, dp]
    L.adc_narrow_fwd.argtypes = [dp, C.c_int, C.c_float, dp]
    L.adc_narrow_rev.argtypes = [dp, C.c_int, C.c_float, dp, dp]
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fd_check(fwd, rev, x, n_out, idx, rng, h=1e-6, rtol=2e-6, atol=1e-8):
    u = rng.standard_normal(n_out)
    got = rev(x, u)
    for i in idx:
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (u @ fwd(xp) - u @ fwd(xm)) / (2 * h)
        assert abs(fd - got[i]) <= atol + rtol * max(abs(fd), abs(got[i])), f"input {i}: fd {fd!r} reverse {got[i]!r}"


def _unit(v):
    return v / np.linalg.norm(v)


def _spd(rng):
    a = rng.standard_normal((3, 3))
    return (a @ a.T + 3 * np.eye(3)) * 0.5


def _contact_input(rng, kind, hasB=True, dynB=True, mu=0.5, approach=-1.0):
    x = np.zeros(60)
    x[0:12] = rng.standard_normal(12) * 0.3
    n = _unit(rng.standard_normal(3))
    t1 = _unit(np.cross(n, rng.standard_normal(3)))
    t2 = np.cross(n, t1)
    # relative normal velocity of the given sign so that the normal clamp is open or closed
    x[0:3] += n * approach
    x[12:16] = [0.3, 0.05, -0.04, 0.1]
    con = np.concatenate([n, rng.standard_normal(3) * 0.1, rng.standard_normal(3) * 0.1, t1, t2,
                          [1.7, 1.9, 2.1, 0.4, mu]])
    x[16:36] = con
    x[36:45] = _spd(rng).ravel()
    x[45:54] = _spd(rng).ravel()
    x[54:56] = [0.8, 1.3]
    x[56:59] = [float(hasB), 1.0, float(dynB and hasB)]
    x[59] = kind
    return x


@pytest.mark.parametrize("kind,hasB,dynB,mu,approach", [
    (0, True, True, 0.5, -1.0), (0, True, True, 0.02, -1.0), (0, True, True, 5.0, -1.0), (0, True, True, 0.5, 1.0),
    (0, False, False, 0.3, -1.0), (0, True, False, 0.3, -1.0),
    (1, True, True, 0.5, -1.0), (1, False, False, 0.5, -0.2), (1, True, True, 0.5, 3.0),
    (2, True, True, 0.5, 0.0), (2, False, False, 0.5, 0.0)])
def test_contact_reverse_matches_central_differences(kind, hasB, dynB, mu, approach):
    L = _lib()
    rng = np.random.default_rng(100 + kind * 10 + int(hasB) * 3 + int(dynB) + int(mu * 7))
    x = _contact_input(rng, kind, hasB, dynB, mu, approach)

    def fwd(v):
        out = np.zeros(16)
        L.adc_contact_fwd(_p(np.ascontiguousarray(v)), _p(out))
        return out

    def rev(v, u):
        inb = np.zeros(60)
        L.adc_contact_rev(_p(np.ascontiguousarray(v)), _p(np.ascontiguousarray(u)), _p(inb))
        return inb

    # inputs: state, impulses, constants, both inverse inertias (not the masses or flags)
    idx = [i for i in range(54) if not (not hasB and (6 <= i < 12 or 22 <= i < 25 or 45 <= i < 54))]
    _fd_check(fwd, rev, x, 16, idx, rng)


def test_prep_reverse_matches_central_differences():
    L = _lib()
    rng = np.random.default_rng(7)
    for hasB in (True, False):
        for nx in (0.3, 0.95):  # both branches of the normal-derived tangent
            x = np.zeros(39)
            n = _unit(np.array([nx, 0.4, 0.6]) + rng.standard_normal(3) * 0.02)
            x[0:3] = n
            x[3:15] = rng.standard_normal(12) * 0.2
            x[15:24] = _spd(rng).ravel()
            x[24:33] = _spd(rng).ravel()
            x[33:39] = [2e-3, 0.45, 1 / 480, 0.7, 1.1, float(hasB)]

            def fwd(v):
                out = np.zeros(20)
                L.adc_prep_fwd(_p(np.ascontiguousarray(v)), _p(out))
                return out

            def rev(v, u):
                inb = np.zeros(39)
                L.adc_prep_rev(_p(np.ascontiguousarray(v)), _p(np.ascontiguousarray(u)), _p(inb))
                return inb

            idx = [i for i in range(35) if hasB or not (6 <= i < 9 or 12 <= i < 15 or 24 <= i < 33)]
            _fd_check(fwd, rev, x, 20, idx, rng)


def test_integrate_and_inertia_reverse_match_central_differences():
    L = _lib()
    rng = np.random.default_rng(11)
    x = np.zeros(20)
    x[0:3] = rng.standard_normal(3)
    x[3:7] = _unit(rng.standard_normal(4))
    x[7:19] = rng.standard_normal(12)
    x[19] = 1 / 480

    def fwd(v):
        out = np.zeros(7)
        L.adc_integ_fwd(_p(np.ascontiguousarray(v)), _p(out))
        return out

    def rev(v, u):
        inb = np.zeros(19)
        L.adc_integ_rev(_p(np.ascontiguousarray(v)), _p(np.ascontiguousarray(u)), _p(inb))
        return inb

    _fd_check(fwd, rev, x, 7, range(19), rng)

    y = np.concatenate([_unit(rng.standard_normal(4)), _spd(rng).ravel()])

    def fwdI(v):
        out = np.zeros(9)
        L.adc_invI_fwd(_p(np.ascontiguousarray(v)), _p(out))
        return out

    def revI(v, u):
        qb = np.zeros(4)
        L.adc_invI_rev(_p(np.ascontiguousarray(v)), _p(np.ascontiguousarray(u)), _p(qb))
        return qb

    _fd_check(fwdI, revI, y, 9, range(4), rng)


def _yaw_tilt(yaw, tilt):
    a = np.array([np.sin(tilt / 2), 0, 0, np.cos(tilt / 2)])
    b = np.array([0, 0, np.sin(yaw / 2), np.cos(yaw / 2)])
    # Hamilton product b * a (xyzw)
    x1, y1, z1, w1 = b
    x2, y2, z2, w2 = a
    return np.array([w1 * x2 + w2 * x1 + y1 * z2 - y2 * z1, w1 * y2 + w2 * y1 + z1 * x2 - z2 * x1,
                     w1 * z2 + w2 * z1 + x1 * y2 - x2 * y1, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2])


@pytest.mark.parametrize("case", ["ground_tilted", "box_box_face", "box_box_clipped"])
def test_narrow_phase_duals_match_central_differences(case):
    L = _lib()
    rng = np.random.default_rng({"ground_tilted": 1, "box_box_face": 2, "box_box_clipped": 3}[case])
    x = np.zeros(20)
    if case == "ground_tilted":
        x[0:3] = [0.1, -0.2, 0.048]
        x[3:7] = _yaw_tilt(0.3, 0.05)
        x[7:10] = [0.05, 0.04, 0.05]
        ground = 1
    elif case == "box_box_face":
        x[0:3] = [0.003, -0.002, 0.0995]
        x[3:7] = _yaw_tilt(0.2, 0.01)
        x[7:10] = [0.05, 0.05, 0.05]
        x[10:13] = [0.0, 0.0, 0.0]
        x[13:17] = _yaw_tilt(-0.1, 0.0)
        x[17:20] = [0.07, 0.06, 0.05]
        ground = 0
    else:
        x[0:3] = [0.06, 0.05, 0.0985]
        x[3:7] = _yaw_tilt(0.4, 0.02)
        x[7:10] = [0.05, 0.04, 0.05]
        x[10:13] = [0.0, 0.0, 0.0]
        x[13:17] = _yaw_tilt(-0.2, 0.0)
        x[17:20] = [0.06, 0.06, 0.05]
        ground = 0
    out0 = np.zeros(29)
    k = L.adc_narrow_fwd(_p(x), ground, 0.0, _p(out0))
    assert k >= 1
    idx = list(range(7)) + ([] if ground else list(range(10, 17)))

    def fwd(v):
        out = np.zeros(29)
        kk = L.adc_narrow_fwd(_p(np.ascontiguousarray(v)), ground, 0.0, _p(out))
        assert kk == k, "perturbation changed the contact set"
        return out[1:1 + 7 * k]

    def rev(v, u):
        g = np.zeros(14)
        L.adc_narrow_rev(_p(np.ascontiguousarray(v)), ground, 0.0, _p(np.ascontiguousarray(u)), _p(g))
        full = np.zeros(20)
        full[0:7] = g[0:7]
        full[10:17] = g[7:14]
        return full

    _fd_check(fwd, rev, x, 7 * k, idx, rng, h=1e-7, rtol=1e-5, atol=1e-7)
