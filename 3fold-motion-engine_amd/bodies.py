"""Host-side body construction (float64 on the host, float32 on the device).

Two body models share one device layout (SoA float32: x[N,3], q[N,4] xyzw, v, w, invM[N],
IbInv[N,3,3], half[N,3], mu[N], kin[N], scene[N]):

* ``bead`` -- the reference's own model: every body is the same voxelised box of bead particles
  of radius R on a 2R lattice (contact_engine_gpu.py:186-188 ``_voxbox``), mass and inertia taken
  from the beads (contact_engine_gpu.py:209-215).
* ``box`` -- per-body box extents with solid-box inertia ``m/3 * diag(hy^2+hz^2, ...)`` in terms
  of half extents (the reference's ncp/ncp_ref.py:1025-1055 ``Box.I_body``). The bead model cannot
  represent boxes with edges below 3R (a 0.1 m cube is one bead at the centre and has zero
  inertia), so the box workloads need this model.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

G = 9.81
R_BEAD = 0.05


def voxbox(w: float, h: float, d: float, res: float = R_BEAD * 2) -> np.ndarray:
    """Bead rest positions, z-major then y then x (contact_engine_gpu.py:186-188)."""
    a = lambda L: np.arange(-L / 2 + res / 2, L / 2, res)
    return np.array([[x, y, z] for z in a(d) for y in a(h) for x in a(w)], float)


@dataclass
class BodySet:
    """Plain host arrays of a body set, ready for upload."""
    x: np.ndarray          # (N,3) float32 centres
    q: np.ndarray          # (N,4) float32 xyzw
    v: np.ndarray          # (N,3) float32
    w: np.ndarray          # (N,3) float32
    invM: np.ndarray       # (N,) float32
    IbInv: np.ndarray      # (N,3,3) float32 body-frame inverse inertia
    kin: np.ndarray        # (N,) int32 1 = kinematic
    half: np.ndarray       # (N,3) float32 box half extents (zeros in bead mode)
    mu: np.ndarray         # (N,) float32 friction coefficient
    scene: np.ndarray      # (N,) int32 scene id (bodies of different scenes never touch)
    M: np.ndarray          # (N,) float64 mass (inf for kinematic), used for contact forces
    rest: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))  # bead mode

    @property
    def n(self) -> int:
        return int(self.x.shape[0])


def bead_bodies(centers, density, kin, dims=(0.3, 0.3, 0.2), mu=0.5) -> BodySet:
    """The reference build (contact_engine_gpu.py:209-239): bead masses mp = rho*vol/P, shape
    inertia summed over the beads of the first body, IbInv = inv((mp_i/mp_0)*Ib) for dynamic
    bodies and the identity for kinematic ones, invM = 0 for kinematic bodies."""
    rest = voxbox(*dims)
    P = len(rest)
    N = len(centers)
    w, h, d = dims
    vol = w * h * d
    mp = np.array(density, float) * vol / P
    M = mp * P
    kinarr = np.array(kin, int)
    invM = np.where(kinarr == 1, 0.0, 1.0 / M)
    Ib = sum(mp[0] * ((r @ r) * np.eye(3) - np.outer(r, r)) for r in rest)
    IbInv = np.array([np.linalg.inv((mp[i] / mp[0]) * Ib) if kinarr[i] == 0 else np.eye(3)
                      for i in range(N)])
    if not np.all(np.isfinite(IbInv)):
        raise ValueError("bead_bodies: IbInv contains non-finite entries -- degenerate body "
                         "mass/inertia; refusing to build a corrupted rigid-body simulation")
    Mh = M.copy()
    Mh[kinarr == 1] = np.inf
    return BodySet(
        x=np.array(np.array(centers, float), np.float32),
        q=np.array(np.tile([0, 0, 0, 1.0], (N, 1)), np.float32),
        v=np.zeros((N, 3), np.float32), w=np.zeros((N, 3), np.float32),
        invM=np.array(invM, np.float32), IbInv=np.array(IbInv, np.float32),
        kin=kinarr.astype(np.int32), half=np.zeros((N, 3), np.float32),
        mu=np.full(N, mu, np.float32), scene=np.zeros(N, np.int32), M=Mh,
        rest=np.array(rest, np.float32))


def box_bodies(centers, half, density, kin=None, quats=None, mu=0.5, scene=None,
               v=None, w=None) -> BodySet:
    """Solid boxes: M = rho * 8*hx*hy*hz, Ib = M/3 diag(hy^2+hz^2, hx^2+hz^2, hx^2+hy^2)
    (ncp/ncp_ref.py:1032-1037), IbInv its exact diagonal inverse; float64 on the host, rounded
    once to float32. Kinematic bodies get invM = 0 and IbInv = I (as contact_engine_gpu.py:213-215)."""
    c = np.asarray(centers, float).reshape(-1, 3)
    N = c.shape[0]
    h = np.broadcast_to(np.asarray(half, float), (N, 3)).copy()
    if np.any(~np.isfinite(h)) or np.any(h <= 0):
        raise ValueError("box_bodies: half extents must be finite and positive")
    rho = np.broadcast_to(np.asarray(density, float), (N,)).copy()
    kinarr = np.zeros(N, int) if kin is None else np.broadcast_to(np.asarray(kin, int), (N,)).copy()
    M = rho * 8.0 * h[:, 0] * h[:, 1] * h[:, 2]
    if np.any(~np.isfinite(M)) or np.any((M <= 0) & (kinarr == 0)):
        raise ValueError("box_bodies: non-finite or non-positive mass")
    invM = np.where(kinarr == 1, 0.0, 1.0 / M)
    hx2, hy2, hz2 = h[:, 0] ** 2, h[:, 1] ** 2, h[:, 2] ** 2
    Ib = np.stack([M / 3.0 * (hy2 + hz2), M / 3.0 * (hx2 + hz2), M / 3.0 * (hx2 + hy2)], axis=1)
    IbInv = np.zeros((N, 3, 3))
    for k in range(3):
        IbInv[:, k, k] = 1.0 / Ib[:, k]
    IbInv[kinarr == 1] = np.eye(3)
    if not np.all(np.isfinite(IbInv)):
        raise ValueError("box_bodies: IbInv contains non-finite entries")
    if quats is None:
        q = np.tile([0, 0, 0, 1.0], (N, 1))
    else:
        q = np.asarray(quats, float).reshape(N, 4)
    mu_arr = np.broadcast_to(np.asarray(mu, float), (N,)).copy()
    sc = np.zeros(N, int) if scene is None else np.asarray(scene, int).reshape(N)
    Mh = M.copy()
    Mh[kinarr == 1] = np.inf
    return BodySet(
        x=np.array(c, np.float32), q=np.array(q, np.float32),
        v=np.zeros((N, 3), np.float32) if v is None else np.array(v, np.float32).reshape(N, 3),
        w=np.zeros((N, 3), np.float32) if w is None else np.array(w, np.float32).reshape(N, 3),
        invM=np.array(invM, np.float32), IbInv=np.array(IbInv, np.float32),
        kin=kinarr.astype(np.int32), half=np.array(h, np.float32),
        mu=np.array(mu_arr, np.float32), scene=sc.astype(np.int32), M=Mh)


def quat_to_R_host(q: np.ndarray) -> np.ndarray:
    """get_state's host rotation from float32 quaternions, in float32 arithmetic
    (contact_engine_gpu.py:283-287 ``_quat2R``)."""
    q = np.asarray(q, np.float32)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    one, two = np.float32(1), np.float32(2)
    R = np.empty((q.shape[0], 3, 3), np.float32)
    R[:, 0, 0] = one - two * (y * y + z * z)
    R[:, 0, 1] = two * (x * y - z * w)
    R[:, 0, 2] = two * (x * z + y * w)
    R[:, 1, 0] = two * (x * y + z * w)
    R[:, 1, 1] = one - two * (x * x + z * z)
    R[:, 1, 2] = two * (y * z - x * w)
    R[:, 2, 0] = two * (x * z - y * w)
    R[:, 2, 1] = two * (y * z + x * w)
    R[:, 2, 2] = one - two * (x * x + y * y)
    return R
