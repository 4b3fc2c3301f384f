"""Host-side worlds of the exact-cone PGS solver: rigid boxes on a ground plane, stepped through
`pgs.NCPSolver` on the GPU; the two-layer N-body lattice; the test scenes.

The reference keeps these in numpy too (src/motion_engine/ncp/ncp_gpu.py:1083-1860); this module
is its own design -- the bodies are arrays (one row per box), and the contact generation works on
every body pair at once: batched separating-axis tests over the 6 face and 9 edge axes, a batched
Sutherland-Hodgman clip of the incident face against the reference face's side slabs, all array
operations over the pairs. What callers observe is kept from the reference, because its recorded
scene numbers (data/ncp/gpu_scene_reference.json) depend on it:

- the tolerances: contact margin 2e-3 m; a face axis of the second box replaces the first box's
  best face only when it separates by 1e-4 more; an edge axis wins over the best face only when it
  separates by 1e-3 more; cross products shorter than 1e-6 are no axis;
- the numbering: corner k = 4 (x > 0) + 2 (y > 0) + (z > 0); face (axis, side) corners in the
  cyclic order (+u+v, -u+v, -u-v, +u-v) with u, v the next two axes; incident face = the first of
  (axis, +) (axis, -) most anti-parallel to the reference normal; point ids (2 axis + side) 8 + n for
  face manifolds (at most the 4 smallest gaps, stably), 100 + 3 i + j for an edge pair, the corner
  index for the ground; feature id a 4096 + (b + 1) 256 + point id;
- the step order: contacts from the poses at the start of the substep, sleeping, the free velocity
  (gravity and external force through the inverse mass), the solve, then positions from the biased
  pass's velocity (soft step) or the solved one, orientation q += 0.5 (w, 0) q dt, normalised.

Products that feed those numbers are formed with batched matrix products (np.matmul), whose
rounding equals the per-body products of the reference; contact ordering is ground contacts by body
then corner, then box pairs (i < j) in index order with their points in clip order.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .protocol import BodyState

G = 9.81
BETA = 0.2        # Baumgarte coefficient (ncp_gpu.py:99)
SLOP = 1e-4       # penetration allowance (ncp_gpu.py:100)
MARGIN = 2e-3     # speculative contact margin of the box world
FACE_B_BIAS = 1e-4
EDGE_BIAS = 1e-3

# corner k: x sign slowest, z sign fastest
CORNER_SIGNS = np.array([[(k >> 2) & 1, (k >> 1) & 1, k & 1] for k in range(8)], np.float64) * 2.0 - 1.0


def _pgs():
    from . import pgs
    return pgs


# ── batched small linear algebra (np.matmul: the same rounding as a per-row product) ──

def _mv(M, v):
    """M [n,3,3] times v [n,3]."""
    return np.matmul(M, v[..., None])[..., 0]


def _dot(a, b):
    """Row-wise dot product of [n,3] arrays."""
    return np.matmul(a[..., None, :], b[..., :, None])[..., 0, 0]


def _norm(a):
    return np.sqrt(_dot(a, a))


def rotation(q):
    """Rotation matrices [n,3,3] of unit quaternions [n,4] (x, y, z, w)."""
    q = np.asarray(q, np.float64).reshape(-1, 4)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((len(q), 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def quat_step(q, w, dt):
    """q + 0.5 (w, 0) q dt, normalised, for rows of q [n,4] and angular velocities w [n,3]."""
    wx, wy, wz = w[:, 0], w[:, 1], w[:, 2]
    qx, qy, qz, qw = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    zero = 0.0 * wx
    prod = np.stack([zero * qx + wx * qw + wy * qz - wz * qy,
                     zero * qy - wx * qz + wy * qw + wz * qx,
                     zero * qz + wx * qy - wy * qx + wz * qw,
                     zero * qw - wx * qx - wy * qy - wz * qz], axis=1)
    qq = q + 0.5 * prod * dt
    return qq / _norm(qq)[:, None]


def corners(x, R, half):
    """World corners [n,8,3] of boxes with centres x, rotations R and half extents half."""
    local = CORNER_SIGNS[None, :, :] * half[:, None, :]
    return x[:, None, :] + np.matmul(local, np.transpose(R, (0, 2, 1)))


def tangents(n):
    """Contact frame tangents of unit normals n [m,3]: t1 from the x axis (y when |n_x| >= 0.9)."""
    a = np.zeros_like(n)
    near_x = np.abs(n[:, 0]) >= 0.9
    a[~near_x, 0] = 1.0
    a[near_x, 1] = 1.0
    t1 = a - _dot(a, n)[:, None] * n
    t1 = t1 / _norm(t1)[:, None]
    return t1, np.cross(n, t1)


# ── contacts ──

@dataclass
class Contacts:
    """Contact set, one row per point: bodies a, b (-1 = ground), point p, normal n from b to a,
    gap (negative = penetration) and the point id within the pair."""
    a: np.ndarray
    b: np.ndarray
    p: np.ndarray
    n: np.ndarray
    gap: np.ndarray
    pid: np.ndarray

    def __len__(self):
        return len(self.a)

    @staticmethod
    def empty():
        return Contacts(np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros((0, 3)), np.zeros((0, 3)),
                        np.zeros(0), np.zeros(0, np.int64))

    def take(self, idx):
        return Contacts(self.a[idx], self.b[idx], self.p[idx], self.n[idx], self.gap[idx], self.pid[idx])

    @staticmethod
    def cat(parts):
        parts = [c for c in parts if len(c)]
        if not parts:
            return Contacts.empty()
        return Contacts(*(np.concatenate([getattr(c, f) for c in parts]) for f in ("a", "b", "p", "n", "gap", "pid")))

    def features(self):
        return (self.a * 4096 + (self.b + 1) * 256 + self.pid).astype(np.int32)


def ground_contacts(x, R, half, margin=MARGIN, active=None):
    """Every box corner below z = margin touches the ground plane (normal +z)."""
    cs = corners(x, R, half)
    hit = cs[:, :, 2] < margin
    if active is not None:
        hit &= active[:, None]
    body, k = np.nonzero(hit)
    p = cs[body, k]
    return Contacts(body.astype(np.int64), np.full(len(body), -1, np.int64), p.copy(),
                    np.tile([0.0, 0.0, 1.0], (len(body), 1)), p[:, 2].copy(), k.astype(np.int64))


def _separation(axes, d, RaT, ha, RbT, hb):
    """Gap of two boxes along unit axes [m,3] and the axis oriented from the first box to the second."""
    ra = np.sum(np.abs(_mv(RaT, axes)) * ha, axis=1)
    rb = np.sum(np.abs(_mv(RbT, axes)) * hb, axis=1)
    proj = _dot(d, axes)
    oriented = np.where((proj >= 0)[:, None], axes, -axes)
    return np.abs(proj) - (ra + rb), oriented


def _face_quad(half, x, R, axis, sgn):
    """Corners [m,4,3] (cyclic: +u+v, -u+v, -u-v, +u-v) of the face (axis, sgn) of each box, its outward
    normal and centre."""
    m = len(x)
    rows = np.arange(m)
    j, k = (axis + 1) % 3, (axis + 2) % 3
    nrm = sgn[:, None] * R[rows, :, axis]
    c = x + half[rows, axis][:, None] * nrm
    u = half[rows, j][:, None] * R[rows, :, j]
    v = half[rows, k][:, None] * R[rows, :, k]
    quad = np.stack([c + u + v, c - u + v, c - u - v, c + u - v], axis=1)
    return quad, nrm, c


def _clip_side(poly, cnt, c, ax, ha, s):
    """One Sutherland-Hodgman pass of every polygon (rows of poly [m,8,3], cnt vertices each) against
    s (p - c) . ax <= ha. A vertex is kept when inside; an edge that crosses adds its crossing point
    right after its start vertex."""
    m, cap = poly.shape[0], poly.shape[1]
    idx = np.arange(cap)[None, :]
    valid = idx < cnt[:, None]
    nxt = np.where(idx + 1 < cnt[:, None], idx + 1, 0)
    q = np.take_along_axis(poly, nxt[:, :, None], axis=1)
    cb = np.broadcast_to(c[:, None, :], poly.shape)
    ab = np.broadcast_to(ax[:, None, :], poly.shape)
    dp = s * _dot(poly - cb, ab) - ha[:, None]
    dq = s * _dot(q - cb, ab) - ha[:, None]
    keep = valid & (dp <= 0)
    cross = valid & (dp * dq < 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        hitp = poly + (q - poly) * (dp / (dp - dq))[:, :, None]
    # output slot of each emitted point: vertex, then its crossing
    n_emit = keep.astype(np.int64) + cross.astype(np.int64)
    start = np.cumsum(n_emit, axis=1) - n_emit
    out = np.zeros_like(poly)
    rows = np.broadcast_to(np.arange(m)[:, None], keep.shape)
    kr, kc = np.nonzero(keep)
    out[rows[kr, kc], start[kr, kc]] = poly[kr, kc]
    xr, xc = np.nonzero(cross)
    out[rows[xr, xc], start[xr, xc] + keep[xr, xc]] = hitp[xr, xc]
    return out, n_emit.sum(axis=1)


def pair_contacts(ia, ib, x, R, half, margin=MARGIN):
    """Contacts of the box pairs (ia[k], ib[k]), the normal pointing from ib to ia."""
    P = len(ia)
    if P == 0:
        return Contacts.empty()
    xa, xb, Ra, Rb, ha, hb = x[ia], x[ib], R[ia], R[ib], half[ia], half[ib]
    RaT, RbT = np.transpose(Ra, (0, 2, 1)), np.transpose(Rb, (0, 2, 1))
    d = xb - xa
    apart = np.zeros(P, bool)
    # best face axis: A's three, then B's three; B needs FACE_B_BIAS more to replace A's best
    f_sep = np.full(P, -np.inf)
    f_owner = np.zeros(P, np.int64)
    f_idx = np.zeros(P, np.int64)
    f_ax = np.zeros((P, 3))
    for owner, Rm in ((0, Ra), (1, Rb)):
        for i in range(3):
            sp, ax = _separation(Rm[:, :, i], d, RaT, ha, RbT, hb)
            apart |= sp > margin
            if owner == 0 and i == 0:
                take = np.ones(P, bool)
            else:
                need = np.where((owner == 1) & (f_owner == 0), FACE_B_BIAS, 1e-12)
                take = sp > f_sep + need
            f_sep = np.where(take, sp, f_sep)
            f_owner = np.where(take, owner, f_owner)
            f_idx = np.where(take, i, f_idx)
            f_ax = np.where(take[:, None], ax, f_ax)
    # best edge-edge axis
    e_sep = np.full(P, -np.inf)
    e_set = np.zeros(P, bool)
    e_idx = np.zeros(P, np.int64)
    e_ax = np.zeros((P, 3))
    for i in range(3):
        for j in range(3):
            cr = np.cross(Ra[:, :, i], Rb[:, :, j])
            nn = _norm(cr)
            ok = nn > 1e-6
            with np.errstate(divide="ignore", invalid="ignore"):
                unit = cr / np.where(ok, nn, 1.0)[:, None]
            sp, ax = _separation(unit, d, RaT, ha, RbT, hb)
            apart |= ok & (sp > margin)
            take = ok & (~e_set | (sp > e_sep + 1e-12))
            e_sep = np.where(take, sp, e_sep)
            e_idx = np.where(take, 3 * i + j, e_idx)
            e_ax = np.where(take[:, None], ax, e_ax)
            e_set |= ok
    touching = ~apart
    edge = touching & e_set & (e_sep > f_sep + EDGE_BIAS)
    face = touching & ~edge
    parts = []
    order_keys = []
    # edge-edge pairs: one point midway between the two boxes' extreme corners along the axis
    if edge.any():
        k = np.flatnonzero(edge)
        nAB = e_ax[k]
        ca = corners(xa[k], Ra[k], ha[k])
        cb = corners(xb[k], Rb[k], hb[k])
        sa = np.matmul(ca, nAB[:, :, None])[..., 0]
        sb = np.matmul(cb, nAB[:, :, None])[..., 0]
        pa = ca[np.arange(len(k)), np.argmax(sa, axis=1)]
        pb = cb[np.arange(len(k)), np.argmin(sb, axis=1)]
        parts.append(Contacts(ia[k].astype(np.int64), ib[k].astype(np.int64), 0.5 * (pa + pb), -nAB, e_sep[k],
                              100 + e_idx[k]))
        order_keys.append(np.stack([k, np.zeros_like(k)], axis=1))
    if face.any():
        k = np.flatnonzero(face)
        m = len(k)
        rows = np.arange(m)
        ownA = f_owner[k] == 0
        idx = f_idx[k]
        nAB = f_ax[k]
        rR = np.where(ownA[:, None, None], Ra[k], Rb[k])
        iR = np.where(ownA[:, None, None], Rb[k], Ra[k])
        rh = np.where(ownA[:, None], ha[k], hb[k])
        ih = np.where(ownA[:, None], hb[k], ha[k])
        rx = np.where(ownA[:, None], xa[k], xb[k])
        ix = np.where(ownA[:, None], xb[k], xa[k])
        toward = np.where(ownA[:, None], nAB, -nAB)   # the reference face faces the other box
        rsgn = np.where(_dot(rR[rows, :, idx], toward) >= 0, 1.0, -1.0)
        nBA = np.where(ownA[:, None], -nAB, nAB)
        _, rn, rc = _face_quad(rh, rx, rR, idx, rsgn)
        # incident face: the first of (axis 0 +, 0 -, 1 +, ...) most anti-parallel to the reference normal
        cand = np.stack([_dot(iR[:, :, a] * sg, rn) for a in range(3) for sg in (1.0, -1.0)], axis=1)
        mi = np.argmin(cand, axis=1)
        quad, _, _ = _face_quad(ih, ix, iR, mi // 2, np.where(mi % 2 == 0, 1.0, -1.0))
        poly = np.zeros((m, 8, 3))
        poly[:, :4] = quad
        cnt = np.full(m, 4, np.int64)
        alive = np.ones(m, bool)
        for off in (1, 2):
            sa = (idx + off) % 3
            ax = rR[rows, :, sa]
            hh = rh[rows, sa]
            for s in (1.0, -1.0):
                poly, cnt = _clip_side(poly, cnt, rc, ax, hh, s)
                alive &= cnt > 0
        g = _dot(poly - rc[:, None, :], np.broadcast_to(rn[:, None, :], poly.shape))
        slot = np.arange(8)[None, :]
        inside = alive[:, None] & (slot < cnt[:, None]) & (g < margin)
        # at most 4 per pair: the smallest gaps, ties in clip order
        rank_key = np.where(inside, g, np.inf)
        order = np.argsort(rank_key, axis=1, kind="stable")
        n_in = inside.sum(axis=1)
        many = n_in > 4
        chosen = inside.copy()
        if many.any():
            keep4 = np.zeros_like(inside)
            np.put_along_axis(keep4, order[:, :4], True, axis=1)
            chosen[many] = keep4[many]
        # output order: clip order, or gap order when more than 4 were found
        out_pos = np.where(many[:, None], np.argsort(order, axis=1), slot)
        r, s_ = np.nonzero(chosen)
        pos = out_pos[r, s_]
        seq = np.lexsort((pos, r))
        r, s_, pos = r[seq], s_[seq], pos[seq]
        # point index within the pair: running count
        first = np.concatenate([[True], r[1:] != r[:-1]]) if len(r) else np.zeros(0, bool)
        start = np.maximum.accumulate(np.where(first, np.arange(len(r)), 0)) if len(r) else np.zeros(0, np.int64)
        nth = np.arange(len(r)) - start
        base = (2 * idx[r] + (rsgn[r] < 0)) * 8
        parts.append(Contacts(ia[k][r].astype(np.int64), ib[k][r].astype(np.int64), poly[r, s_].copy(),
                              nBA[r].copy(), g[r, s_], base + nth))
        order_keys.append(np.stack([k[r], 1 + nth], axis=1))
    if not parts:
        return Contacts.empty()
    allc = Contacts.cat(parts)
    keys = np.concatenate(order_keys)
    seq = np.lexsort((keys[:, 1], keys[:, 0]))
    return allc.take(seq)


def all_pairs(n, active=None):
    """Every pair i < j (of active bodies)."""
    i, j = np.triu_indices(n, k=1)
    if active is not None:
        m = active[i] & active[j]
        i, j = i[m], j[m]
    return i, j


# ── bodies ──

class Box:
    """One row of a world's body arrays, read and written in place (x, q, v, w, half, mass,
    kinematic; Ib the body-frame inertia)."""
    __slots__ = ("_w", "_i")

    def __init__(self, world, i):
        self._w, self._i = world, i

    def _get(name):
        return property(lambda s: getattr(s._w, name)[s._i],
                        lambda s, val: getattr(s._w, name).__setitem__(s._i, val))

    x = _get("x")
    q = _get("q")
    v = _get("v")
    w = _get("w")
    half = _get("half")
    mass = property(lambda s: float(s._w.mass[s._i]))
    kinematic = _get("kin")

    @property
    def Ib(self):
        return self._w.inertia_body()[self._i]


class _BoxArrays:
    """Rows of free boxes: centre, orientation (x, y, z, w), velocities, half extents, mass, kinematic."""

    def _init_arrays(self):
        self.x = np.zeros((0, 3))
        self.q = np.zeros((0, 4))
        self.v = np.zeros((0, 3))
        self.w = np.zeros((0, 3))
        self.half = np.zeros((0, 3))
        self.mass = np.zeros(0)
        self.kin = np.zeros(0, bool)

    def add(self, half, mass, x, v=None):
        row = lambda a, val: np.concatenate([a, np.asarray(val, np.float64).reshape(1, -1)])
        self.x = row(self.x, x)
        self.q = row(self.q, [0.0, 0.0, 0.0, 1.0])
        self.v = row(self.v, np.zeros(3) if v is None else v)
        self.w = row(self.w, np.zeros(3))
        self.half = row(self.half, half)
        self.mass = np.append(self.mass, float(mass))
        self.kin = np.append(self.kin, False)
        return len(self.mass) - 1

    @property
    def n_bodies(self):
        return len(self.mass)

    @property
    def bodies(self):
        return [Box(self, i) for i in range(self.n_bodies)]

    def inertia_body(self):
        e = (2.0 * self.half) ** 2
        diag = np.stack([e[:, 1] + e[:, 2], e[:, 0] + e[:, 2], e[:, 0] + e[:, 1]], axis=1)
        Ib = np.zeros((self.n_bodies, 3, 3))
        Ib[:, [0, 1, 2], [0, 1, 2]] = (self.mass / 12.0)[:, None] * diag
        return Ib

    def inverse_mass(self):
        """World-frame 6x6 inverse mass blocks: 1/m on the linear diagonal, R Ib^-1 R^T; zero when
        kinematic."""
        n = self.n_bodies
        Mi = np.zeros((n, 6, 6))
        dyn = ~self.kin
        if dyn.any():
            R = rotation(self.q[dyn])
            Ii = np.linalg.inv(self.inertia_body()[dyn])
            Mi[dyn, 3:, 3:] = np.matmul(np.matmul(R, Ii), np.transpose(R, (0, 2, 1)))
            lin = np.zeros((int(dyn.sum()), 3, 3))
            lin[:, [0, 1, 2], [0, 1, 2]] = (1.0 / self.mass[dyn])[:, None]
            Mi[dyn, :3, :3] = lin
        return Mi


def contact_blocks(cs: Contacts, x):
    """3x6 Jacobian blocks of contacts: rows (n, t1, t2), linear part D, angular part r x D for the
    first body; minus both for the second."""
    m = len(cs)
    t1, t2 = tangents(cs.n)
    D = np.stack([cs.n, t1, t2], axis=1)
    Ja, Jb = np.zeros((m, 3, 6)), np.zeros((m, 3, 6))
    ra = cs.p - x[cs.a]
    Ja[:, :, :3] = D
    Ja[:, :, 3:] = np.cross(ra[:, None, :], D)
    hasb = cs.b >= 0
    if hasb.any():
        rb = cs.p[hasb] - x[cs.b[hasb]]
        Jb[hasb, :, :3] = -D[hasb]
        Jb[hasb, :, 3:] = -np.cross(rb[:, None, :], D[hasb])
    return Ja, Jb


def baumgarte_bias(gap, dt):
    """A gap may close within the step; a penetration beyond the slop is pushed out at BETA / dt."""
    return np.where(gap > 0, gap / dt, -BETA * np.maximum(-gap - SLOP, 0.0) / dt)


# ── islands and sleeping ──

@dataclass
class SleepConfig:
    enabled: bool = True
    v_sleep: float = 1e-3
    w_sleep: float = 1e-3
    k_steps: int = 8


def connected_islands(n_b: int, pairs: np.ndarray) -> np.ndarray:
    """Island label per body = the smallest body index of its connected component over the pairs of
    two dynamic bodies (a static endpoint, -1, joins nothing)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    m = (pairs[:, 0] >= 0) & (pairs[:, 1] >= 0)
    a, b = pairs[m, 0], pairs[m, 1]
    g = coo_matrix((np.ones(len(a)), (a, b)), shape=(n_b, n_b))
    _, comp = connected_components(g, directed=False)
    low = np.full(comp.max() + 1 if n_b else 0, np.iinfo(np.int64).max, np.int64)
    np.minimum.at(low, comp, np.arange(n_b))
    return low[comp]


class Sleeper:
    """Bodies count the steps they have stayed below (v_sleep, w_sleep); an island whose bodies all
    reach k_steps sleeps. A body touched by a pair that was not in contact the step before, pushed by
    an external force, or moving above the thresholds starts counting again (ncp_gpu.py:1480-1551)."""

    def __init__(self, n_b: int, cfg: SleepConfig):
        self.cfg = cfg
        self.n_b = int(n_b)
        self.still = np.zeros(n_b, np.int64)
        self.asleep = np.zeros(n_b, bool)
        self.labels = np.arange(n_b, dtype=np.int64)
        self._seen = None      # sorted pair keys of the last changed contact set
        self._last = None      # the pair keys as given last time

    def _quiet(self, v, w):
        return (np.linalg.norm(v, axis=1) <= self.cfg.v_sleep) & (np.linalg.norm(w, axis=1) <= self.cfg.w_sleep)

    def begin(self, pairs, v, w, forced=None) -> np.ndarray:
        pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
        keys = pairs[:, 0] * (self.n_b + 1) + pairs[:, 1] + 1
        changed = self._last is None or not np.array_equal(keys, self._last)
        if changed:
            if self._seen is not None and len(keys):
                new = ~np.isin(keys, self._seen)
                ends = pairs[new].reshape(-1)
                self.still[ends[ends >= 0]] = 0
            self._seen = np.unique(keys)
            self._last = keys.copy()
            self.labels = connected_islands(self.n_b, pairs)
        if forced is not None:
            self.still[np.asarray(forced, bool)] = 0
        self.still[~self._quiet(v, w)] = 0
        least = np.full(self.n_b, np.iinfo(np.int64).max, np.int64)
        np.minimum.at(least, self.labels, self.still)
        self.asleep = least[self.labels] >= self.cfg.k_steps
        return self.asleep

    def end(self, v, w):
        awake = ~self.asleep
        quiet = self._quiet(v, w)
        self.still = np.where(awake & quiet, self.still + 1, np.where(awake, 0, self.still))

    def n_islands_asleep(self) -> int:
        return int(np.unique(self.labels[self.asleep]).size)


def contact_labels(lam, mu, tol_lam=1e-10, tol_rel=1e-6) -> np.ndarray:
    """Active set of solved impulses (ncp_gpu.py:1425-1437): 0 open (normal impulse below the scaled
    tolerance), 2 slip (tangential impulse on the cone), 1 stick otherwise."""
    L = np.asarray(lam, np.float64).reshape(-1, 3)
    mu = np.asarray(mu, np.float64).reshape(-1)
    scale = max(float(np.abs(L).max()) if L.size else 0.0, 1e-30)
    on_cone = np.linalg.norm(L[:, 1:], axis=1) >= mu * L[:, 0] * (1.0 - tol_rel) - tol_lam
    opened = L[:, 0] <= tol_lam + tol_rel * scale
    return np.where(opened, 0, np.where(on_cone, 2, 1)).astype(np.int8)


# ── the box world ──

class BoxWorld(_BoxArrays):
    """Free boxes on a static ground plane z = 0, stepped by the exact-cone PGS. `substeps` splits a
    step into S substeps of dt / S with `iters` iterations each; `soft` runs the TGS soft step (warm
    started; biased pass, the position update with its velocity, relax pass); `sleep` puts quiet
    islands to sleep."""

    def __init__(self, dt=1.0 / 240.0, mu=0.5, gravity=True, dtype="float32", solver_kw=None, iters=200,
                 margin=MARGIN, desaxce=True, sleep: Optional[SleepConfig] = None, substeps=1, warm_start=None,
                 soft=None):
        self._init_arrays()
        self.dt, self.mu, self.gravity, self.iters = float(dt), float(mu), bool(gravity), int(iters)
        self.margin, self.dtype, self.desaxce = float(margin), dtype, desaxce
        self.substeps = int(substeps)
        self.soft = soft
        self.warm_start = (soft is not None) if warm_start is None else bool(warm_start)
        self.sleep_cfg = sleep
        self.sleeper: Optional[Sleeper] = None
        self.solver_kw = solver_kw or {}
        self._solver = None
        self._warm_keys = np.zeros(0, np.int64)
        self._warm_lam = np.zeros((0, 3))
        self.last = None
        self.n_contacts = self.n_colors = self.n_asleep = 0

    def _make_solver(self, n_bodies, capacity):
        return _pgs().NCPSolver(n_bodies, capacity, dtype=self.dtype, desaxce=self.desaxce, **self.solver_kw)

    def contact_set(self, active=None) -> Contacts:
        R = rotation(self.q)
        i, j = all_pairs(self.n_bodies, active)
        return Contacts.cat([ground_contacts(self.x, R, self.half, self.margin, active),
                             pair_contacts(i, j, self.x, R, self.half, self.margin)])

    def contacts(self):
        """The contact list as tuples (a, b, point, normal, gap, point id)."""
        cs = self.contact_set()
        return [(int(cs.a[c]), int(cs.b[c]), cs.p[c], cs.n[c], float(cs.gap[c]), int(cs.pid[c]))
                for c in range(len(cs))]

    def _free_velocity(self, dt, f_ext, frozen):
        n = self.n_bodies
        f = np.zeros((n, 3))
        if self.gravity:
            f[:, 2] = -G * self.mass
        if f_ext is not None:
            f = f + np.array([np.asarray(f_ext(i), float) for i in range(n)]).reshape(n, 3)
        lin = np.zeros((n, 3, 3))
        lin[:, [0, 1, 2], [0, 1, 2]] = (1.0 / self.mass)[:, None]
        # (dt M^-1) f, the product in the order the recorded scenes were made with
        vf = np.concatenate([self.v + _mv(dt * lin, f), self.w], axis=1)
        vf[frozen] = 0.0
        return vf

    def build_scene(self, f_ext=None, dt=None, contacts=None, asleep=None):
        pgs = _pgs()
        dt = self.dt if dt is None else float(dt)
        frozen = self.kin.copy() if asleep is None else (self.kin | np.asarray(asleep, bool))
        cs = self.contact_set() if contacts is None else contacts
        if isinstance(cs, list):
            cs = Contacts(*(np.array([c[k] for c in cs]).reshape((len(cs),) + s) if cs else np.zeros((0,) + s)
                            for k, s in ((0, ()), (1, ()), (2, (3,)), (3, (3,)), (4, ()), (5, ()))))
        Ja, Jb = contact_blocks(cs, self.x)
        nc = len(cs)
        bias = np.zeros((nc, 3))
        if self.soft is None:
            bias[:, 0] = baumgarte_bias(cs.gap, dt)
        return pgs.BlockScene(Ja=Ja, Jb=Jb, Minv=self.inverse_mass(), v_free=self._free_velocity(dt, f_ext, frozen),
                              mu=np.full(nc, self.mu), body_pairs=np.stack([cs.a, cs.b], axis=1).astype(np.int32),
                              dt=dt, bias=bias, feature=cs.features(), sep=cs.gap.astype(np.float64))

    def step(self, f_ext=None, history_every=0):
        r = None
        S = max(1, self.substeps)
        for _ in range(S):
            r = self._substep(f_ext, self.dt / S, history_every)
        return r

    def _warm_for(self, feature):
        if not (self.warm_start and len(self._warm_keys)):
            return None
        pos = np.clip(np.searchsorted(self._warm_keys, feature), 0, len(self._warm_keys) - 1)
        hit = self._warm_keys[pos] == feature
        return np.where(hit[:, None], self._warm_lam[pos], 0.0)

    def _substep(self, f_ext, dt, history_every=0):
        pgs = _pgs()
        n = self.n_bodies
        cs = self.contact_set()
        asleep = None
        if self.sleep_cfg is not None and self.sleep_cfg.enabled:
            if self.sleeper is None:
                self.sleeper = Sleeper(n, self.sleep_cfg)
            forced = np.zeros(n, bool) if f_ext is None else \
                np.array([np.any(np.asarray(f_ext(i), float) != 0.0) for i in range(n)])
            asleep = self.sleeper.begin(np.stack([cs.a, cs.b], axis=1), self.v, self.w, forced)
            self.v[asleep] = 0.0
            self.w[asleep] = 0.0
            touched = asleep[cs.a] | np.where(cs.b >= 0, asleep[np.maximum(cs.b, 0)], False)
            cs = cs.take(np.flatnonzero(~touched))
            self.n_asleep = int(asleep.sum())
        bs = self.build_scene(f_ext, dt=dt, contacts=cs, asleep=asleep)
        if self._solver is None:
            self._solver = self._make_solver(n, max(4096, len(cs)))
        if bs.n_c:
            warm = self._warm_for(bs.feature)
            if self.soft is None:
                r = self._solver.solve(bs, iters=self.iters, warm=warm, history_every=history_every)
            else:
                biased, relax = self.soft.split(self.iters)
                r = self._solver.solve(bs, iters=biased, warm=warm, history_every=history_every,
                                       soft=self.soft.params(dt), relax_iters=relax)
                self.last_split = (biased, relax)
            if self.warm_start:
                order = np.argsort(bs.feature, kind="stable")
                self._warm_keys = np.asarray(bs.feature, np.int64)[order]
                self._warm_lam = np.asarray(r.lam)[order]
            v6 = r.v
            self.n_colors = r.n_colors
        else:
            v6 = bs.v_free
            self.n_colors = 0
            r = pgs.SolveResult(np.zeros((0, 3)), v6, 0.0, 0, 0, False, v_bias=v6)
        self.n_contacts = bs.n_c
        self.last = r
        # the bodies keep the solved (relaxed) velocity; positions move with the biased pass's
        vp = v6 if r.v_bias is None else r.v_bias
        mv = ~self.kin if asleep is None else ~(self.kin | asleep)
        self.v[mv] = v6[mv, :3]
        self.w[mv] = v6[mv, 3:]
        self.x[mv] = self.x[mv] + vp[mv, :3] * dt
        if mv.any():
            self.q[mv] = quat_step(self.q[mv], vp[mv, 3:], dt)
        if self.sleeper is not None:
            self.sleeper.end(self.v, self.w)
        return r

    def penetration(self):
        cs = self.contact_set()
        return max(0.0, float(np.max(-cs.gap))) if len(cs) else 0.0

    def kinetic_energy(self):
        Ib = self.inertia_body()
        lin = 0.5 * self.mass * np.sum(self.v * self.v, axis=1)
        ang = 0.5 * np.sum(self.w * np.matmul(Ib, self.w[:, :, None])[..., 0], axis=1)
        return float(np.sum(lin + ang))


# ── the two-layer N-body lattice (ncp_gpu.py:1554-1713) ──

class LatticeWorld:
    """Cubes of half size `half` on a square grid of pitch `spacing`: a lower layer at height z0 on the
    ground, and an upper cube dz above each of the first n_up lower ones. The contact set is fixed by
    construction: each lower cube touches the ground and each upper cube its lower one at 4 bottom
    corners (or 1 point under the centre), normals world +z, a stacked pair's gap measured to the
    lower cube's top face."""

    def __init__(self, n=10000, half=0.1, spacing=0.35, z0=0.0999, dz=0.1999, mass=1.0, dt=1.0 / 240.0, mu=0.5,
                 iters=40, dtype="float32", sleep=None, solver_kw=None, points_per_pair=4):
        side = int(np.ceil(np.sqrt(n / 2.0)))
        self.n = int(n)
        self.n_low = min(side * side, self.n)
        self.n_up = self.n - self.n_low
        if self.n_up > self.n_low:
            raise ValueError("lattice: more upper than lower bodies")
        self.half, self.dt, self.mu, self.iters, self.dtype = float(half), float(dt), float(mu), int(iters), dtype
        self.mass = np.full(self.n, float(mass))
        col, row = np.divmod(np.arange(self.n_low), side)
        centre = (side - 1) / 2.0
        self.x = np.zeros((self.n, 3))
        self.x[:self.n_low, 0] = (col - centre) * spacing
        self.x[:self.n_low, 1] = (row - centre) * spacing
        self.x[:self.n_low, 2] = z0
        self.x[self.n_low:, :2] = self.x[:self.n_up, :2]
        self.x[self.n_low:, 2] = z0 + dz
        self.v = np.zeros((self.n, 3))
        self.w = np.zeros((self.n, 3))
        self.q = np.zeros((self.n, 4))
        self.q[:, 3] = 1.0
        h = self.half
        self.Ib = np.full(3, mass / 12.0 * ((2 * h) ** 2 + (2 * h) ** 2))
        if points_per_pair == 4:
            local = np.array([[-h, -h, -h], [h, -h, -h], [-h, h, -h], [h, h, -h]])
        else:
            local = np.array([[0.0, 0.0, -h]])
        K = len(local)
        owners = np.concatenate([np.arange(self.n_low), np.arange(self.n_low, self.n)])
        below = np.concatenate([np.full(self.n_low, -1), np.arange(self.n_up)])
        self.ca = np.repeat(owners, K).astype(np.int64)
        self.cb = np.repeat(below, K).astype(np.int64)
        self.off = np.tile(local, (len(owners), 1))
        self.feat = (np.tile(np.arange(K), len(owners)) + 8 * (self.cb + 1)).astype(np.int32)
        self.pairs = np.stack([self.ca, self.cb], axis=1)
        self.sleep_cfg = sleep
        self.sleeper = Sleeper(self.n, sleep) if (sleep is not None and sleep.enabled) else None
        self.solver = _pgs().NCPSolver(self.n, len(self.ca), dtype=dtype, **(solver_kw or {}))
        self.n_contacts = self.n_colors = self.n_asleep = 0
        self.last = None

    def _gaps_points(self, idx=None):
        sel = slice(None) if idx is None else idx
        ca, cb, off = self.ca[sel], self.cb[sel], self.off[sel]
        R = rotation(self.q)
        p = self.x[ca] + np.einsum("nij,nj->ni", R[ca], off)
        ground = cb < 0
        top = self.x[np.maximum(cb, 0), 2] + R[np.maximum(cb, 0), 2, 2] * self.half
        gap = np.where(ground, p[:, 2], p[:, 2] - top)
        return gap, p

    def build_scene(self, f_ext=None, keep=None, asleep=None):
        pgs = _pgs()
        n, dt = self.n, self.dt
        idx = np.arange(len(self.ca)) if keep is None else np.flatnonzero(keep)
        gap, p = self._gaps_points(idx)
        nc = len(idx)
        R = rotation(self.q)
        Minv = np.zeros((n, 6, 6))
        Minv[:, [0, 1, 2], [0, 1, 2]] = (1.0 / self.mass)[:, None]
        Minv[:, 3:, 3:] = R @ (np.diag(1.0 / self.Ib)[None] @ np.transpose(R, (0, 2, 1)))
        f = np.zeros((n, 3))
        f[:, 2] = -G * self.mass
        if f_ext is not None:
            f = f + f_ext
        vf = np.concatenate([self.v + dt * f / self.mass[:, None], self.w], axis=1)
        if asleep is not None:
            vf[asleep] = 0.0
        a, b = self.ca[idx], self.cb[idx]
        cs = Contacts(a, b, p, np.tile([0.0, 0.0, 1.0], (nc, 1)), gap, np.zeros(nc, np.int64))
        # the lattice's fixed frame: (n, t1, t2) = (z, x, y)
        D = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
        Ja, Jb = np.zeros((nc, 3, 6)), np.zeros((nc, 3, 6))
        Ja[:, :, :3] = D
        Ja[:, :, 3:] = np.cross((p - self.x[a])[:, None, :], D[None, :, :])
        hasb = b >= 0
        Jb[hasb, :, :3] = -D
        Jb[hasb, :, 3:] = -np.cross((p[hasb] - self.x[b[hasb]])[:, None, :], D[None, :, :])
        bias = np.zeros((nc, 3))
        bias[:, 0] = baumgarte_bias(cs.gap, dt)
        return pgs.BlockScene(Ja=Ja, Jb=Jb, Minv=Minv, v_free=vf, mu=np.full(nc, self.mu),
                              body_pairs=np.stack([a, b], axis=1).astype(np.int32), dt=dt, bias=bias,
                              feature=self.feat[idx])

    def step(self, f_ext=None):
        pgs = _pgs()
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
            r = pgs.SolveResult(np.zeros((0, 3)), v6, 0.0, 0, 0, False)
        self.last = r
        mv = np.ones(self.n, bool) if asleep is None else ~asleep
        self.v[mv] = v6[mv, :3]
        self.w[mv] = v6[mv, 3:]
        self.x[mv] = self.x[mv] + self.v[mv] * self.dt
        self.q[mv] = quat_step(self.q[mv], self.w[mv], self.dt)
        if self.sleeper is not None:
            self.sleeper.end(self.v, self.w)
        return r


def lattice_pushed_forces(w: LatticeWorld, frac=0.1, fx=40.0):
    """A horizontal push fx on every (1/frac)-th stacked island, its lower and upper cube alike."""
    f = np.zeros((w.n, 3))
    picked = np.arange(0, w.n_up, int(round(1.0 / frac)))
    f[picked, 0] = fx
    f[w.n_low + picked, 0] = fx
    return f


# ── scenes of the reference's recorded data (ncp_gpu.py:1719-1854; gpu_scene_reference.json) ──

def _world(world_cls, **kw):
    return (world_cls or BoxWorld)(**kw)


def scene_pushed_cube(dtype="float32", steps=500, iters=200, solver_kw=None, world_cls=None):
    """A 0.2 m, 15 kg cube (mu 0.5, dt 4 ms) pushed by 147.15 N, along x then y then x ... every 0.5 s,
    the push ramped up over the first 0.3 s of each half second."""
    dt, F, period, ramp = 0.004, 147.15, 0.5, 0.3
    w = _world(world_cls, dt=dt, mu=0.5, dtype=dtype, iters=iters, solver_kw=solver_kw)
    w.add([0.1, 0.1, 0.1], 15.0, [0.0, 0.0, 0.1])
    traj, pen, res = [], [], []
    for s in range(steps):
        t = s * dt
        phase = int(t // period)
        amp = F * min(1.0, (t - phase * period) / ramp)
        push = np.array([amp, 0.0, 0.0]) if phase % 2 == 0 else np.array([0.0, amp, 0.0])
        r = w.step(f_ext=lambda i, push=push: push)
        traj.append(w.x[0].copy())
        pen.append(w.penetration())
        res.append(r.residual)
    return w, np.array(traj), np.array(pen), np.array(res)


def scene_sliding_box(dtype="float32", steps=200, iters=200, desaxce=True, v0=1.0, mu=0.3, dt=1.0 / 240.0,
                      substeps=1, soft=None, world_cls=None):
    """A 0.2 m, 1 kg box launched at v0 along x on friction mu: its height above rest and its speed."""
    w = _world(world_cls, dt=dt, mu=mu, dtype=dtype, iters=iters, desaxce=desaxce, substeps=substeps, soft=soft)
    w.add([0.1, 0.1, 0.1], 1.0, [0.0, 0.0, 0.1], v=[v0, 0.0, 0.0])
    gap, vx = np.zeros(steps), np.zeros(steps)
    for s in range(steps):
        w.step()
        gap[s] = w.x[0, 2] - 0.1
        vx[s] = w.v[0, 0]
    return w, gap, vx


def sliding_distance(dtype="float32", iters=400, desaxce=True, v0=1.0, mu=0.3, dt=1.0 / 240.0, max_steps=4000,
                     substeps=1, soft=None, world_cls=None):
    """The distance the box slides until its speed first drops to zero or below, the closed form
    v0^2 / (2 mu g), and the number of steps taken."""
    w = _world(world_cls, dt=dt, mu=mu, dtype=dtype, iters=iters, desaxce=desaxce, substeps=substeps, soft=soft)
    w.add([0.1, 0.1, 0.1], 1.0, [0.0, 0.0, 0.1], v=[v0, 0.0, 0.0])
    x0 = w.x[0, 0]
    steps = 0
    while steps < max_steps:
        steps += 1
        w.step()
        if w.v[0, 0] <= 0.0:
            break
    return float(w.x[0, 0] - x0), v0 * v0 / (2.0 * mu * G), steps


def _column(world_cls, masses, steps, **kw):
    w = _world(world_cls, dt=1.0 / 240.0, mu=0.5, **kw)
    edge = 0.2
    for k, m in enumerate(masses):
        w.add([edge / 2] * 3, m, [0.0, 0.0, (k + 0.5) * edge])
    return w


def scene_massratio(ratio, dtype="float32", steps=400, iters=200, substeps=1, soft=None, world_cls=None):
    """Three stacked 0.2 m boxes, the top one `ratio` times heavier: the deepest penetration per step."""
    w = _column(world_cls, [1.0, 1.0, ratio], steps, dtype=dtype, iters=iters, substeps=substeps, soft=soft)
    pen = np.zeros(steps)
    for s in range(steps):
        w.step()
        pen[s] = w.penetration()
    return w, pen


def scene_tower(K=8, dtype="float32", steps=400, iters=200, sleep=None, substeps=1, soft=None, world_cls=None):
    """A tower of K 0.2 m unit boxes: penetration per step; the fastest body's speed and spin and the
    sleeping count per step in w.vmax_history, w.wmax_history, w.asleep_history."""
    w = _column(world_cls, [1.0] * K, steps, dtype=dtype, iters=iters, sleep=sleep, substeps=substeps, soft=soft)
    pen, asleep, vmax, wmax = (np.zeros(steps) for _ in range(4))
    for s in range(steps):
        w.step()
        pen[s] = w.penetration()
        asleep[s] = w.n_asleep
        vmax[s] = float(np.max(np.linalg.norm(w.v, axis=1)))
        wmax[s] = float(np.max(np.linalg.norm(w.w, axis=1)))
    w.asleep_history, w.vmax_history, w.wmax_history = asleep, vmax, wmax
    return w, pen


# ── synthetic block scenes (the reference selftest's workload, ncp_gpu.py:1872-1880) ──

def _free_body_blocks(nrm, ra, rb):
    t1, t2 = tangents(nrm)
    D = np.stack([nrm, t1, t2], axis=1)
    m = len(nrm)
    Ja, Jb = np.zeros((m, 3, 6)), np.zeros((m, 3, 6))
    Ja[:, :, :3] = D
    Ja[:, :, 3:] = np.cross(ra[:, None, :], D)
    Jb[:, :, :3] = -D
    Jb[:, :, 3:] = -np.cross(rb[:, None, :], D)
    return Ja, Jb


def synthetic_scene(n_c, n_b=512, seed=0, mu=0.5):
    """n_c random contacts between n_b unit-mass free bodies, as blocks. Draws, in order: body a
    (n_c), body offset (n_c), normals (n_c, 3); then per contact its two lever arms U(-0.1, 0.1)^3;
    then the free velocities (n_b, 6) ~ N(0, 0.2)."""
    from . import pgs
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n_b, n_c)
    b = (a + 1 + rng.integers(0, n_b - 1, n_c)) % n_b
    nrm = rng.normal(size=(n_c, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    arms = np.array([rng.uniform(-0.1, 0.1, 3) for _ in range(2 * n_c)]).reshape(n_c, 2, 3)
    Ja, Jb = _free_body_blocks(nrm, arms[:, 0], arms[:, 1])
    Iinv = np.linalg.inv(1.0 / 12.0 * np.diag([0.08, 0.08, 0.08]))
    blk = np.eye(6)
    blk[3:, 3:] = Iinv
    vf = rng.normal(scale=0.2, size=(n_b, 6))
    return pgs.BlockScene(Ja=Ja, Jb=Jb, Minv=np.tile(blk, (n_b, 1, 1)), v_free=vf, mu=np.full(n_c, mu),
                          body_pairs=np.stack([a, b], axis=1).astype(np.int32), dt=1.0 / 240.0,
                          bias=np.zeros((n_c, 3)), feature=np.arange(n_c, dtype=np.int32))


def random_wellposed_scene(n_c=8, n_b=8, seed=1, mu=0.4):
    """3 n_c <= 6 n_b random full-rank blocks (SPD Delassus, unique NCP solution). Draws, in order: a
    body permutation, Ja, Jb (n_c, 3, 6) ~ N(0, 1), mass scales U(0.5, 2), free velocities N(0, 0.5),
    normal biases -U(0.1, 1)."""
    from . import pgs
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n_b)
    k = np.arange(n_c)
    a, b = perm[k % n_b], perm[(k + 3) % n_b]
    Ja = rng.normal(size=(n_c, 3, 6))
    Jb = rng.normal(size=(n_c, 3, 6))
    Mi = np.eye(6)[None] * rng.uniform(0.5, 2.0, (n_b, 1, 1))
    vf = rng.normal(scale=0.5, size=(n_b, 6))
    bias = np.zeros((n_c, 3))
    bias[:, 0] = -rng.uniform(0.1, 1.0, n_c)
    return pgs.BlockScene(Ja=Ja, Jb=Jb, Minv=Mi, v_free=vf, mu=np.full(n_c, mu),
                          body_pairs=np.stack([a, b], axis=1).astype(np.int32), dt=1.0 / 240.0, bias=bias,
                          feature=np.arange(n_c, dtype=np.int32))


# ── the fourth ContactEngine member ──

class ExactConeContactEngine:
    """A box world stepped by the exact-cone PGS behind the ContactEngine protocol
    (contact_engine.py:49-55): step(dt, substeps), get_state, set_kinematic, contact_forces.
    Kinematic bodies enter with zero inverse mass (static to the solver) and keep the pose
    set_kinematic gives them; their force is reported as zero."""
    name = "exact_cone_pgs"

    def __init__(self, mu=0.5, iters=200, dtype="float32", desaxce=True, soft=None, margin=MARGIN,
                 solver_kw=None):
        self.world = BoxWorld(mu=mu, iters=iters, dtype=dtype, desaxce=desaxce, soft=soft, margin=margin,
                              solver_kw=solver_kw)
        self._force = np.zeros((0, 3))

    def add_box(self, center, half, mass, v=None):
        return self.world.add(half, mass, center, v=v)

    def step(self, dt: float, substeps: int = 1) -> None:
        w = self.world
        w.dt, w.substeps = float(dt), int(substeps)
        before = w.v.copy()
        w.step()
        g = np.array([0.0, 0.0, -G]) if w.gravity else np.zeros(3)
        # contact force = M (v_after - v_before - g dt) / dt; none on kinematic bodies
        self._force = w.mass[:, None] * (w.v - before - g[None, :] * dt) / dt
        self._force[w.kin] = 0.0

    def get_state(self) -> BodyState:
        w = self.world
        return BodyState(xc=w.x.copy(), Rm=rotation(w.q).reshape(-1, 3, 3), vc=w.v.copy(), om=w.w.copy())

    def set_kinematic(self, i: int, xc) -> None:
        w = self.world
        w.x[i] = np.asarray(xc, float)
        w.v[i] = 0.0
        w.w[i] = 0.0
        w.kin[i] = True

    def contact_forces(self) -> np.ndarray:
        return self._force.copy()
