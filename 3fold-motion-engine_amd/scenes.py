"""Seeded synthetic workloads (numpy default_rng, seed 0 unless given; draw order documented per
function). Every scene is a BodySet of the box model.

* ``box_stack``  -- 16 unit-mass cubes, edge 0.1 m, on the ground with gap 1e-4 m, xy jitter.
* ``box_pile``   -- N boxes, per-axis edges U(0.05, 0.15) m, density 1000, 5 layers of
                    cells on a 0.2 m grid (layer cells = ceil(N/5), grid width 100 or
                    ceil(sqrt(cells)) when fewer), jitter U(+-0.02)^3, yaw U(0, 2pi), mu U(0.3, 0.8);
                    optional log-uniform mass ratio (density scale).
* ``batched_piles`` -- independent piles with N_s = round(exp(U(ln lo, ln hi))) bodies, one scene
                    id each (each with its own ground).
* ``lattice`` / ``stack`` -- the reference's bead fixtures (scripts/colored_vs_jacobi.py:100-112).
"""
from __future__ import annotations

import math

import numpy as np

from .bodies import BodySet, box_bodies

PITCH = 0.205


def stack(K):
    """Reference bead stack: K boxes at 0.101 + k*0.205 (scripts/colored_vs_jacobi.py:100-101)."""
    return [[0, 0, 0.101 + k * PITCH] for k in range(K)]


def lattice(n_target, layers=2):
    """Reference bead lattice 0.35 m apart (scripts/colored_vs_jacobi.py:104-112)."""
    side = int(np.ceil(np.sqrt(n_target / layers)))
    out = []
    for lz in range(layers):
        for iy in range(side):
            for ix in range(side):
                out.append([ix * 0.35, iy * 0.35, 0.101 + lz * PITCH])
    return out[:n_target]


def box_stack(n=16, edge=0.1, gap=1e-4, jitter=1e-3, mu=0.5, seed=0) -> BodySet:
    """Draws: jitter (n, 2) U(-jitter, jitter). Centres z_k = edge/2 + gap + k*(edge + gap)."""
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-jitter, jitter, size=(n, 2))
    c = np.zeros((n, 3))
    c[:, :2] = jit
    c[:, 2] = edge / 2 + gap + np.arange(n) * (edge + gap)
    rho = 1.0 / edge ** 3
    return box_bodies(c, [edge / 2] * 3, rho, mu=mu)


def _yaw_quats(yaw):
    q = np.zeros((len(yaw), 4))
    q[:, 2] = np.sin(yaw / 2)
    q[:, 3] = np.cos(yaw / 2)
    return q


def pile_layout(n, layers=5, pitch=0.2, width=100, z0=0.1):
    cells = int(math.ceil(n / layers))
    w = min(width, int(math.ceil(math.sqrt(cells))))
    k = np.arange(n)
    cell = k % cells
    layer = k // cells
    return np.stack([(cell % w) * pitch, (cell // w) * pitch, z0 + layer * pitch], axis=1)


def box_pile(n=10000, seed=0, rng=None, layers=5, mass_ratio=None, scene=0, mu_range=(0.3, 0.8),
             origin=(0.0, 0.0, 0.0)) -> BodySet:
    """Draws, in order: edges (n,3) U(0.05,0.15); jitter (n,3) U(-0.02,0.02); yaw (n) U(0,2pi);
    mu (n) U(mu_range); then, when mass_ratio is set, density scale (n) = exp(U(0, ln ratio))."""
    rng = np.random.default_rng(seed) if rng is None else rng
    edges = rng.uniform(0.05, 0.15, size=(n, 3))
    jit = rng.uniform(-0.02, 0.02, size=(n, 3))
    yaw = rng.uniform(0.0, 2 * np.pi, size=n)
    mu = rng.uniform(mu_range[0], mu_range[1], size=n)
    rho = np.full(n, 1000.0)
    if mass_ratio:
        rho = rho * np.exp(rng.uniform(0.0, math.log(mass_ratio), size=n))
    c = pile_layout(n, layers) + jit + np.asarray(origin, float)
    return box_bodies(c, edges / 2, rho, quats=_yaw_quats(yaw), mu=mu, scene=np.full(n, scene))


def batched_sizes(n_scenes=256, lo=8, hi=2048, seed=0):
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), size=n_scenes))).astype(int), rng


def batched_piles(n_scenes=256, lo=8, hi=2048, seed=0):
    """Scene sizes first (n_scenes draws), then each scene's pile draws from the same generator.
    Returns (BodySet with scene ids 0..n_scenes-1, sizes)."""
    sizes, rng = batched_sizes(n_scenes, lo, hi, seed)
    parts = [box_pile(int(ns), rng=rng, scene=s) for s, ns in enumerate(sizes)]
    return concat(parts), sizes


def fixed_piles(n_scenes=1024, n=64, seed=0, vel_sigma=0.1, target_half=0.5):
    """n_scenes independent piles of n boxes (box_pile draws per scene from one generator, scene ids
    0..n_scenes-1), then initial velocities v0, w0 ~ N(0, vel_sigma) (N*n_scenes x 3 each) and targets
    U(-target_half, target_half)^3 per box, in that order. Returns (BodySet, targets)."""
    rng = np.random.default_rng(seed)
    b = concat([box_pile(n, rng=rng, scene=s) for s in range(n_scenes)])
    N = b.n
    b.v[:] = rng.normal(0.0, vel_sigma, size=(N, 3))
    b.w[:] = rng.normal(0.0, vel_sigma, size=(N, 3))
    targets = rng.uniform(-target_half, target_half, size=(N, 3))
    return b, targets


def concat(parts) -> BodySet:
    cat = lambda k: np.concatenate([getattr(b, k) for b in parts])
    return BodySet(**{k: cat(k) for k in ("x", "q", "v", "w", "invM", "IbInv", "kin", "half", "mu",
                                           "scene", "M")})


def ramp_box(deg, mu=0.5, half=(0.15, 0.1, 0.1), slab_half=(2.0, 0.5, 0.1), gap=1e-4, density=1000.0) -> BodySet:
    """A box resting on a kinematic slab tilted by `deg` about y, the physics check of the reference's
    ramp (tests/test_contact_and_grasp.py:20-31: a 0.3 x 0.2 x 0.2 box, mu 0.5, downhill direction
    (-cos deg, 0, -sin deg)). Body 0 is the box, body 1 the slab; both carry mu, so the pair's mu is
    mu. The slab's centre is high enough that it clears the ground plane. No random draws."""
    th = math.radians(deg)
    n = np.array([-math.sin(th), 0.0, math.cos(th)])       # slab top normal
    q = np.array([0.0, math.sin(-th / 2), 0.0, math.cos(-th / 2)])  # rotation about y by -deg
    cs = np.array([0.0, 0.0, slab_half[0] * math.sin(th) + slab_half[2] + 0.2])
    cb = cs + (slab_half[2] + half[2] + gap) * n
    return box_bodies([cb, cs], [half, slab_half], density, kin=[0, 1], quats=[q, q], mu=mu)


def ramp_downhill(deg):
    th = math.radians(deg)
    return np.array([-math.cos(th), 0.0, -math.sin(th)])
