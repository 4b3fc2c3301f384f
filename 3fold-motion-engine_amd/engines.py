"""The three GPU solvers behind the ContactEngine protocol, on libmotion_hip.so.

Class names, constructor keywords, ``add_body`` and the harness-visible attributes follow the
reference classes:
  RelaxedJacobiContactEngine  contact_engine_gpu.py:191-299
  GraphColoredContactEngine   contact_engine_gpu_colored.py:793-1410 (pair-chunk solve)
  ColorCacheContactEngine     contact_engine_gpu_color_cache.py:21-54 (FixedFrame -> AppliedWarm ->
                              DeferredForce -> ColorCache chain)
Body state lives in torch tensors on the GPU (``engine.x``, ``q``, ``v``, ``w``, ...); the library
works on those tensors and on one workspace tensor. ``get_state`` / ``contact_forces`` return host
numpy copies, as the protocol requires.

Two body models: ``add_body(center, density, kin)`` is the reference's bead model (every body the
same voxelised box of dims), ``add_box`` / ``add_boxes`` the per-body box model of the box
workloads. One engine holds one model.
"""
from __future__ import annotations

import ctypes as C
import time
from contextlib import contextmanager

import numpy as np

from . import _lib
from .bodies import BodySet, R_BEAD, bead_bodies, box_bodies, quat_to_R_host
from .protocol import BodyState

_MAXC = 400000


def _torch():
    import torch
    return torch


def bead_feature_keys(keys, n_bodies, n_beads):
    """The reference's feature keys fa*(N*P+1) + fb + 1 (contact_engine_gpu_colored.py:903-917) of
    bead contacts keyed pair-major (pairkey*P*(P+1) + la*(P+1) + lb + 1, pairkey = a*(N+1) + b + 1)."""
    keys = np.asarray(keys, np.int64)
    N, P = int(n_bodies), int(n_beads)
    pk, rem = np.divmod(keys, np.int64(P * (P + 1)))
    la, lb1 = np.divmod(rem, np.int64(P + 1))
    a, b1 = np.divmod(pk, np.int64(N + 1))
    fa = a * P + la
    fb = np.where(b1 == 0, -1, (b1 - 1) * P + (lb1 - 1))
    return fa * np.int64(N * P + 1) + fb + 1


class _HipContactEngine:
    name = "hip_contact"
    _solver = _lib.SOLVER_JACOBI

    def __init__(s, dims=(0.3, 0.3, 0.2), vit=40, pit=20, mu=0.5, relax=0.2, t2_damp=1.0,
                 warm_start=True, margin=0.0, n_colors=0, tiled=False, tile_bodies=128,
                 max_pairs=None, max_contacts=None, device="cuda:0", check_every_step=True,
                 profile=False, per_contact=False, graph_capture=False):
        if vit < 0 or pit < 0:
            raise ValueError("iteration counts must be >= 0")
        if not 0 <= int(n_colors) <= 512:
            raise ValueError("n_colors must be in [0, 512]")
        s.dims = tuple(dims)
        s.vit, s.pit, s.mu, s.relax, s.t2_damp = int(vit), int(pit), float(mu), float(relax), float(t2_damp)
        s.warm_start = bool(warm_start)
        s.margin = float(margin)
        s.n_colors_cfg = int(n_colors)
        s.tiled = bool(tiled)
        s.tile_bodies = int(tile_bodies)
        s.max_pairs, s.max_contacts = max_pairs, max_contacts
        s.device = device
        s.check_every_step = bool(check_every_step)
        s.per_contact = bool(per_contact)
        # whole-step hipGraph replay (the tiled coloured solve; otherwise eager)
        s.graph_capture = bool(graph_capture)
        s.profile = bool(profile)
        s.prof_ms = {}
        s._centers, s._dens, s._kin = [], [], []
        s._boxsets = []
        s._geometry = None
        s._built = False
        s._h = None
        s._sdt = None
        s._counters = dict.fromkeys(_lib.COUNTERS, 0)

    # ── construction (data in) ──
    def add_body(s, center, density=700., kin=False):
        """Reference body model: a voxelised box of ``dims`` (contact_engine_gpu.py:205-207)."""
        if s._geometry == "box":
            raise ValueError("this engine holds box bodies; add_body adds bead bodies")
        s._geometry = "bead"
        s._centers.append([float(c) for c in center])
        s._dens.append(float(density))
        s._kin.append(1 if kin else 0)
        s._built = False
        return len(s._centers) - 1

    def add_box(s, center, half, density=1000., kin=False, quat=None, mu=None, scene=0, v=None, w=None):
        """Box body model: centre, half extents, density (solid-box mass and inertia)."""
        b = box_bodies([center], [half], [density], kin=[1 if kin else 0],
                       quats=None if quat is None else [quat], mu=[s.mu if mu is None else mu],
                       scene=[scene], v=None if v is None else [v], w=None if w is None else [w])
        return s.add_boxes(b)

    def add_boxes(s, bodies: BodySet):
        if s._geometry == "bead":
            raise ValueError("this engine holds bead bodies; add_boxes adds box bodies")
        s._geometry = "box"
        s._boxsets.append(bodies)
        s._built = False
        return sum(b.n for b in s._boxsets) - 1

    def _bodyset(s) -> BodySet:
        if s._geometry == "bead":
            return bead_bodies(s._centers, s._dens, s._kin, dims=s.dims, mu=s.mu)
        if s._geometry == "box":
            if len(s._boxsets) == 1:
                return s._boxsets[0]
            cat = lambda k: np.concatenate([getattr(b, k) for b in s._boxsets])
            return BodySet(**{k: cat(k) for k in ("x", "q", "v", "w", "invM", "IbInv", "kin", "half",
                                                   "mu", "scene", "M")})
        raise ValueError("no bodies added")

    def _config(s, N, P):
        flags = (_lib.FLAG_TILED if s.tiled else 0) | (_lib.FLAG_PER_CONTACT if s.per_contact else 0)
        if s.graph_capture and s.tiled and s._solver != _lib.SOLVER_JACOBI:
            flags |= _lib.FLAG_GRAPH
        if getattr(s, "fast_math", False):
            flags |= _lib.FLAG_FAST_MATH
        return _lib.Config(s._solver, N, P, s.vit, s.pit, s.relax, s.t2_damp, int(s.warm_start),
                           R_BEAD, s.margin,
                           int(s.max_pairs or 0), int(s.max_contacts or (_MAXC if P else 0)),
                           s.n_colors_cfg, s.tile_bodies, flags)

    def _build(s):
        torch = _torch()
        if not torch.cuda.is_available():
            raise _lib.MotionError("no ROCm GPU visible: the contact step runs only on the GPU")
        L = _lib.lib()
        bs = s._bodyset()
        for k in ("x", "q", "v", "w", "invM", "IbInv", "half", "mu"):
            if not np.all(np.isfinite(getattr(bs, k))):
                raise ValueError(f"{s.name}._build: non-finite {k}")
        dev = torch.device(s.device)
        f32 = lambda a, shape: torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).reshape(shape))).to(dev)
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.int32))).to(dev)
        N = bs.n
        P = len(bs.rest)
        s.N, s.P = N, P
        s.x, s.q, s.v, s.w = f32(bs.x, (N, 3)), f32(bs.q, (N, 4)), f32(bs.v, (N, 3)), f32(bs.w, (N, 3))
        s.invM, s.IbInv = f32(bs.invM, (N,)), f32(bs.IbInv, (N, 3, 3))
        s.half, s.mu_t = f32(bs.half, (N, 3)), f32(bs.mu, (N,))
        s.kin, s.scene = i32(bs.kin), i32(bs.scene)
        s.rest = f32(bs.rest if P else np.zeros((1, 3)), (max(P, 1), 3))
        s._M = np.array(bs.M, np.float64)
        s._kinarr = np.array(bs.kin, int)
        s._cfg = s._config(N, P)
        nbytes = C.c_size_t()
        _lib.check(L.me_workspace_size(C.byref(s._cfg), C.byref(nbytes)), "me_workspace_size")
        s._ws = torch.empty(int(nbytes.value) + 256, dtype=torch.uint8, device=dev)
        base = s._ws.data_ptr()
        s._ws_off = (-base) % 256
        s._ws_ptr = base + s._ws_off
        s._bodies = s._bodies_struct()
        h = C.c_void_p()
        _lib.check(L.me_create(C.byref(s._cfg), C.byref(s._bodies), C.c_void_p(s._ws_ptr), nbytes.value,
                               C.c_void_p(s._stream()), C.byref(h)), "me_create")
        s._h = h
        s._sdt = None
        s._built = True

    def _bodies_struct(s):
        p = lambda t: C.c_void_p(t.data_ptr())
        return _lib.Bodies(p(s.x), p(s.q), p(s.v), p(s.w), p(s.invM), p(s.IbInv), p(s.half), p(s.mu_t),
                           p(s.kin), p(s.scene), p(s.rest))

    def _stream(s):
        torch = _torch()
        return torch.cuda.current_stream(torch.device(s.device)).cuda_stream

    def __del__(s):
        try:
            if s._h is not None:
                _lib.lib().me_destroy(s._h)
                s._h = None
        except Exception:
            pass

    def set_kinematic(s, i, xc):
        """Kinematic body: driven position, zero velocity, invM = 0 (contact_engine_gpu.py:241-249)."""
        if s._geometry == "bead" or not s._built:
            if s._geometry == "bead":
                s._kin[i] = 1
                s._centers[i] = [float(c) for c in xc]
        if s._built:
            torch = _torch()
            with torch.no_grad():
                s.x[i] = torch.tensor([float(c) for c in xc], dtype=torch.float32, device=s.x.device)
                s.invM[i] = 0.0
                s.v[i] = 0.0
                s.w[i] = 0.0
                s.kin[i] = 1
            s._kinarr[i] = 1
            s._M[i] = np.inf
        elif s._geometry == "box":
            bs = s._bodyset()
            bs.kin[i] = 1; bs.invM[i] = 0.0; bs.v[i] = 0.0; bs.w[i] = 0.0; bs.M[i] = np.inf
            bs.x[i] = np.asarray(xc, np.float32); bs.IbInv[i] = np.eye(3, dtype=np.float32)
            s._boxsets = [bs]

    # ── advance ──
    @contextmanager
    def _stage(s, name):
        if not s.profile:
            yield
            return
        torch = _torch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        yield
        torch.cuda.synchronize()
        s.prof_ms[name] = s.prof_ms.get(name, 0.0) + (time.perf_counter() - t0) * 1e3

    def step(s, dt, substeps=1):
        if not s._built:
            s._build()
        with s._stage("step"):
            _lib.check(_lib.lib().me_step(s._h, float(dt), int(substeps), C.c_void_p(s._stream())), "me_step")
        s._sdt = float(dt) / int(substeps)
        if s.check_every_step:
            s.refresh_counters()

    def refresh_counters(s, raise_on_overflow=True):
        out = (C.c_int64 * len(_lib.COUNTERS))()
        _lib.check(_lib.lib().me_read_counters(s._h, out, C.c_void_p(s._stream())), "me_read_counters")
        s._counters = dict(zip(_lib.COUNTERS, (int(v) for v in out)))
        ovf = s._counters["overflow"]
        if ovf and raise_on_overflow:
            s.raise_overflow(ovf)
        return s._counters

    def raise_overflow(s, ovf):
        if ovf:
            what = [n for bit, n in ((1, "contacts"), (2, "body pairs"), (4, "broad-phase candidates"),
                                     (8, "colours"), (16, "tile"), (32, "recorded tile for the adjoint"))
                    if ovf & bit]
            if ovf & 256:
                raise RuntimeError(f"{s.name}.step: a one-pass scan waited past its bound for an earlier block; "
                                   f"the step's state is not valid")
            if ovf & 64:
                raise RuntimeError(f"{s.name}.step: the grid sweep of an island beyond one workgroup timed out "
                                   f"waiting for its workgroups; the step's state is not valid")
            if ovf & 128:
                raise RuntimeError(f"{s.name}: an adjoint contribution reached 2^30, beyond the range of the "
                                   f"fixed-point body sums; the gradient is not valid")
            raise RuntimeError(f"{s.name}.step: capacity exceeded ({', '.join(what)}); raise max_pairs / "
                               f"max_contacts")

    # harness-visible counters (contact_engine_gpu_colored.py:824-829)
    n_contacts = property(lambda s: s._counters["n_contacts"])
    n_contacts_solved = property(lambda s: s._counters["n_solved"])
    n_pairs = property(lambda s: s._counters["n_pairs"])
    n_colors = property(lambda s: s._counters["n_colors"])
    color_cache_hits = property(lambda s: s._counters["cache_hits"])
    color_cache_misses = property(lambda s: s._counters["cache_misses"])
    n_graph_captures = property(lambda s: s._counters["graph_captures"])

    # ── data out ──
    def get_state(s):
        if not s._built:
            s._build()
        q = s.q.cpu().numpy()
        return BodyState(xc=s.x.cpu().numpy(), Rm=quat_to_R_host(q), vc=s.v.cpu().numpy(), om=s.w.cpu().numpy())

    def contact_forces(s):
        """Net contact force of the last substep: M (v_end - v_after_gravity) / dt_sub, float32
        difference then float64 (contact_engine_gpu_deferred_force.py:66-74)."""
        if not s._built or s._sdt is None:
            return np.zeros((len(s._M) if s._built else 0, 3))
        before = s.buffer(_lib.BUF_VSNAP, np.float32, (s.N, 3))
        after = s.v.cpu().numpy()
        F = np.zeros((s.N, 3))
        dyn = s._kinarr == 0
        F[dyn] = s._M[dyn, None] * (after[dyn] - before[dyn]) / s._sdt
        return F

    def com(s, i):
        return s.x[i].cpu().numpy()

    def state_numpy(s):
        return s.x.cpu().numpy(), s.q.cpu().numpy(), s.v.cpu().numpy(), s.w.cpu().numpy()

    # ── solver arrays (views into the workspace) ──
    def buffer_tensor(s, bid, dtype, count=None):
        torch = _torch()
        ptr = C.c_void_p()
        n = C.c_int64()
        _lib.check(_lib.lib().me_buffer(s._h, bid, C.byref(ptr), C.byref(n)), "me_buffer")
        tdt = {np.float32: torch.float32, np.int32: torch.int32, np.int64: torch.int64}[dtype]
        item = np.dtype(dtype).itemsize
        off = (ptr.value or 0) - s._ws.data_ptr()
        cnt = int(n.value) if count is None else int(count)
        return s._ws[off:off + cnt * item].view(tdt)

    def buffer(s, bid, dtype, shape):
        cnt = int(np.prod(shape))
        return s.buffer_tensor(bid, dtype, cnt).cpu().numpy().reshape(shape)

    def contacts(s):
        """Contact list of the last substep with its impulses: pair-major, sorted by key (box geometry:
        pair key * 8192 + feature; bead geometry under a coloured solver: the pair-major composite key
        of the bead features, see bead_feature_keys)."""
        Cn = s._counters["n_solved"]
        b = lambda bid, dt, shape: s.buffer(bid, dt, shape)
        return dict(key=b(_lib.BUF_KEY, np.int64, (Cn,)), bi=b(_lib.BUF_BI, np.int32, (Cn,)),
                    bj=b(_lib.BUF_BJ, np.int32, (Cn,)), pA=b(_lib.BUF_PA, np.float32, (Cn, 3)),
                    pB=b(_lib.BUF_PB, np.float32, (Cn, 3)), n=b(_lib.BUF_N, np.float32, (Cn, 3)),
                    pen=b(_lib.BUF_PEN, np.float32, (Cn,)), jn=b(_lib.BUF_JN, np.float32, (Cn,)),
                    jt1=b(_lib.BUF_JT1, np.float32, (Cn,)), jt2=b(_lib.BUF_JT2, np.float32, (Cn,)),
                    jp=b(_lib.BUF_JP, np.float32, (Cn,)), mu=b(_lib.BUF_MU, np.float32, (Cn,)))

    def pairs(s):
        P = s._counters["n_pairs"]
        b = lambda bid: s.buffer(bid, np.int32, (P,))
        return dict(pbi=b(_lib.BUF_PBI), pbj=b(_lib.BUF_PBJ), pstart=b(_lib.BUF_PSTART),
                    pcount=b(_lib.BUF_PCOUNT), color=b(_lib.BUF_COLOR))

    def aux(s):
        return (s.buffer(_lib.BUF_PV, np.float32, (s.N, 3)), s.buffer(_lib.BUF_PO, np.float32, (s.N, 3)),
                s.buffer(_lib.BUF_INVIW, np.float32, (s.N, 3, 3)))


class RelaxedJacobiContactEngine(_HipContactEngine):
    """Relaxed Jacobi with int64 fixed-point accumulation (contact_engine_gpu.py:191-299)."""
    name = "relaxed_jacobi_gpu"
    _solver = _lib.SOLVER_JACOBI

    def __init__(s, dims=(0.3, 0.3, 0.2), relax=0.2, vit=40, pit=20, mu=0.5, **kw):
        super().__init__(dims=dims, vit=vit, pit=pit, mu=mu, relax=relax, **kw)


class GraphColoredContactEngine(_HipContactEngine):
    """Graph-coloured Gauss-Seidel (contact_engine_gpu_colored.py:793-1410), with the reference's
    constructor and defaults (:798-823):

    - ``pair_chunks=False`` (default): Jones-Plassmann colouring of the contact graph, every contact
      its own chunk, contacts in feature-key order (:903-917, 1105-1136), ``vit`` 20;
    - ``pair_chunks=True``: colouring of the body-pair graph, a pair's <= 4 contacts solved as one
      chunk; needs the device manifold reduction (``manifold_reduce=True``,
      ``manifold_on_host=False``), ``vit`` 40.

    Slip-direction friction frame, warm start by feature key without applying it to the
    velocities. ``graph_mode`` and ``chunk_mode`` choose a launch schedule in the reference and
    leave the results unchanged; they are validated as there. Box bodies always carry at most 4
    narrow-phase points per pair, so ``manifold_reduce`` selects no further reduction for them."""
    name = "graph_colored_gpu"
    _solver = _lib.SOLVER_PAIR_CHUNKS

    def __init__(s, dims=(0.3, 0.3, 0.2), vit=None, pit=20, mu=0.5, warm_start=True, t2_damp=1.0,
                 manifold_reduce=False, manifold_max_points=4, manifold_on_host=False, profile=False,
                 graph_capture=False, graph_mode="device", pair_chunks=False, chunk_mode="registers", **kw):
        if graph_mode not in ("device", "spans"):
            raise ValueError("contact_engine_gpu_colored: graph_mode is 'device' or 'spans'")
        pair_chunks = bool(pair_chunks)
        if pair_chunks:
            if chunk_mode not in ("registers", "global"):
                raise ValueError("contact_engine_gpu_colored: chunk_mode is 'registers' or 'global'")
            if not (manifold_reduce and not manifold_on_host):
                raise ValueError("contact_engine_gpu_colored: pair_chunks=True needs the device manifold "
                                 "reduction (manifold_reduce=True, manifold_on_host=False): a chunk is the "
                                 "<= manifold_max_points contacts of one body pair")
        if manifold_reduce and int(manifold_max_points) != 4:
            raise NotImplementedError("contact_engine_gpu_colored: only manifold_max_points=4 is built")
        if manifold_reduce and not pair_chunks:
            raise NotImplementedError("contact_engine_gpu_colored: per-contact colouring of manifold-reduced "
                                      "contacts (pair_chunks=False with manifold_reduce=True) is not built; use "
                                      "pair_chunks=True or manifold_reduce=False")
        s.manifold_reduce = bool(manifold_reduce)
        s.manifold_max_points = int(manifold_max_points)
        s.manifold_on_host = bool(manifold_on_host)
        s.pair_chunks = pair_chunks
        s.graph_mode, s.chunk_mode = graph_mode, chunk_mode
        # the island-tiled solve is the default schedule: bit-identical to the per-colour launches
        # (tiled=False), which stay available
        kw.setdefault("tiled", True)
        vit = (40 if manifold_reduce else 20) if vit is None else int(vit)
        super().__init__(dims=dims, vit=vit, pit=pit, mu=mu, warm_start=warm_start, t2_damp=t2_damp,
                         per_contact=not pair_chunks, profile=profile, graph_capture=graph_capture, **kw)


class ColorCacheContactEngine(GraphColoredContactEngine):
    """ColorCache (contact_engine_gpu_color_cache.py:21-54 and its chain): normal-derived friction
    frame, cached impulses applied per colour before the sweeps, pair colouring reused while the
    pair list and invM are unchanged, force from the post-gravity velocity snapshot. As in the
    chain's FixedFrameContactEngine (contact_engine_gpu_fixed_frame.py:83-86) the manifold
    reduction and pair chunks are forced on, so ``vit`` defaults to 40 and ``manifold_on_host=True``
    is refused by the pair-chunk rule."""
    _solver = _lib.SOLVER_COLOR_CACHE

    def __init__(s, fast_math=False, **kw):
        """``fast_math=True`` (not in the reference; the tiled solve only): the same sweeps in a shorter
        instruction chain -- each body's angular response to a unit impulse and the reciprocal
        effective masses computed once per substep, fused multiply-adds, the hardware reciprocal
        square root in the friction cone. Bit-identical from run to run but not to the reference
        kernels: after one step the impulses agree with the exact solve's within 1e-5 of the
        largest and the positions within 1e-6 m. The adjoint refuses such an engine."""
        if fast_math and not kw.get("tiled", False):
            raise ValueError("fast_math needs the tiled solve (tiled=True)")
        s.fast_math = bool(fast_math)
        kw.update(manifold_reduce=True, pair_chunks=True, chunk_mode="registers")
        if kw.get("t2_damp", 1.0) != 1.0:
            raise ValueError("Applied warm start requires full tangent updates")
        # the reference's chain always captures its solve (contact_engine_gpu_fixed_frame.py:84); here
        # graph_capture=True captures the whole step once per (dt, substeps) and replays it, but the
        # default stays eager: on MI355X the replay measured 5 % slower than the launch sequence (the
        # reference's N = 10 000 lattice: 0.573 against 0.547 ms per step)
        super().__init__(**kw)
