"""Differentiable rollout of the ColorCache contact step (hand-written adjoint on the GPU).

``Rollout(bodies, steps, ...)`` steps a box scene with the tiled ColorCache solver and returns the
final centres; ``Rollout.grad`` (or the ``torch.autograd.Function`` ``ContactRollout``) returns the
gradient of a loss of the final centres with respect to the initial linear and angular velocities
and the per-body friction coefficients.

Memory is bounded by checkpoints every ``k`` steps: the forward keeps the engine state (body
tensors + workspace) at every k-th step only; the backward replays one segment at a time, recording
every substep into a tape (``me_step_taped``: sweep-start states and impulses, warm-start matches,
substep-start centres and orientations) and keeping the state after it, then reverses the segment
substep by substep, last first (``me_adjoint_substep``). The impulse adjoint travels from a substep
to the one before it through the warm start (matched feature keys), so gradients flow across steps
through the carried impulses as well as through the body state.

The differentiated function is the step exactly as the engine computes it: contact detection, the
choice of reference face, clipping side and kept points, the colouring, and the clamps of the
normal and friction impulses are the discrete branches the forward took, and the reverse is exact
for them (piecewise differentiable). A perturbation that changes one of those branches (a contact
appearing, a stick/slip switch) is outside what a derivative describes.

The reference's own adjoint is an NCP-level sensitivity (ncp/ncp_batch_gpu.py:1172-1187,
ncp/ncp_ref.py:786-962: dlambda/db and dlambda/dmu through the active set of a converged solve);
this rollout adjoint differentiates the actual iteration instead.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .bodies import BodySet
from .engines import ColorCacheContactEngine


def _torch():
    import torch
    return torch


def _vp(t):
    return C.c_void_p(t.data_ptr())


class Rollout:
    """T steps of a box scene under the tiled ColorCache solver, with gradients.

    bodies: the scene (any number of independent scenes by scene id; every island within one small
    tile, at most 256 pairs and 512 bodies). steps, dt, substeps, vit, pit, n_colors: the step.
    k: checkpoint interval in steps. max_contacts: contact capacity (default 8 per body)."""

    def __init__(self, bodies: BodySet, steps=32, dt=1 / 240, substeps=2, vit=16, pit=10, n_colors=16, k=8,
                 tile_bodies=128, max_contacts=None, max_pairs=None, device="cuda:0"):
        torch = _torch()
        self.T, self.dt, self.S, self.k = int(steps), float(dt), int(substeps), int(k)
        self.sdt = self.dt / self.S
        N = bodies.n
        self.eng = ColorCacheContactEngine(vit=vit, pit=pit, n_colors=n_colors, tiled=True, graph_capture=False,
                                           tile_bodies=tile_bodies,
                                           max_contacts=max_contacts or 8 * N, max_pairs=max_pairs,
                                           check_every_step=False, device=device)
        self.eng.add_boxes(bodies)
        self.eng._build()
        e = self.eng
        self.N, self.capC = N, None
        L = _lib.lib()
        sb = C.c_size_t()
        _lib.check(L.me_state_size(e._h, C.byref(sb)), "me_state_size")
        self._state = e._ws[e._ws_off:e._ws_off + int(sb.value)]  # the workspace's state block
        self._init = self._snap()
        nb = C.c_size_t()
        _lib.check(L.me_tape_size(e._h, C.byref(nb)), "me_tape_size")
        self.tape_bytes = int(nb.value)
        cnt = C.c_int64()
        ptr = C.c_void_p()
        _lib.check(L.me_buffer(e._h, _lib.BUF_JN, C.byref(ptr), C.byref(cnt)), "me_buffer")
        self.capC = int(cnt.value)
        dev = e.x.device
        self._jb = [torch.zeros(3 * self.capC, dtype=torch.float32, device=dev) for _ in range(2)]
        self._ckpt = []
        self.substeps_total = self.T * self.S

    # engine state: body tensors + the workspace's state block (status words, warm store, colour
    # cache, and what the reverse of a recorded substep reads); the rest of the workspace is scratch
    def _snap(self):
        """Bodies and the state block; the block by the used items of its arrays (me_state_copy)."""
        e = self.eng
        blk = _torch().empty_like(self._state)
        _lib.check(_lib.lib().me_state_copy(e._h, C.c_void_p(blk.data_ptr()), C.c_void_p(self._state.data_ptr()),
                                            C.c_void_p(e._stream())), "me_state_copy")
        return tuple(t.clone() for t in (e.x, e.q, e.v, e.w)) + (blk,)

    def _restore(self, s):
        e = self.eng
        for dst, src in zip((e.x, e.q, e.v, e.w), s[:4]):
            dst.copy_(src)
        _lib.check(_lib.lib().me_state_copy(e._h, C.c_void_p(self._state.data_ptr()), C.c_void_p(s[4].data_ptr()),
                                            C.c_void_p(e._stream())), "me_state_copy")

    def _check(self):
        ovf = self.eng.refresh_counters(raise_on_overflow=False)["overflow"]
        if ovf & _lib.OVF_TAPE:
            raise RuntimeError("adjoint: an island exceeds one small tile (256 pairs / 512 bodies)")
        self.eng.raise_overflow(ovf)

    def run(self, v0, w0, mu):
        """Forward: set initial velocities (N,3), (N,3) and friction (N,), step T steps; returns x_T."""
        torch = _torch()
        e = self.eng
        _lib.check(_lib.lib().me_adopt(e._h), "me_adopt")  # forward and backward may run on different threads
        with torch.no_grad():
            self._restore(self._init)
            e.v.copy_(v0.reshape(-1, 3))
            e.w.copy_(w0.reshape(-1, 3))
            e.mu_t.copy_(mu.reshape(-1))
            self._ckpt = []
            for t in range(self.T):
                if t % self.k == 0:
                    self._ckpt.append(self._snap())
                e.step(self.dt, self.S)
            self._check()
            return e.x.clone()

    def _tape(self):
        torch = _torch()
        buf = torch.empty(self.tape_bytes + 256, dtype=torch.uint8, device=self.eng.x.device)
        off = (-buf.data_ptr()) % 256
        return buf, buf.data_ptr() + off

    def grad(self, xbar_T):
        """Backward: adjoint of x_T (N,3) -> (v0bar, w0bar, mubar)."""
        torch = _torch()
        e = self.eng
        L = _lib.lib()
        _lib.check(L.me_adopt(e._h), "me_adopt")
        dev = e.x.device
        xbar = xbar_T.reshape(-1, 3).to(torch.float32).clone().contiguous()
        qbar = torch.zeros((self.N, 4), dtype=torch.float32, device=dev)
        vbar = torch.zeros((self.N, 3), dtype=torch.float32, device=dev)
        wbar = torch.zeros((self.N, 3), dtype=torch.float32, device=dev)
        mubar = torch.zeros(self.N, dtype=torch.float32, device=dev)
        jend, jprev = self._jb
        jend.zero_()
        stream = C.c_void_p(e._stream())
        with torch.no_grad():
            for seg in range(len(self._ckpt) - 1, -1, -1):
                self._restore(self._ckpt[seg])
                t0 = seg * self.k
                n_sub = (min(self.T, t0 + self.k) - t0) * self.S
                tapes, after = [], []
                for i in range(n_sub):
                    buf, ptr = self._tape()
                    _lib.check(L.me_step_taped(e._h, self.sdt, C.c_void_p(ptr), self.tape_bytes, stream),
                               "me_step_taped")
                    tapes.append((buf, ptr))
                    # the state after the segment's last substep is the engine's own: no copy
                    after.append(self._snap() if i < n_sub - 1 else None)
                self._check()
                for i in range(n_sub - 1, -1, -1):
                    if after[i] is not None:
                        self._restore(after[i])
                    _, ptr = tapes[i]
                    _lib.check(L.me_adjoint_substep(e._h, self.sdt, C.c_void_p(ptr), self.tape_bytes, _vp(xbar),
                                                    _vp(qbar), _vp(vbar), _vp(wbar), _vp(jend), _vp(jprev),
                                                    _vp(mubar), stream), "me_adjoint_substep")
                    jend, jprev = jprev, jend
                    tapes[i] = None
                    after[i] = None
            self._check()
        self._jb = [jend, jprev]
        return vbar, wbar, mubar


def contact_rollout(rollout: Rollout):
    """torch.autograd.Function over a Rollout: f(v0, w0, mu) = x_T."""
    torch = _torch()

    class ContactRollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v0, w0, mu):
            return rollout.run(v0.detach(), w0.detach(), mu.detach())

        @staticmethod
        def backward(ctx, xbar):
            vb, wb, mb = rollout.grad(xbar.contiguous())
            return vb, wb, mb

    return ContactRollout.apply
