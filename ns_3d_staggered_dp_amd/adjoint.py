"""Discrete adjoint of the time step, with checkpoint/recompute over steps.

The library provides one reverse step (``ns_adjoint_step``): given the
gradient of a scalar objective with respect to the state at the end of a
forward step, it returns the gradient with respect to the state at its start
and with respect to that step's boundary plane set.  Walking a rollout
backwards needs the forward velocity at the start of every step; the
rollout keeps the full state every ``checkpoint_every`` steps on the device
and recomputes each segment forward before reversing it (memory: one state
per checkpoint plus u, v, w for the steps of one segment).

``NSRollout`` wraps a rollout as a ``torch.autograd.Function`` (FP64 CUDA
tensors in, FP64 CUDA tensors out), so an objective written in torch (e.g.
a sub-domain mismatch) differentiates through the solver.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from .config import NSConfig
from .grid import field_shapes, plane_set_length
from .solver import NSSolver, _ptr

STATE = ("u", "v", "w", "p", "hx", "hy", "hz")


def _comp(k):
    return {"u": "u", "v": "v", "w": "w", "p": "p", "hx": "u", "hy": "v", "hz": "w"}[k]


class Rollout:
    """nsteps forward steps of one context with per-step plane sets, and the
    reverse sweep.  State arrays may be numpy (host) or torch CUDA tensors
    holding the Fortran-ordered ghosted arrays."""

    def __init__(self, cfg: NSConfig, nsteps: int, checkpoint_every: int = 10):
        self.cfg = cfg
        self.n = int(nsteps)
        self.K = max(1, int(checkpoint_every))
        self.L = plane_set_length(cfg.nx, cfg.ny, cfg.nz)
        self.s = NSSolver(cfg)
        self.sh = field_shapes(cfg.nx, cfg.ny, cfg.nz)

    def close(self):
        self.s.close()

    def _empty_state(self, zero=False):
        import torch

        alloc = torch.zeros if zero else torch.empty
        return {k: alloc(int(np.prod(self.sh[_comp(k)])), dtype=torch.float64, device="cuda") for k in STATE}

    @staticmethod
    def _sync_torch():
        # the context runs on its own non-blocking stream: torch's pending work
        # on the tensors handed to it (inputs, freshly allocated outputs) must
        # be complete first
        import torch

        torch.cuda.current_stream().synchronize()

    def _download(self, out):
        _abi.call("ns_download_state", self.s.h, *(ctypes.c_void_p(out[k].data_ptr()) for k in STATE))

    def _upload(self, st, first):
        _abi.call("ns_upload_state", self.s.h, *(ctypes.c_void_p(st[k].data_ptr()) for k in STATE), int(first))

    def _planes(self, planes, n):
        if planes is None:
            return None
        return planes[n * self.L:(n + 1) * self.L]

    def forward(self, state0: dict, planes=None, first: bool = True) -> dict:
        """state0: dict of flat CUDA tensors (u, v, w, p, hx, hy, hz); planes:
        flat CUDA tensor of nsteps plane sets (or None: zero planes).  Returns
        the final state; checkpoints are kept for backward()."""
        import torch

        self.first = bool(first)
        self.planes = planes
        self.ckpt = {}
        st = {k: state0[k].detach().clone() for k in STATE}
        self._sync_torch()
        self._upload(st, first)
        for n in range(self.n):
            if n % self.K == 0:
                self.ckpt[n] = st if n == 0 else self._snapshot()
            p = self._planes(planes, n)
            if p is not None:
                _abi.call("ns_set_bc_planes", self.s.h, ctypes.c_void_p(p.data_ptr()))
            self.s.step(1)
        out = self._empty_state()
        self._download(out)
        torch.cuda.synchronize()
        return out

    def _snapshot(self):
        st = self._empty_state()
        self._download(st)
        return st

    def backward(self, grad_state: dict):
        """grad_state: gradients of the objective w.r.t. the final state (flat
        CUDA tensors, missing keys = zero).  Returns (gradients w.r.t. the
        initial state, gradients w.r.t. the plane sets (flat, nsteps sets))."""
        import torch

        g = {k: (grad_state[k].contiguous() if grad_state.get(k) is not None else None) for k in STATE}
        pbar = torch.zeros(self.n * self.L, dtype=torch.float64, device="cuda")
        self._sync_torch()
        _abi.call("ns_adjoint_set_state", self.s.h, *(ctypes.c_void_p(g[k].data_ptr()) if g[k] is not None else None
                                                      for k in STATE))
        starts = sorted(self.ckpt)
        for s0 in reversed(starts):
            s1 = min(s0 + self.K, self.n)
            # recompute the segment forward, keeping the velocity at each step's start
            self._upload(self.ckpt[s0], self.first and s0 == 0)
            vel = []
            for n in range(s0, s1):
                uvw = {k: torch.empty(int(np.prod(self.sh[k])), dtype=torch.float64, device="cuda")
                       for k in ("u", "v", "w")}
                _abi.call("ns_download_state", self.s.h, *(ctypes.c_void_p(uvw[k].data_ptr()) for k in ("u", "v", "w")),
                          None, None, None, None)
                vel.append(uvw)
                if n + 1 < s1:
                    p = self._planes(self.planes, n)
                    if p is not None:
                        _abi.call("ns_set_bc_planes", self.s.h, ctypes.c_void_p(p.data_ptr()))
                    self.s.step(1)
            for n in range(s1 - 1, s0 - 1, -1):
                uvw = vel[n - s0]
                _abi.call("ns_adjoint_step", self.s.h, *(ctypes.c_void_p(uvw[k].data_ptr()) for k in ("u", "v", "w")),
                          int(self.first and n == 0), ctypes.c_void_p(pbar[n * self.L:(n + 1) * self.L].data_ptr()))
        gs = self._empty_state()
        _abi.call("ns_adjoint_get_state", self.s.h, *(ctypes.c_void_p(gs[k].data_ptr()) for k in STATE))
        torch.cuda.synchronize()
        return gs, pbar


def make_rollout_function():
    """torch.autograd.Function over a Rollout: inputs (u0, v0, w0, planes) as
    flat FP64 CUDA tensors (p0 and the AB2 history start at zero, Euler first
    step, as a fresh run), outputs the final (u, v, w)."""
    import torch

    class NSRollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u0, v0, w0, planes, rollout: Rollout):
            st = rollout._empty_state(zero=True)
            st["u"].copy_(u0)
            st["v"].copy_(v0)
            st["w"].copy_(w0)
            out = rollout.forward(st, planes, first=True)
            ctx.rollout = rollout
            return out["u"], out["v"], out["w"]

        @staticmethod
        def backward(ctx, gu, gv, gw):
            gs, pbar = ctx.rollout.backward({"u": gu, "v": gv, "w": gw})
            return gs["u"], gs["v"], gs["w"], pbar, None

    return NSRollout
