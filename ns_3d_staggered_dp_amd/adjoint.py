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


class Tape:
    """What one forward call leaves for its reverse sweep: the checkpointed
    states, the plane sets it stepped with (None: zero planes) and whether its
    first step was the Euler start.  Each call has its own, so several
    forward calls on one Rollout can be reversed in any order."""

    def __init__(self, first: bool, planes):
        self.first = bool(first)
        self.planes = planes
        self.ckpt = {}


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
        self.tape = None  # the last forward call's
        self._zeros = None

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

    def _set_planes(self, planes, n):
        """load step n's plane set; with no planes, the zero set (the
        context otherwise keeps whatever set it was given last)"""
        import torch

        if planes is None:
            if self._zeros is None:
                self._zeros = torch.zeros(self.L, dtype=torch.float64, device="cuda")
                torch.cuda.current_stream().synchronize()
            p = self._zeros
        else:
            p = planes[n * self.L:(n + 1) * self.L]
        _abi.call("ns_set_bc_planes", self.s.h, ctypes.c_void_p(p.data_ptr()))

    def forward(self, state0: dict, planes=None, first: bool = True) -> dict:
        """state0: dict of flat CUDA tensors (u, v, w, p, hx, hy, hz); planes:
        flat CUDA tensor of nsteps plane sets (or None: zero planes).  Returns
        the final state; the call's Tape (checkpoints for the reverse sweep)
        is kept as self.tape.  forward_with_tape returns it instead."""
        out, self.tape = self.forward_with_tape(state0, planes, first)
        return out

    def forward_with_tape(self, state0: dict, planes=None, first: bool = True):
        import torch

        tape = Tape(first, planes)
        st = {k: state0[k].detach().clone() for k in STATE}
        self._sync_torch()
        self._upload(st, first)
        for n in range(self.n):
            if n % self.K == 0:
                tape.ckpt[n] = st if n == 0 else self._snapshot()
            if planes is not None or n == 0:
                self._set_planes(planes, n)
            self.s.step(1)
        out = self._empty_state()
        self._download(out)
        torch.cuda.synchronize()
        return out, tape

    def _snapshot(self):
        st = self._empty_state()
        self._download(st)
        return st

    def backward(self, grad_state: dict, tape: Tape | None = None):
        """grad_state: gradients of the objective w.r.t. the final state (flat
        CUDA tensors, missing keys = zero) of the forward call that made
        `tape` (default: the last one).  Returns (gradients w.r.t. the
        initial state, gradients w.r.t. the plane sets (flat, nsteps sets))."""
        import torch

        tape = self.tape if tape is None else tape
        if tape is None:
            raise RuntimeError("Rollout.backward before any forward call")
        g = {k: (grad_state[k].contiguous() if grad_state.get(k) is not None else None) for k in STATE}
        pbar = torch.zeros(self.n * self.L, dtype=torch.float64, device="cuda")
        self._sync_torch()
        _abi.call("ns_adjoint_set_state", self.s.h, *(ctypes.c_void_p(g[k].data_ptr()) if g[k] is not None else None
                                                      for k in STATE))
        starts = sorted(tape.ckpt)
        for s0 in reversed(starts):
            s1 = min(s0 + self.K, self.n)
            # recompute the segment forward, keeping the velocity at each step's start
            self._upload(tape.ckpt[s0], tape.first and s0 == 0)
            vel = []
            for n in range(s0, s1):
                uvw = {k: torch.empty(int(np.prod(self.sh[k])), dtype=torch.float64, device="cuda")
                       for k in ("u", "v", "w")}
                _abi.call("ns_download_state", self.s.h, *(ctypes.c_void_p(uvw[k].data_ptr()) for k in ("u", "v", "w")),
                          None, None, None, None)
                vel.append(uvw)
                if n + 1 < s1:
                    if tape.planes is not None or n == s0:
                        self._set_planes(tape.planes, n)
                    self.s.step(1)
            for n in range(s1 - 1, s0 - 1, -1):
                uvw = vel[n - s0]
                _abi.call("ns_adjoint_step", self.s.h, *(ctypes.c_void_p(uvw[k].data_ptr()) for k in ("u", "v", "w")),
                          int(tape.first and n == 0), ctypes.c_void_p(pbar[n * self.L:(n + 1) * self.L].data_ptr()))
        gs = self._empty_state()
        _abi.call("ns_adjoint_get_state", self.s.h, *(ctypes.c_void_p(gs[k].data_ptr()) for k in STATE))
        torch.cuda.synchronize()
        return gs, pbar


def make_rollout_function():
    """torch.autograd.Function over a Rollout: inputs (u0, v0, w0, planes) as
    flat FP64 CUDA tensors (planes may be None: zero planes; p0 and the AB2
    history start at zero, Euler first step, as a fresh run), outputs the
    final (u, v, w).  Each application keeps its own tape, so a loss built
    from several applications on one Rollout differentiates correctly."""
    import torch

    class NSRollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u0, v0, w0, planes, rollout: Rollout):
            st = rollout._empty_state(zero=True)
            st["u"].copy_(u0)
            st["v"].copy_(v0)
            st["w"].copy_(w0)
            out, tape = rollout.forward_with_tape(st, planes, first=True)
            ctx.rollout = rollout
            ctx.tape = tape
            ctx.no_planes = planes is None
            return out["u"], out["v"], out["w"]

        @staticmethod
        def backward(ctx, gu, gv, gw):
            gs, pbar = ctx.rollout.backward({"u": gu, "v": gv, "w": gw}, tape=ctx.tape)
            return gs["u"], gs["v"], gs["w"], (None if ctx.no_planes else pbar), None

    return NSRollout
