"""Batched re-simulation: many sub-boxes of uneven sizes stepped together.

``Batch`` wraps ``ns_batch``: every kernel of the time step is launched once
for all boxes (each block looks its box up in a prefix of per-box block
counts), and steady steps are replayed from one HIP graph, so a step of 64
boxes costs 13 launches instead of 64 x 13.  Each box is an ordinary
``NSSolver`` (its own state, plane set, upload/download)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi

# the transform lengths the DCT/FFT plans factor (radices 2, 3, 5, 7)
def _smooth(n: int) -> bool:
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def box_sizes(count: int = 64, lo: int = 32, hi: int = 128, seed: int = 4):
    """count (nx, ny, nz) drawn uniformly per axis from the 7-smooth sizes in
    [lo, hi] (the lengths the transforms support)."""
    allowed = [n for n in range(lo, hi + 1) if _smooth(n)]
    rng = np.random.default_rng(seed)
    return [tuple(int(x) for x in rng.choice(allowed, 3)) for _ in range(count)]


class Batch:
    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.h for s in self.solvers])
        h = ctypes.c_void_p()
        _abi.call("ns_batch_create", arr, len(self.solvers), ctypes.byref(h))
        self.h = h

    def step(self, nsteps: int = 1):
        _abi.call("ns_batch_step", self.h, int(nsteps))

    def tune(self, trial_steps: int = 3) -> dict:
        """Pick the launch shapes per size class by timing steady batch steps
        (ns_batch_tune); the boxes' states are unchanged.  Returns the choice
        per class: {"small"|"medium"|"large": (rounds, dct_drop, x_sweep_lsh,
        tile)}, tile 0 meaning each box's own shape (the one with fewer idle
        lanes), else the predictor tile width (64 or 32)."""
        k = (ctypes.c_int * 12)()
        _abi.call("ns_batch_tune", self.h, int(trial_steps), k)
        return {name: tuple(k[4 * i + j] for j in range(4)) for i, name in enumerate(("small", "medium", "large"))}

    def close(self):
        if self.h:
            _abi.call("ns_batch_destroy", self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
