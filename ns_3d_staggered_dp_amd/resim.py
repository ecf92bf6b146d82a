"""Re-simulation drivers: boundary-data interpolation in time and the
sub-step schedule of re_simulation_sub_tstep_interBC_combine.f90.

Snapshots are plane sets (``grid.plane_set_length`` doubles each) at times
k*snap_dt, k = 0..nsnap-1, stored contiguously (host or device).  The
interpolation runs on the device (``ns_resim_advance``: linear,
``ns_resim_advance_spline`` / ``ns_bc_interp_spline``: MATLAB's not-a-knot
spline over six snapshots, sub_tstep_BC_combined_interpolation.m:47-64).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _abi


def spline_weights(t, x: float) -> np.ndarray:
    """w with spline(x) = sum_i w[i]*y[i] for the not-a-knot cubic spline
    through (t[i], y[i]) (host computation in libns_hip, no GPU)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    w = np.zeros(len(t))
    _abi.call("ns_spline_weights", t.ctypes.data_as(ctypes.c_void_p), len(t), float(x),
              w.ctypes.data_as(ctypes.c_void_p))
    return w


def spline_window(nsnap: int, snap_dt: float, t: float):
    """First snapshot and count of the spline window used for time t: six
    snapshots, three at or before t (…interpolation.m:56-57), kept inside
    the history (the same rule as the library)."""
    m = min(nsnap, 6)
    kle = min(int(math.floor(t / snap_dt)), nsnap - 1)
    s0 = max(0, min(kle - 2, nsnap - m))
    return s0, m


def sub_tstep_times(t_start: float, dt0: float, sub_tstep: int, nfull: int):
    """End times of the steps of a re-simulation started at t_start:
    sub_tstep steps of dt0/sub_tstep up to t_start + dt0, then nfull steps of
    dt0 (time_array, …interBC_combine.f90:146, without its t=0 entry)."""
    ts = [t_start + dt0 / sub_tstep * i for i in range(1, sub_tstep + 1)]
    ts += [t_start + dt0 * (1 + i) for i in range(1, nfull + 1)]
    return ts


def run_sub_tstep(solver, snapshots, nsnap: int, snap_dt: float, t_start: float, dt0: float, sub_tstep: int,
                  nfull: int, interp: str = "spline"):
    """Advance ``solver`` from t_start (its uploaded state) through
    sub_tstep sub-steps of dt0/sub_tstep with boundary planes interpolated in
    time, then nfull steps of dt0 with the snapshot planes; times are
    relative to snapshot 0.  The AB2 history is saved after the first
    sub-step and restored after the last (…interBC_combine.f90:602-612), so
    the first full step extrapolates over one dt0.  interp: "spline" or
    "linear" for the sub-steps."""
    if sub_tstep > 1:
        solver.set_dt(dt0 / sub_tstep)
        for i in range(1, sub_tstep + 1):
            t = t_start + dt0 / sub_tstep * i
            if interp == "spline":
                solver.bc_interp_spline(snapshots, nsnap, snap_dt, t)
            else:
                k = min(max(int(math.floor(t / snap_dt)), 0), nsnap - 2)
                th = (t - k * snap_dt) / snap_dt
                solver.bc_interp(_slice(snapshots, k, nsnap), _slice(snapshots, k + 1, nsnap), th)
            solver.step(1)
            if i == 1:
                solver.history_save()
            if i == sub_tstep:
                solver.history_restore()
        solver.set_dt(dt0)
        t0 = t_start + dt0
    else:
        solver.set_dt(dt0)
        t0 = t_start
    if nfull > 0:
        # full steps end on snapshot times: linear interpolation with theta = 0 is the snapshot itself
        solver.resim_advance(snapshots, nsnap, snap_dt, t0, nfull)


def _slice(snapshots, k: int, nsnap: int):
    """plane set k of a flat (host numpy or device torch) history"""
    n = (snapshots.numel() if hasattr(snapshots, "numel") else len(snapshots)) // nsnap
    return snapshots[k * n:(k + 1) * n]
