"""Re-simulation error metrics and the reference's text output formats.

Per step the reference compares the re-simulated fields (u, v, w, p) with
the parent's sub-box fields (u_sub, ...) over the full ghosted arrays
(re_simulation_sub_tstep_interBC_combine.f90:826-833):

    rms(a)  = sqrt(sum(a*a)/size(a))                  (other_utility.f90:124-129)
    err(1:4) = maxval(|u - u_sub|) / rms(u_sub), ... p
    err(5:7) = mean(|u - u_sub|)  / rms(u_sub), ... w  (mean: other_utility.f90:113-118)
    err(8)   = maxval(|oz - oz_sub|) / rms(oz_sub),  oz = diff(v,1,1)/dx - diff(u,1,2)/dy

The device computes them in one pass (``NSSolver.error_metrics``); this
module holds the host-side formats: the file names (…interBC:146-158,
907-913) and the ``I0, 8 x ES12.5E2`` comma-separated tables (:914-927).
"""
from __future__ import annotations

import math
import os

import numpy as np

ERR_HEADER = ("time_step, rel_err_u, rel_err_v, rel_err_w, rel_err_p, "
              "rel_err_u_m, rel_err_v_m, rel_err_w_m, rel_err_p_m")
RMS_HEADER = "time_step, rms_u, rms_v, rms_w, rms_p"


def fortran_es(x: float, w: int = 12, d: int = 5, e: int = 2) -> str:
    """Fortran ``ESw.dEe`` edit descriptor: d.ddddd mantissa, signed
    exponent of e digits, right-justified in w columns; non-finite values
    print as Infinity / -Infinity / NaN; a field too narrow prints w '*'."""
    x = float(x)
    if math.isnan(x):
        s = "NaN"
    elif math.isinf(x):
        s = "Infinity" if x > 0 else "-Infinity"
    else:
        mant, ex = f"{x:.{d}E}".split("E")
        ev = int(ex)
        if abs(ev) >= 10 ** e:
            return "*" * w
        s = f"{mant}E{'-' if ev < 0 else '+'}{abs(ev):0{e}d}"
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_es_short(x: float) -> str:
    """``ES5.0E1`` as used in the reference's file names (``1.E-3``)."""
    mant, ex = f"{float(x):.0E}".split("E")
    ev = int(ex)
    return f"{mant}.E{'-' if ev < 0 else '+'}{abs(ev)}"


def run_tag(timescheme: str, dt0: float, pbc_x: int, using_ustar: bool = True, toffset: bool = True,
            restart: bool = False, noise: float = 0.0, sub_tstep: int = 1, bc_interp: bool = False) -> str:
    """The run tag string_var2 (re_simulation_sub_tstep_interBC_combine.f90:146-158)."""
    s = f"{timescheme}_{fortran_es_short(dt0)}"
    s += {2: "_D", 3: "_N", 4: "_Dg"}.get(pbc_x, "")
    s += "_Ustar" if using_ustar else "_U"
    if toffset:
        s += "_TOffset"
    if restart:
        s += "_restart"
    if noise != 0.0:
        s += f"_noise{fortran_es_short(noise)}"
    if sub_tstep != 1:
        s += f"_sub_tstep{sub_tstep}"
    if bc_interp:
        s += "_BCinterp"
    return s


def err_file_path(interp_scheme: str, tag: str, dp_flag: int, root: str = ".") -> str:
    """err_file/<interp>/err_vel_<tag>.txt (perr_vel_ when dp_flag = 0), …interBC:907-911."""
    name = ("err_vel_" if dp_flag else "perr_vel_") + tag + ".txt"
    return os.path.join(root, "err_file", interp_scheme, name)


def rms_file_path(interp_scheme: str, timescheme: str, dt0: float, root: str = ".") -> str:
    """err_file/<interp>/rms_vel_<scheme>_<dt0>.txt (…interBC:920)."""
    return os.path.join(root, "err_file", interp_scheme, f"rms_vel_{timescheme}_{fortran_es_short(dt0)}.txt")


def format_table(header: str, rows) -> str:
    """One header line, then ``step, v1, ..., vk`` per row with ES12.5E2 values."""
    rows = np.asarray(rows, dtype=np.float64)
    out = [header]
    for j, r in enumerate(rows):
        out.append(f"{j}," + ",".join(fortran_es(x) for x in r))
    return "\n".join(out) + "\n"


def write_err_file(path: str, err_rows) -> None:
    """err_rows: (time_length+1, 8) per-step errors, row j = time step j."""
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as f:
        f.write(format_table(ERR_HEADER, err_rows))


def write_rms_file(path: str, rms_rows) -> None:
    """rms_rows: (time_length+1, 4) rms of the parent's u, v, w, p per step."""
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as f:
        f.write(format_table(RMS_HEADER, rms_rows))


def read_table(path: str):
    """Parse a table written as above (Infinity/NaN accepted); returns (header, steps, values)."""
    with open(path) as f:
        lines = f.read().splitlines()
    steps, vals = [], []
    for ln in lines[1:]:
        if not ln.strip():
            continue
        parts = ln.split(",")
        steps.append(int(parts[0]))
        vals.append([float(p.replace("Infinity", "inf")) for p in parts[1:]])
    return lines[0], np.asarray(steps), np.asarray(vals, dtype=np.float64)


def reference_metrics(u, v, w, p, u_sub, v_sub, w_sub, p_sub, dx, dy):
    """The same metrics in numpy, evaluated literally (for checks and small runs)."""
    def rms(a):
        return np.sqrt(np.sum(a * a) / a.size)

    rmsv = np.array([rms(u_sub), rms(v_sub), rms(w_sub), rms(p_sub)])
    err = np.zeros(8)
    for i, (a, b) in enumerate(((u, u_sub), (v, v_sub), (w, w_sub), (p, p_sub))):
        err[i] = np.max(np.abs(a - b)) / rmsv[i]
    for i, (a, b) in enumerate(((u, u_sub), (v, v_sub), (w, w_sub))):
        err[4 + i] = (np.sum(np.abs(a - b)) / a.size) / rmsv[i]
    oz = np.diff(v, axis=0) / dx - np.diff(u, axis=1) / dy
    oz_sub = np.diff(v_sub, axis=0) / dx - np.diff(u_sub, axis=1) / dy
    err[7] = np.max(np.abs(oz - oz_sub)) / rms(oz_sub)
    return err, rmsv
