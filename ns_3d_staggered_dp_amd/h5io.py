"""The reference's HDF5 files (libns_h5, include/ns_h5.h).

File kinds, as the reference writes and reads them:

- checkpoint ``<scheme>_result/HIT_<nx>^3_decay_<dt0>_<scheme>_dp_t_<t>.h5``:
  root attributes nx, dt, nu, time_scheme, time_step, time; datasets u, v,
  w, p, rhs_x_previous, rhs_y_previous, rhs_z_previous
  (big_periodic_3D.f90:620-644; restart read at :309-325);
- parent sub-box history ``..._dp_x0_<x0>_nx0_<nx0>_sub.h5``: root
  attributes (+ x0, nx0); one group ``t_<F0.4>`` per step with u_sub, v_sub,
  w_sub, p_sub, u_star_sub, v_star_sub, w_star_sub, dp_sub (+ RHS_poisson_sub),
  rhs_*_previous at integer times, and for AB2-CN the 18 post-sweep planes
  bx_u_1s ... bz_w_nzs (:589-712; read at re_simulation_sub_tstep_interBC_combine.f90:403-476);
- spline boundary file ``..._sub_BC_spline.h5``: groups ``t_<F0.5>`` with
  u_star_sub, v_star_sub, w_star_sub, dp_sub (sub_tstep_BC_combined_interpolation.m:66-94;
  read at …interBC:486-493);
- sub-step boundary arrays: group ``sub_t_<n>`` with time-first arrays
  (n, nx+1, ny+2, nz+2) ... (NS_functions.f90:1633-1664).

Arrays are numpy float64 in Fortran order with the reference's shapes; the
library stores them as the Fortran HDF5 API does (dims reversed in the file).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .grid import field_shapes, plane_layout

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libns_h5.so")
SUB_FIELDS = ("u_sub", "v_sub", "w_sub", "p_sub", "u_star_sub", "v_star_sub", "w_star_sub", "dp_sub")
BC_FIELDS = ("u_star_sub", "v_star_sub", "w_star_sub", "dp_sub")
HISTORY = ("rhs_x_previous", "rhs_y_previous", "rhs_z_previous")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (python -m ns_3d_staggered_dp_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        P, I, D, C = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
        LL = ctypes.c_longlong
        sig = {
            "nsh5_create": [C, ctypes.POINTER(LL)], "nsh5_open": [C, I, ctypes.POINTER(LL)], "nsh5_close": [LL],
            "nsh5_exists": [LL, C], "nsh5_write_double": [LL, C, I, P, P], "nsh5_read_double": [LL, C, I, P, P],
            "nsh5_get_dims": [LL, C, ctypes.POINTER(I), P], "nsh5_set_attr_int": [LL, C, C, I],
            "nsh5_set_attr_double": [LL, C, C, D], "nsh5_set_attr_string": [LL, C, C, C],
            "nsh5_get_attr_int": [LL, C, C, ctypes.POINTER(I)], "nsh5_get_attr_double": [LL, C, C, ctypes.POINTER(D)],
            "nsh5_get_attr_string": [LL, C, C, ctypes.c_char_p, ctypes.c_long],
            "nsh5_list": [LL, C, ctypes.c_char_p, ctypes.c_long],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = I
        _lib = L
    return _lib


class H5Error(RuntimeError):
    pass


def _chk(rc, what):
    if rc != 0:
        raise H5Error(f"{what}: status {rc}")


def _b(s: str) -> bytes:
    return s.encode()


class H5File:
    """mode "r" (read-only), "r+" (read-write), "x" (create; fails if it exists)."""

    def __init__(self, path: str, mode: str = "r"):
        self.path = path
        h = ctypes.c_longlong()
        if mode == "x":
            d = os.path.dirname(path)
            if d:
                os.makedirs(d, exist_ok=True)
            _chk(lib().nsh5_create(_b(path), ctypes.byref(h)), f"create {path}")
        else:
            _chk(lib().nsh5_open(_b(path), 1 if mode == "r+" else 0, ctypes.byref(h)), f"open {path}")
        self.h = h.value

    def close(self):
        if self.h is not None:
            _chk(lib().nsh5_close(self.h), f"close {self.path}")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def exists(self, path: str) -> bool:
        return lib().nsh5_exists(self.h, _b(path)) == 1

    def write(self, path: str, a):
        a = np.asfortranarray(a, dtype=np.float64)
        dims = (ctypes.c_long * a.ndim)(*a.shape)
        _chk(lib().nsh5_write_double(self.h, _b(path), a.ndim, dims, a.ctypes.data_as(ctypes.c_void_p)),
             f"write {path}")

    def dims(self, path: str):
        r = ctypes.c_int()
        d = (ctypes.c_long * 8)()
        _chk(lib().nsh5_get_dims(self.h, _b(path), ctypes.byref(r), d), f"dims {path}")
        return tuple(d[:r.value])

    def read(self, path: str, shape=None) -> np.ndarray:
        shape = tuple(shape) if shape is not None else self.dims(path)
        a = np.zeros(shape, dtype=np.float64, order="F")
        dims = (ctypes.c_long * len(shape))(*shape)
        _chk(lib().nsh5_read_double(self.h, _b(path), len(shape), dims, a.ctypes.data_as(ctypes.c_void_p)),
             f"read {path} as {shape}")
        return a

    def set_attr(self, obj: str, name: str, value):
        if isinstance(value, str):
            _chk(lib().nsh5_set_attr_string(self.h, _b(obj), _b(name), _b(value)), name)
        elif isinstance(value, (int, np.integer)) and not isinstance(value, bool):
            _chk(lib().nsh5_set_attr_int(self.h, _b(obj), _b(name), int(value)), name)
        else:
            _chk(lib().nsh5_set_attr_double(self.h, _b(obj), _b(name), float(value)), name)

    def get_attr(self, obj: str, name: str, kind=float):
        if kind is str:
            buf = ctypes.create_string_buffer(4096)
            _chk(lib().nsh5_get_attr_string(self.h, _b(obj), _b(name), buf, 4096), name)
            return buf.value.decode()
        if kind is int:
            v = ctypes.c_int()
            _chk(lib().nsh5_get_attr_int(self.h, _b(obj), _b(name), ctypes.byref(v)), name)
            return v.value
        v = ctypes.c_double()
        _chk(lib().nsh5_get_attr_double(self.h, _b(obj), _b(name), ctypes.byref(v)), name)
        return v.value

    def list(self, group: str = "/"):
        buf = ctypes.create_string_buffer(1 << 20)
        _chk(lib().nsh5_list(self.h, _b(group), buf, 1 << 20), f"list {group}")
        s = buf.value.decode()
        return s.split("\n") if s else []


# ---- names -------------------------------------------------------------

def fortran_f0(x: float, d: int) -> str:
    """Fortran ``F0.d`` as the reference's compiler writes it: minimal width,
    no leading zero before the point for |x| < 1 (the MATLAB readers strip it
    the same way, sub_tstep_BC_combined_interpolation.m:37-38)."""
    s = f"{float(x):.{d}f}"
    if s.startswith("0."):
        s = s[1:]
    elif s.startswith("-0."):
        s = "-" + s[2:]
    return s


def t_group(t: float, digits: int = 4) -> str:
    """``t_<F0.4>`` (parent sub-box steps) or ``t_<F0.5>`` (interpolated sub-steps)."""
    return "t_" + fortran_f0(t, digits)


def _es5(x: float) -> str:
    from .metrics import fortran_es_short

    return fortran_es_short(x)


def sub_file_name(timescheme: str, nx: int, dt0: float, dp_flag: int, x0: int, nx0: int,
                  result_dir: str | None = None) -> str:
    """big_periodic_3D.f90:588-592 (``_result`` directory; the re-sims read ``_result.MARCC``)."""
    d = result_dir or f"{timescheme}_result"
    kind = "dp" if dp_flag else "p"
    return f"{d}/HIT_{nx}^3_decay_{_es5(dt0)}_{timescheme}_{kind}_x0_{x0}_nx0_{nx0}_sub.h5"


def checkpoint_file_name(timescheme: str, nx: int, dt0: float, dp_flag: int, t: float) -> str:
    """big_periodic_3D.f90:621-625."""
    kind = "dp" if dp_flag else "p"
    return f"{timescheme}_result/HIT_{nx}^3_decay_{_es5(dt0)}_{timescheme}_{kind}_t_{fortran_f0(t, 4)}.h5"


# ---- file kinds ----------------------------------------------------------

def write_run_attrs(f: H5File, nx: int, dt0: float, nu: float, timescheme: str, time_step: int, time: float,
                    x0: int | None = None, nx0: int | None = None):
    f.set_attr("/", "nx", int(nx))
    f.set_attr("/", "dt", float(dt0))
    f.set_attr("/", "nu", float(nu))
    f.set_attr("/", "time_scheme", timescheme)
    f.set_attr("/", "time_step", int(time_step))
    f.set_attr("/", "time", float(time))
    if x0 is not None:
        f.set_attr("/", "x0", int(x0))
        f.set_attr("/", "nx0", int(nx0))


def write_checkpoint(path: str, state: dict, nx: int, dt0: float, nu: float, timescheme: str, time_step: int,
                     time: float):
    """state: u, v, w, p, hx, hy, hz (NSSolver.download keys)."""
    with H5File(path, "x") as f:
        write_run_attrs(f, nx, dt0, nu, timescheme, time_step, time)
        for k in ("u", "v", "w", "p"):
            f.write(k, state[k])
        for k, name in zip(("hx", "hy", "hz"), HISTORY):
            f.write(name, state[k])


def read_checkpoint(path: str, nx: int, ny: int, nz: int) -> dict:
    sh = field_shapes(nx, ny, nz)
    with H5File(path) as f:
        out = {k: f.read(k, sh[k]) for k in ("u", "v", "w", "p")}
        for k, name, c in zip(("hx", "hy", "hz"), HISTORY, ("u", "v", "w")):
            out[k] = f.read(name, sh[c])
        out["attrs"] = {"nx": f.get_attr("/", "nx", int), "dt": f.get_attr("/", "dt"), "nu": f.get_attr("/", "nu"),
                        "time_scheme": f.get_attr("/", "time_scheme", str),
                        "time_step": f.get_attr("/", "time_step", int), "time": f.get_attr("/", "time")}
    return out


def _sub_shapes(nx0: int, ny0: int, nz0: int) -> dict:
    sh = field_shapes(nx0, ny0, nz0)
    return {"u_sub": sh["u"], "v_sub": sh["v"], "w_sub": sh["w"], "p_sub": sh["p"], "u_star_sub": sh["u"],
            "v_star_sub": sh["v"], "w_star_sub": sh["w"], "dp_sub": sh["p"], "RHS_poisson_sub": sh["p"],
            "rhs_x_previous": sh["u"], "rhs_y_previous": sh["v"], "rhs_z_previous": sh["w"]}


def sweep_plane_names(nx0: int, ny0: int, nz0: int):
    """(dataset name, packed-set name, shape) of the 18 post-sweep velocity
    planes in the reference's write order (big_periodic_3D.f90:686-706)."""
    lay, _ = plane_layout(nx0, ny0, nz0)
    shapes = {name: shape for name, shape, _ in lay}
    out = []
    for ax in ("x", "y", "z"):
        for c in ("u", "v", "w"):
            for end in ("1", "n" + ax):
                pname = f"b{ax}_{c}_{end}"
                out.append((pname + "s", pname, shapes[pname]))
    return out


def write_sub_group(f: H5File, t: float, fields: dict, history: dict | None = None, planes: dict | None = None):
    """One parent step of the sub-box history: fields holds SUB_FIELDS (and
    optionally RHS_poisson_sub), history the three rhs_*_previous boxes
    (written at integer times), planes the packed-set planes by name (AB2-CN)."""
    g = t_group(t, 4)
    for k in SUB_FIELDS + ("RHS_poisson_sub",):
        if k in fields:
            f.write(f"{g}/{k}", fields[k])
    for k in HISTORY:
        if history and k in history:
            f.write(f"{g}/{k}", history[k])
    if planes:
        for dname, pname, _ in sweep_plane_names(*_box_dims(fields["u_sub"])):
            f.write(f"{g}/{dname}", planes[pname])


def _box_dims(u_sub):
    s = u_sub.shape
    return s[0] - 1, s[1] - 2, s[2] - 2


def read_sub_group(f: H5File, t: float, nx0: int, ny0: int, nz0: int, names=SUB_FIELDS) -> dict:
    g = t_group(t, 4)
    sh = _sub_shapes(nx0, ny0, nz0)
    return {k: f.read(f"{g}/{k}", sh[k]) for k in names if f.exists(f"{g}/{k}")}


def write_bc_group(f: H5File, t: float, fields: dict):
    g = t_group(t, 5)
    for k in BC_FIELDS:
        f.write(f"{g}/{k}", fields[k])


def read_bc_group(f: H5File, t: float, nx0: int, ny0: int, nz0: int) -> dict:
    g = t_group(t, 5)
    sh = _sub_shapes(nx0, ny0, nz0)
    return {k: f.read(f"{g}/{k}", sh[k]) for k in BC_FIELDS}


def write_sub_tstep_group(f: H5File, sub_tstep: int, series: dict):
    """series[name]: (sub_tstep, *box shape) time-first arrays (NS_functions.f90:1650-1659)."""
    for k, a in series.items():
        f.write(f"sub_t_{sub_tstep}/{k}", a)


def read_sub_tstep_group(f: H5File, sub_tstep: int, nx0: int, ny0: int, nz0: int) -> dict:
    sh = _sub_shapes(nx0, ny0, nz0)
    return {k: f.read(f"sub_t_{sub_tstep}/{k}", (sub_tstep,) + sh[k]) for k in SUB_FIELDS}
