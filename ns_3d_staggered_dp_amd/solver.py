"""Host-side mirror of the reference's step and subroutine interfaces.

``NSSolver`` owns one ``ns_ctx`` (device-resident state, one HIP stream).
Its methods carry the reference's subroutine names and argument meaning
(NS_functions.f90): ``cal_conv``, ``cal_diff``, ``vel_bc_staggered``,
``vel_bc_staggered_CN2``, ``get_velpr_bc``, ``get_pr_bc``,
``pr_bc_staggered_modify_rhs``, ``DCT_poisson_solver``, ``pr_bc_staggered``,
plus the inline driver pieces (divergence, FFT Poisson, projection, ADI
sweep) and the fused ``step``.  Arrays are numpy float64 in Fortran order
(host) or torch float64 tensors on the GPU laid out the same way
(column-major = a C-contiguous tensor of the reversed shape, see
``to_device``/``from_device``).  Everything runs through libns_hip.so.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from .config import NSConfig
from .grid import field_shapes, plane_set_length


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if a.dtype != np.float64:
            raise TypeError("arrays must be float64")
        if not (a.flags.f_contiguous or a.ndim <= 1):
            raise ValueError("numpy arrays must be Fortran-contiguous")
        return ctypes.c_void_p(a.ctypes.data)
    # torch tensor (device or host)
    try:
        import torch
    except ImportError:  # pragma: no cover
        raise TypeError(f"unsupported array type {type(a)}")
    if a.dtype.__repr__() != "torch.float64":
        raise TypeError("tensors must be float64")
    if not a.is_contiguous():
        raise ValueError("tensors must be contiguous")
    if a.is_cuda:
        # the context works on its own non-blocking stream: torch's pending
        # work on the tensor must be complete before the library touches it
        torch.cuda.current_stream(a.device).synchronize()
    return ctypes.c_void_p(a.data_ptr())


def fzeros(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def to_device(a: np.ndarray, device="cuda"):
    """Fortran-ordered numpy array -> contiguous torch tensor with reversed
    shape (same bytes, column-major view of the reference's array)."""
    import torch

    return torch.from_numpy(np.ascontiguousarray(a.T)).to(device)


def from_device(t) -> np.ndarray:
    return np.asfortranarray(t.detach().cpu().numpy().T)


class NSSolver:
    def __init__(self, cfg: NSConfig):
        self.cfg = cfg
        self._c = cfg.to_c()
        h = ctypes.c_void_p()
        _abi.call("ns_create", ctypes.byref(self._c), ctypes.byref(h))
        self.h = h
        self.shapes = field_shapes(cfg.nx, cfg.ny, cfg.nz)
        self.nplanes = plane_set_length(cfg.nx, cfg.ny, cfg.nz)

    def close(self):
        if getattr(self, "h", None):
            _abi.lib().ns_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- state ----
    def upload(self, u, v, w, p=None, hx=None, hy=None, hz=None, first=True):
        _abi.call("ns_upload_state", self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(p), _ptr(hx), _ptr(hy), _ptr(hz),
                  int(bool(first)))

    def download(self):
        s = self.shapes
        u, v, w, p = fzeros(s["u"]), fzeros(s["v"]), fzeros(s["w"]), fzeros(s["p"])
        hx, hy, hz = fzeros(s["u"]), fzeros(s["v"]), fzeros(s["w"])
        _abi.call("ns_download_state", self.h, *(_ptr(x) for x in (u, v, w, p, hx, hy, hz)))
        return dict(u=u, v=v, w=w, p=p, hx=hx, hy=hy, hz=hz)

    def download_work(self):
        s = self.shapes
        us, vs, ws, dp = fzeros(s["u"]), fzeros(s["v"]), fzeros(s["w"]), fzeros(s["p"])
        _abi.call("ns_download_work", self.h, *(_ptr(x) for x in (us, vs, ws, dp)))
        return dict(u_star=us, v_star=vs, w_star=ws, dp=dp)

    def state_pointers(self):
        arr = (ctypes.c_void_p * 11)()
        _abi.call("ns_state_pointers", self.h, arr)
        keys = ["u", "v", "w", "p", "hx", "hy", "hz", "u_star", "v_star", "w_star", "dp"]
        return {k: arr[i] for i, k in enumerate(keys)}

    def state_pitch(self):
        """Row pitch (elements) of the arrays state_pointers() points at: entry
        (i, j, k) of an array with y extent n2 is at (i-1) + pitch*((j-1) + n2*(k-1))."""
        out = ctypes.c_int(0)
        _abi.call("ns_state_pitch", self.h, ctypes.byref(out))
        return out.value

    def synchronize(self):
        _abi.call("ns_synchronize", self.h)

    # ---- fused step ----
    def set_bc_planes(self, planes):
        _abi.call("ns_set_bc_planes", self.h, _ptr(planes))

    def get_bc_planes(self):
        out = np.zeros(self.nplanes)
        _abi.call("ns_get_bc_planes", self.h, _ptr(out))
        return out

    def bc_interp(self, a, b, theta: float):
        _abi.call("ns_bc_interp", self.h, _ptr(a), _ptr(b), float(theta))

    def step(self, nsteps: int = 1):
        _abi.call("ns_step", self.h, int(nsteps))

    def resim_advance(self, snapshots, nsnap: int, snap_dt: float, t0: float, nsteps: int):
        _abi.call("ns_resim_advance", self.h, _ptr(snapshots), int(nsnap), float(snap_dt), float(t0), int(nsteps))

    def resim_advance_spline(self, snapshots, nsnap: int, snap_dt: float, t0: float, nsteps: int):
        _abi.call("ns_resim_advance_spline", self.h, _ptr(snapshots), int(nsnap), float(snap_dt), float(t0),
                  int(nsteps))

    def bc_interp_spline(self, snapshots, nsnap: int, snap_dt: float, t: float):
        _abi.call("ns_bc_interp_spline", self.h, _ptr(snapshots), int(nsnap), float(snap_dt), float(t))

    def set_dt(self, dt: float):
        _abi.call("ns_set_dt", self.h, float(dt))
        self.cfg = self.cfg.replace(dt=float(dt))

    def history_save(self):
        _abi.call("ns_history_save", self.h)

    def history_restore(self):
        _abi.call("ns_history_restore", self.h)

    def extract_subbox(self, x0: int, y0: int, z0: int, nx0: int, ny0: int, nz0: int, field: str) -> np.ndarray:
        """sub-box of a resident field (ghosts included), e.g. u_sub of the parent"""
        comp = {"u": "u", "v": "v", "w": "w", "p": "p", "u_star": "u", "v_star": "v", "w_star": "w", "dp": "p",
                "hx": "u", "hy": "v", "hz": "w"}[field]
        shape = field_shapes(nx0, ny0, nz0)[comp]
        out = fzeros(shape)
        sb = _abi.SubBox(x0, y0, z0, nx0, ny0, nz0)
        _abi.call("ns_extract_subbox", self.h, ctypes.byref(sb), _abi.FIELDS[field], _ptr(out))
        return out

    def max_div(self) -> float:
        out = ctypes.c_double()
        _abi.call("ns_max_div", self.h, ctypes.byref(out))
        return out.value

    def error_metrics(self, u_sub, v_sub, w_sub, p_sub):
        """(err[8], rms[4]) of the current state against the parent's sub-box
        fields, as the reference's per-step error table (see metrics.py)."""
        err = (ctypes.c_double * 8)()
        rms = (ctypes.c_double * 4)()
        _abi.call("ns_error_metrics", self.h, _ptr(u_sub), _ptr(v_sub), _ptr(w_sub), _ptr(p_sub), err, rms)
        return np.array(err[:]), np.array(rms[:])

    def check_finite(self) -> bool:
        rc = _abi.lib().ns_check_finite(self.h)
        if rc == _abi.NS_OK:
            return True
        if rc == _abi.NS_ERR_NONFINITE:
            return False
        raise _abi.NSError("ns_check_finite", rc)

    # ---- mirrors (NS_functions.f90 names) ----
    def _conv_shapes(self, bc):
        c = self.cfg
        m = [c.nx, c.ny, c.nz]
        out = []
        for k in range(3):
            s = list(m)
            if bc[k] != 1:
                s[k] -= 1
            out.append(tuple(s))
        return out

    def cal_conv(self, u, v, w, bc_x, bc_y, bc_z, dx, dy, dz):
        sx, sy, sz = self._conv_shapes((bc_x, bc_y, bc_z))
        cx, cy, cz = fzeros(sx), fzeros(sy), fzeros(sz)
        _abi.call("ns_cal_conv", self.h, _ptr(u), _ptr(v), _ptr(w), bc_x, bc_y, bc_z, dx, dy, dz, _ptr(cx), _ptr(cy),
                  _ptr(cz))
        return cx, cy, cz

    def cal_diff(self, u, v, w, bc_x, bc_y, bc_z, dx2, dy2, dz2):
        sx, sy, sz = self._conv_shapes((bc_x, bc_y, bc_z))
        cx, cy, cz = fzeros(sx), fzeros(sy), fzeros(sz)
        _abi.call("ns_cal_diff", self.h, _ptr(u), _ptr(v), _ptr(w), bc_x, bc_y, bc_z, dx2, dy2, dz2, _ptr(cx),
                  _ptr(cy), _ptr(cz))
        return cx, cy, cz

    def vel_bc_staggered(self, u_star, v_star, w_star, planes, bc_x, bc_y, bc_z):
        _abi.call("ns_vel_bc_staggered", self.h, _ptr(u_star), _ptr(v_star), _ptr(w_star), _ptr(planes), bc_x, bc_y,
                  bc_z)

    def vel_bc_staggered_CN2(self, u_star, v_star, w_star, planes, bc, direction):
        _abi.call("ns_vel_bc_staggered_cn2", self.h, _ptr(u_star), _ptr(v_star), _ptr(w_star), _ptr(planes), bc,
                  direction)

    def adi_sweep(self, u_star, v_star, w_star, direction):
        _abi.call("ns_adi_sweep", self.h, _ptr(u_star), _ptr(v_star), _ptr(w_star), direction)

    def get_velpr_bc(self, u_bc, v_bc, w_bc, p_bc, dt, dp_flag, using_ustar, p_previous=None):
        planes = np.zeros(self.nplanes)
        _abi.call("ns_get_velpr_bc", self.h, _ptr(u_bc), _ptr(v_bc), _ptr(w_bc), _ptr(p_bc), float(dt), int(dp_flag),
                  int(using_ustar), _ptr(p_previous), _ptr(planes))
        return planes

    def get_pr_bc(self, p_bc, planes=None):
        if planes is None:
            planes = np.zeros(self.nplanes)
        _abi.call("ns_get_pr_bc", self.h, _ptr(p_bc), _ptr(planes))
        return planes

    def divergence(self, u_star, v_star, w_star, dt):
        c = self.cfg
        rhs = fzeros((c.nx, c.ny, c.nz))
        _abi.call("ns_divergence", self.h, _ptr(u_star), _ptr(v_star), _ptr(w_star), float(dt), _ptr(rhs))
        return rhs

    def pr_bc_staggered_modify_rhs(self, rhs, planes):
        _abi.call("ns_pr_bc_staggered_modify_rhs", self.h, _ptr(rhs), _ptr(planes))

    def poisson_fft(self, rhs):
        sol = fzeros(rhs.shape)
        _abi.call("ns_poisson_fft", self.h, _ptr(np.asfortranarray(rhs)), _ptr(sol))
        return sol

    def DCT_poisson_solver(self, rhs, mean_dp):
        out = np.array(rhs, dtype=np.float64, order="F", copy=True)
        _abi.call("ns_dct_poisson_solver", self.h, _ptr(out), float(mean_dp))
        return out

    def pr_bc_staggered(self, dp, planes=None):
        _abi.call("ns_pr_bc_staggered", self.h, _ptr(dp), _ptr(planes))

    def project(self, u_star, v_star, w_star, dp, dt, dp_flag, p):
        s = self.shapes
        u, v, w = fzeros(s["u"]), fzeros(s["v"]), fzeros(s["w"])
        p = np.array(p, dtype=np.float64, order="F", copy=True)
        _abi.call("ns_project", self.h, _ptr(u_star), _ptr(v_star), _ptr(w_star), _ptr(dp), float(dt), int(dp_flag),
                  _ptr(u), _ptr(v), _ptr(w), _ptr(p))
        return u, v, w, p
