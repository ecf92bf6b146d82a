"""GPU: ns_error_metrics_noalloc (block partials in a static device array,
combined on the device) against ns_error_metrics (library-allocated partials,
host combination in extended precision) and the numpy statement of the
reference's metrics (re_simulation_sub_tstep_interBC_combine.f90:826-833).

Two calls in a row on one GPU, the second on a smaller box: the first box
needs more partial rows than the scratch holds (grid-stride), the second
far fewer, so a combine that read rows left from the first call would fail.
The maxima are exact; the sums differ from the other two only in summation
order, hence the relative tolerance of 1e-13.
"""
import ctypes

import numpy as np
import pytest

from ns_3d_staggered_dp_amd import NSConfig, _abi, fields, metrics
from ns_3d_staggered_dp_amd.solver import NSSolver

pytestmark = pytest.mark.gpu

RTOL = 1e-13


def _case(nx, ny, nz, seed):
    cfg = NSConfig(nx=nx, ny=ny, nz=nz, dx=0.31, dy=0.27, dz=0.23, nu=0.01, dt=1e-3, bc=(2, 2, 2), pbc=(3, 3, 3),
                   dp_flag=1)
    u, v, w, p = fields.random_ghosted(cfg, seed)
    sub = tuple(x + 1e-3 * y for x, y in zip((u, v, w, p), fields.random_ghosted(cfg, seed + 1)))
    return cfg, (u, v, w, p), sub


def _noalloc(s, dev_sub):
    err = (ctypes.c_double * 8)()
    rms = (ctypes.c_double * 4)()
    ptrs = [ctypes.c_void_p(t.data_ptr()) for t in dev_sub]
    _abi.call("ns_error_metrics_noalloc", s.h, *ptrs, err, rms)
    return np.array(err[:]), np.array(rms[:])


def test_noalloc_metrics_match_allocating_entry_point():
    import torch

    torch.cuda.synchronize()
    for nx, ny, nz, seed in ((128, 64, 80, 301), (20, 12, 18, 303)):
        print(f"noalloc metrics: box {nx}x{ny}x{nz} seed {seed}", flush=True)
        cfg, state, sub = _case(nx, ny, nz, seed)
        dev_sub = [torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).cuda() for a in sub]
        torch.cuda.synchronize()
        with NSSolver(cfg) as s:
            s.upload(*state, first=True)
            err_n, rms_n = _noalloc(s, dev_sub)
            err_a, rms_a = s.error_metrics(*dev_sub)
        err_ref, rms_ref = metrics.reference_metrics(*state, *sub, cfg.dx, cfg.dy)
        assert np.all(np.isfinite(err_n)) and np.all(rms_n > 0)
        assert np.allclose(rms_n, rms_a, rtol=RTOL, atol=0), (rms_n, rms_a)
        assert np.allclose(err_n, err_a, rtol=RTOL, atol=0), (err_n, err_a)
        assert np.allclose(rms_n, rms_ref, rtol=RTOL, atol=0)
        assert np.allclose(err_n, err_ref, rtol=RTOL, atol=0)


def test_noalloc_metrics_rejects_host_arrays():
    """host arrays would need staging (an allocation): refused before any launch"""
    cfg, state, sub = _case(12, 10, 8, 305)
    with NSSolver(cfg) as s:
        s.upload(*state, first=True)
        err = (ctypes.c_double * 8)()
        rms = (ctypes.c_double * 4)()
        ptrs = [ctypes.c_void_p(a.ctypes.data) for a in sub]
        rc = _abi.lib().ns_error_metrics_noalloc(s.h, *ptrs, err, rms)
    assert rc == _abi.NS_ERR_INVALID
