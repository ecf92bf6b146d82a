"""GPU: ns_error_metrics (one device pass + host combination of block
partials) against the literal numpy statement of the reference's metrics
(metrics.reference_metrics).  The maxima are exact; sums differ only in
summation order (relative 1e-13)."""
import numpy as np
import pytest

from ns_3d_staggered_dp_amd import AB2, NSConfig, fields, metrics
from ns_3d_staggered_dp_amd.solver import NSSolver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("device_inputs", [False, True])
def test_error_metrics_match_numpy(device_inputs):
    import torch

    cfg = NSConfig(nx=40, ny=12, nz=18, dx=0.31, dy=0.27, dz=0.23, nu=0.01, dt=1e-3, bc=(2, 2, 2), pbc=(3, 3, 3),
                   dp_flag=1)
    u, v, w, p = fields.random_ghosted(cfg, 91)
    us, vs, ws, ps = (x + 1e-3 * y for x, y in zip((u, v, w, p), fields.random_ghosted(cfg, 92)))
    with NSSolver(cfg) as s:
        s.upload(u, v, w, p, first=True)
        args = (us, vs, ws, ps)
        if device_inputs:
            args = tuple(torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).cuda() for a in args)
        err, rms = s.error_metrics(*args)
    err_ref, rms_ref = metrics.reference_metrics(u, v, w, p, us, vs, ws, ps, cfg.dx, cfg.dy)
    assert np.allclose(rms, rms_ref, rtol=1e-13, atol=0)
    assert np.allclose(err, err_ref, rtol=1e-13, atol=0)


def test_error_metrics_after_steps_periodic():
    """a lazily projected periodic run: the metrics see the materialised u, v, w, p"""
    cfg = NSConfig(nx=64, ny=16, nz=24, dx=2 * np.pi / 64, dy=0.5, dz=0.35, nu=2e-3, dt=2e-3, scheme=AB2)
    u, v, w, p = fields.decaying_turbulence(cfg, seed=11)
    with NSSolver(cfg) as s:
        s.upload(u, v, w, p, first=True)
        s.step(1)
        ref = s.download()
        s.step(2)
        st = s.download()
        err, rms = s.error_metrics(ref["u"], ref["v"], ref["w"], ref["p"])
    err_ref, rms_ref = metrics.reference_metrics(st["u"], st["v"], st["w"], st["p"], ref["u"], ref["v"], ref["w"],
                                                 ref["p"], cfg.dx, cfg.dy)
    assert np.allclose(rms, rms_ref, rtol=1e-13, atol=0)
    assert np.allclose(err, err_ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("case", ["clean", "nan_in_u", "inf_in_p", "nan_in_history", "lazy_periodic_nan"])
def test_check_finite_reports_injected_nonfinite(case):
    """ns_check_finite returns NS_ERR_NONFINITE for a state holding a NaN or
    an Inf, uploaded (or carried through steps that spread it), and NS_OK
    for a clean one"""
    from ns_3d_staggered_dp_amd import FLAG_GAUGE_CORNER, _abi

    if case == "lazy_periodic_nan":
        cfg = NSConfig(nx=32, ny=16, nz=12, dx=2 * np.pi / 32, dy=0.4, dz=0.5, nu=2e-3, dt=2e-3, scheme=AB2,
                       flags=FLAG_GAUGE_CORNER)
        u, v, w, p = fields.decaying_turbulence(cfg, seed=12)
    else:
        cfg = NSConfig(nx=20, ny=12, nz=10, dx=0.31, dy=0.27, dz=0.23, nu=0.01, dt=1e-3, scheme=AB2, bc=(2, 2, 2),
                       pbc=(3, 3, 3), dp_flag=1)
        u, v, w, p = fields.random_ghosted(cfg, 93)
        u, v, w = 0.3 * u, 0.3 * v, 0.3 * w
    hx = np.zeros_like(u)
    if case in ("nan_in_u", "lazy_periodic_nan"):
        u[5, 3, 4] = np.nan
    elif case == "inf_in_p":
        p[2, 2, 2] = np.inf
    elif case == "nan_in_history":
        hx[6, 5, 4] = np.nan
    want = _abi.NS_OK if case == "clean" else _abi.NS_ERR_NONFINITE
    with NSSolver(cfg) as s:
        s.upload(u, v, w, p, hx, None, None, first=False)
        # the history enters the state only through a step
        if case != "nan_in_history":
            assert _abi.lib().ns_check_finite(s.h) == want
        s.step(2)
        assert _abi.lib().ns_check_finite(s.h) == want
