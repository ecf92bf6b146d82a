"""Host logic of the re-simulation drivers (no GPU): the spline weights of
libns_hip (ns_spline_weights) against scipy's not-a-knot CubicSpline (the
published algorithm MATLAB's interp1 'spline' uses) on the golden vectors
and random windows, exactness on cubics, the window rule and the sub-step
time schedule."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from ns_3d_staggered_dp_amd import resim
from ns_3d_staggered_dp_amd.resim import spline_weights


def test_weights_reproduce_golden_vectors(golden):
    t0, y, tt, out = golden["spl_t0"], golden["spl_y"], golden["spl_tt"], golden["spl_out"]
    for i, x in enumerate(tt):
        got = spline_weights(t0, x) @ y
        assert np.allclose(got, out[i], rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("m", [4, 5, 6, 7])
def test_weights_match_scipy_random(m):
    rng = np.random.default_rng(m)
    for _ in range(20):
        t = np.cumsum(rng.uniform(0.2, 1.5, m)) - 3.0
        y = rng.standard_normal((m, 3))
        cs = CubicSpline(t, y, bc_type="not-a-knot", axis=0)
        for x in np.concatenate([rng.uniform(t[0] - 0.3, t[-1] + 0.3, 5), t]):
            assert np.allclose(spline_weights(t, x) @ y, cs(x), rtol=1e-12, atol=1e-12)


def test_weights_exact_on_cubics_and_low_orders():
    t = np.array([0.0, 0.3, 0.7, 1.2, 1.6, 2.5])
    for x in (0.1, 0.95, 2.0, 2.8):
        w = spline_weights(t, x)
        assert abs(w.sum() - 1.0) < 1e-14
        for k in range(4):  # not-a-knot reproduces every cubic exactly
            assert abs(w @ t ** k - x ** k) < 1e-12 * max(1.0, abs(x) ** k)
    w3 = spline_weights(t[:3], 0.5)
    assert np.allclose(w3 @ t[:3] ** 2, 0.25, atol=1e-15)  # parabola through 3 points
    assert np.allclose(spline_weights(t[:2], 0.15), [0.5, 0.5], atol=1e-15)


def test_window_rule():
    # six snapshots, three at or before t, clamped into the history
    assert resim.spline_window(20, 0.1, 0.55) == (3, 6)
    assert resim.spline_window(20, 0.1, 0.05) == (0, 6)
    assert resim.spline_window(20, 0.1, 1.9) == (14, 6)
    assert resim.spline_window(4, 0.1, 0.25) == (0, 4)


def test_sub_tstep_times():
    ts = resim.sub_tstep_times(1.0, 0.004, 4, 3)
    assert np.allclose(ts, [1.001, 1.002, 1.003, 1.004, 1.008, 1.012, 1.016], atol=1e-15)
