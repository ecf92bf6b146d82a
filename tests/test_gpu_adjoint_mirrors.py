"""GPU: each adjoint mirror (ns_*_adj) is the transpose of its forward mirror.

For a linear map F and random x, y the identity <y, F x> = <F^T y, x> holds
exactly in exact arithmetic; in FP64 both sides agree to rounding, so the bar
is 1e-12 relative to the sum of the magnitudes of the products.  Every mirror
is checked on walls (DCT, DST and mixed codes, AB2-CN for the sweeps and the
CN boundary rows) and, where it applies, on a periodic box; the sweeps and CN
rows also on periodic lines (the reduced cyclic solve, with the reference's
row-1 coupling and with the cyclic operator)."""
import ctypes

import numpy as np
import pytest

from ns_3d_staggered_dp_amd import AB2, AB2_CN, FLAG_CN_CYCLIC, NSConfig, _abi
from ns_3d_staggered_dp_amd.grid import field_shapes, plane_set_length
from ns_3d_staggered_dp_amd.solver import NSSolver, _ptr

pytestmark = pytest.mark.gpu


def box(nx, ny, nz, **kw):
    return NSConfig(nx=nx, ny=ny, nz=nz, dx=0.31, dy=0.27, dz=0.23, nu=0.05, dt=2e-3, dp_flag=1, **kw)


WALLS = [
    ("dct", lambda: box(12, 10, 8, bc=(2, 2, 2), pbc=(3, 3, 3), scheme=AB2_CN)),
    ("dst", lambda: box(10, 9, 7, bc=(4, 4, 4), pbc=(2, 2, 2), scheme=AB2_CN)),
    ("mixed", lambda: box(10, 8, 9, bc=(2, 4, 2), pbc=(3, 2, 3), scheme=AB2_CN)),
    ("long_lines", lambda: box(40, 6, 36, bc=(4, 2, 2), pbc=(3, 3, 3), scheme=AB2_CN)),
]
PERIODIC = lambda: NSConfig(nx=16, ny=12, nz=10, dx=2 * np.pi / 16, dy=0.4, dz=0.5, nu=0.02, dt=2e-3, bc=(1, 1, 1),
                            pbc=(1, 1, 1), scheme=AB2, dp_flag=1)
# AB2-CN sweeps with periodic lines: the reference's row-1 coupling, the cyclic operator, periodic x beside walls
CN_PERIODIC = [
    ("periodic_cn", lambda: PERIODIC().replace(scheme=AB2_CN)),
    ("periodic_cn_cyclic", lambda: PERIODIC().replace(scheme=AB2_CN, flags=FLAG_CN_CYCLIC)),
    ("cn_periodic_x", lambda: box(12, 10, 8, bc=(1, 2, 4), pbc=(3, 3, 2), scheme=AB2_CN)),
]


def rnd(rng, shape):
    return np.asfortranarray(rng.standard_normal(shape))


def check(lhs_terms, rhs_terms):
    lhs = sum(float(np.sum(a * b)) for a, b in lhs_terms)
    rhs = sum(float(np.sum(a * b)) for a, b in rhs_terms)
    scale = sum(float(np.sum(np.abs(a * b))) for a, b in lhs_terms + rhs_terms)
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)


def vel(cfg, rng):
    sh = field_shapes(cfg.nx, cfg.ny, cfg.nz)
    return [rnd(rng, sh[k]) for k in ("u", "v", "w")]


@pytest.mark.parametrize("name,mk", WALLS + [("periodic", PERIODIC)], ids=[c[0] for c in WALLS] + ["periodic"])
def test_velocity_fill_adjoint(name, mk):
    cfg = mk()
    rng = np.random.default_rng(1)
    L = plane_set_length(cfg.nx, cfg.ny, cfg.nz)
    x, planes, y = vel(cfg, rng), rng.standard_normal(L), vel(cfg, rng)
    with NSSolver(cfg) as s:
        fx = [a.copy(order="F") for a in x]
        _abi.call("ns_vel_bc_staggered", s.h, *(_ptr(a) for a in fx), _ptr(planes), *cfg.bc)
        yb, pb = [a.copy(order="F") for a in y], np.zeros(L)
        _abi.call("ns_vel_bc_staggered_adj", s.h, *(_ptr(a) for a in yb), _ptr(pb), *cfg.bc)
    check(list(zip(y, fx)), list(zip(yb, x)) + [(pb, planes)])


@pytest.mark.parametrize("name,mk", WALLS + CN_PERIODIC, ids=[c[0] for c in WALLS + CN_PERIODIC])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_cn_rows_and_sweep_adjoint(name, mk, d):
    cfg = mk()
    rng = np.random.default_rng(2 + d)
    L = plane_set_length(cfg.nx, cfg.ny, cfg.nz)
    x, planes, y = vel(cfg, rng), rng.standard_normal(L), vel(cfg, rng)
    with NSSolver(cfg) as s:
        fx = [a.copy(order="F") for a in x]
        _abi.call("ns_vel_bc_staggered_cn2", s.h, *(_ptr(a) for a in fx), _ptr(planes), cfg.bc[d - 1], d)
        yb, pb = [a.copy(order="F") for a in y], np.zeros(L)
        _abi.call("ns_vel_bc_staggered_cn2_adj", s.h, *(_ptr(a) for a in yb), _ptr(pb), cfg.bc[d - 1], d)
        check(list(zip(y, fx)), list(zip(yb, x)) + [(pb, planes)])
        # the sweep (a linear solve per line) and the solve with the transposed systems
        fx = [a.copy(order="F") for a in x]
        _abi.call("ns_adi_sweep", s.h, *(_ptr(a) for a in fx), d)
        yb = [a.copy(order="F") for a in y]
        _abi.call("ns_adi_sweep_adj", s.h, *(_ptr(a) for a in yb), d)
    check(list(zip(y, fx)), list(zip(yb, x)))


@pytest.mark.parametrize("name,mk", WALLS, ids=[c[0] for c in WALLS])
def test_divergence_rhs_terms_and_dct_solve_adjoint(name, mk):
    cfg = mk()
    rng = np.random.default_rng(7)
    n = (cfg.nx, cfg.ny, cfg.nz)
    L = plane_set_length(*n)
    x, y = vel(cfg, rng), rnd(rng, n)
    with NSSolver(cfg) as s:
        rhs = np.zeros(n, order="F")
        _abi.call("ns_divergence", s.h, *(_ptr(a) for a in x), cfg.dt, _ptr(rhs))
        ub = [np.zeros_like(a, order="F") for a in x]
        _abi.call("ns_divergence_adj", s.h, _ptr(y), cfg.dt, *(_ptr(a) for a in ub))
        check([(y, rhs)], list(zip(ub, x)))
        # boundary terms: rhs -> rhs + M planes
        r0, planes = rnd(rng, n), rng.standard_normal(L)
        r1 = r0.copy(order="F")
        _abi.call("ns_pr_bc_staggered_modify_rhs", s.h, _ptr(r1), _ptr(planes))
        pb = np.zeros(L)
        _abi.call("ns_pr_bc_staggered_modify_rhs_adj", s.h, _ptr(y), _ptr(pb))
        check([(y, r1 - r0)], [(pb, planes)])
        # the transform solve and mean_dp
        r, mean_dp = rnd(rng, n), 0.37
        sol = r.copy(order="F")
        _abi.call("ns_dct_poisson_solver", s.h, _ptr(sol), mean_dp)
        yb, mb = y.copy(order="F"), ctypes.c_double(0.0)
        _abi.call("ns_dct_poisson_solver_adj", s.h, _ptr(yb), ctypes.byref(mb))
    check([(y, sol)], [(yb, r), (np.array(mb.value), np.array(mean_dp))])


def test_periodic_solve_adjoint():
    cfg = PERIODIC()
    rng = np.random.default_rng(9)
    n = (cfg.nx, cfg.ny, cfg.nz)
    r, y = rnd(rng, n), rnd(rng, n)
    with NSSolver(cfg) as s:
        sol, rb = np.zeros(n, order="F"), np.zeros(n, order="F")
        _abi.call("ns_poisson_fft", s.h, _ptr(r), _ptr(sol))
        _abi.call("ns_poisson_fft_adj", s.h, _ptr(y), _ptr(rb))
    check([(y, sol)], [(rb, r)])


@pytest.mark.parametrize("name,mk", WALLS + [("periodic", PERIODIC)], ids=[c[0] for c in WALLS] + ["periodic"])
@pytest.mark.parametrize("dp_flag", [0, 1])
def test_pressure_fill_and_projection_adjoint(name, mk, dp_flag):
    cfg = mk()
    rng = np.random.default_rng(11 + dp_flag)
    sh = field_shapes(cfg.nx, cfg.ny, cfg.nz)
    L = plane_set_length(cfg.nx, cfg.ny, cfg.nz)
    dp, planes, ydp = rnd(rng, sh["p"]), rng.standard_normal(L), rnd(rng, sh["p"])
    with NSSolver(cfg) as s:
        f = dp.copy(order="F")
        _abi.call("ns_pr_bc_staggered", s.h, _ptr(f), _ptr(planes))
        yb, pb = ydp.copy(order="F"), np.zeros(L)
        _abi.call("ns_pr_bc_staggered_adj", s.h, _ptr(yb), _ptr(pb))
        check([(ydp, f)], [(yb, dp), (pb, planes)])
        # projection: (u*, v*, w*, dp, p) -> (u, v, w, p)
        us, p_in = vel(cfg, rng), rnd(rng, sh["p"])
        yu, yp = vel(cfg, rng), rnd(rng, sh["p"])
        out = [np.zeros_like(a, order="F") for a in us]
        p_out = p_in.copy(order="F")
        _abi.call("ns_project", s.h, *(_ptr(a) for a in us), _ptr(dp), cfg.dt, dp_flag, *(_ptr(a) for a in out),
                  _ptr(p_out))
        usb = [np.zeros_like(a, order="F") for a in us]
        dpb, ypb = np.zeros(sh["p"], order="F"), yp.copy(order="F")
        _abi.call("ns_project_adj", s.h, *(_ptr(a) for a in yu), _ptr(ypb), cfg.dt, dp_flag, *(_ptr(a) for a in usb),
                  _ptr(dpb))
    check(list(zip(yu, out)) + [(yp, p_out)], list(zip(usb, us)) + [(dpb, dp), (ypb, p_in)])


def test_adjoint_mirrors_reject_unsupported():
    # the sweep's adjoint needs the CN systems of an AB2-CN context
    with NSSolver(PERIODIC()) as s:
        sh = field_shapes(16, 12, 10)
        a = [np.zeros(sh[k], order="F") for k in ("u", "v", "w")]
        assert _abi.lib().ns_adi_sweep_adj(s.h, *(_ptr(x) for x in a), 1) == _abi.NS_ERR_UNSUPPORTED
        assert _abi.lib().ns_adi_sweep_adj(s.h, *(_ptr(x) for x in a), 4) == _abi.NS_ERR_INVALID
