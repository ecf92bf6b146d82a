"""GPU parity: libns_hip.so (through the C ABI) against the CPU oracle on the
same seeded inputs.

Bars (written in each test):
- pure copies / affine boundary fills and the stencils: bitwise equal
  (same IEEE operations in the same order, no FMA contraction on either side);
- transform-based solves (FFT/DCT, ADI): relative error <= 1e-12 per call;
- whole steps: max relative error of u, v, w, p <= 1e-11 after N steps and
  max |div u| <= 1e-12 (the tolerance of the FP64 target).
"""
import numpy as np
import pytest

from ns_3d_staggered_dp_amd import AB2, AB2_CN, EULER, FLAG_CN_CYCLIC, FLAG_GAUGE_CORNER, NSConfig
from ns_3d_staggered_dp_amd import fields
from ns_3d_staggered_dp_amd.grid import field_shapes, plane_set_length
from ns_3d_staggered_dp_amd.solver import NSSolver, fzeros

pytestmark = pytest.mark.gpu


def rel(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)


def rand_fields(cfg, seed):
    return fields.random_ghosted(cfg, seed)


def rand_planes(cfg, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return scale * rng.standard_normal(plane_set_length(cfg.nx, cfg.ny, cfg.nz))


def cfg_box(nx, ny, nz, **kw):
    return NSConfig(nx=nx, ny=ny, nz=nz, dx=0.31, dy=0.27, dz=0.23, nu=0.01, dt=2e-3, **kw)


BCS = [(1, 1, 1), (2, 2, 2), (4, 4, 4), (1, 2, 4)]


@pytest.mark.parametrize("bc", BCS)
def test_cal_conv_cal_diff_bitwise(oracle, bc):
    cfg = cfg_box(10, 8, 6, bc=bc)
    u, v, w, _ = rand_fields(cfg, 11)
    with NSSolver(cfg.replace(bc=(1, 1, 1))) as s:
        g = s.cal_conv(u, v, w, *bc, cfg.dx, cfg.dy, cfg.dz)
        r = oracle.cal_conv(u, v, w, bc, cfg.dx, cfg.dy, cfg.dz)
        for a, b in zip(g, r):
            assert a.shape == b.shape
            assert np.array_equal(a, b)
        g = s.cal_diff(u, v, w, *bc, cfg.dx ** 2, cfg.dy ** 2, cfg.dz ** 2)
        r = oracle.cal_diff(u, v, w, bc, cfg.dx ** 2, cfg.dy ** 2, cfg.dz ** 2)
        for a, b in zip(g, r):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("bc", BCS)
def test_vel_bc_and_cn2_bitwise(oracle, bc):
    cfg = cfg_box(9, 7, 6, bc=bc)
    planes = rand_planes(cfg, 3)
    with NSSolver(cfg) as s:
        u, v, w, _ = rand_fields(cfg, 5)
        gu, gv, gw = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
        s.vel_bc_staggered(gu, gv, gw, planes, *bc)
        oracle.vel_bc(u, v, w, planes, bc)
        assert np.array_equal(gu, u) and np.array_equal(gv, v) and np.array_equal(gw, w)
        for d in (1, 2, 3):
            u, v, w, _ = rand_fields(cfg, 7 + d)
            gu, gv, gw = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
            s.vel_bc_staggered_CN2(gu, gv, gw, planes, bc[d - 1], d)
            oracle.vel_bc_cn2(u, v, w, planes, bc[d - 1], d)
            assert np.array_equal(gu, u) and np.array_equal(gv, v) and np.array_equal(gw, w)


@pytest.mark.parametrize("pbc", [(1, 1, 1), (2, 2, 2), (3, 3, 3), (4, 4, 4), (3, 2, 4)])
def test_pressure_bc_bitwise(oracle, pbc):
    cfg = cfg_box(9, 7, 6, pbc=pbc)
    with NSSolver(cfg) as s:
        _, _, _, p = rand_fields(cfg, 21)
        g = s.get_pr_bc(p)
        r = oracle.get_pr_bc(p, pbc, cfg.dx, cfg.dy, cfg.dz)
        assert np.array_equal(g, r)
        planes = rand_planes(cfg, 22)
        rhs = np.asfortranarray(np.random.default_rng(23).standard_normal((9, 7, 6)))
        a, b = rhs.copy(order="F"), rhs.copy(order="F")
        s.pr_bc_staggered_modify_rhs(a, planes)
        oracle.pr_bc_modify_rhs(b, planes, pbc, cfg.dx, cfg.dy, cfg.dz)
        assert np.array_equal(a, b)
        a, b = p.copy(order="F"), p.copy(order="F")
        s.pr_bc_staggered(a, planes)
        oracle.pr_bc(b, planes, pbc, cfg.dx, cfg.dy, cfg.dz)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("bc,using_ustar,dp_flag", [((2, 2, 2), 1, 1), ((4, 2, 4), 0, 0), ((2, 4, 2), 0, 1)])
def test_get_velpr_bc_bitwise(oracle, bc, using_ustar, dp_flag):
    pbc = (3, 3, 3)
    cfg = cfg_box(9, 7, 6, bc=bc, pbc=pbc)
    u, v, w, p = rand_fields(cfg, 31)
    pprev = rand_fields(cfg, 32)[3]
    with NSSolver(cfg) as s:
        g = s.get_velpr_bc(u, v, w, p, 1e-3, dp_flag, using_ustar, pprev)
    r = oracle.get_velpr_bc(u, v, w, p, bc, pbc, cfg.dx, cfg.dy, cfg.dz, 1e-3, dp_flag, using_ustar, pprev)
    assert np.array_equal(g, r)


def test_divergence_and_projection_bitwise(oracle):
    cfg = cfg_box(10, 8, 6)
    u, v, w, p = rand_fields(cfg, 41)
    dp = rand_fields(cfg, 42)[3]
    with NSSolver(cfg) as s:
        g = s.divergence(u, v, w, cfg.dt)
        assert np.array_equal(g, oracle.divergence(u, v, w, cfg.dx, cfg.dy, cfg.dz, cfg.dt))
        gu, gv, gw, gp = s.project(u, v, w, dp, cfg.dt, 1, p)
    ou, ov, ow, op = fzeros(u.shape), fzeros(v.shape), fzeros(w.shape), p.copy(order="F")
    oracle.lib().nso_project(*(oracle._p(x) for x in (ou, ov, ow, op, u, v, w, dp)), cfg.nx, cfg.ny, cfg.nz,
                             *(oracle.ctypes.c_double(x) for x in (cfg.dx, cfg.dy, cfg.dz, cfg.dt)), 1)
    for a, b in ((gu, ou), (gv, ov), (gw, ow), (gp, op)):
        assert np.array_equal(a, b)


# 64 and 256 long y/z lines take the register-resident two-step kernel, 512
# long ones the three-step kernel (ragged wavenumber groups: nx = 10, 42, 20, 6)
@pytest.mark.parametrize("n", [(16, 16, 16), (12, 10, 8), (32, 8, 24), (64, 64, 64), (10, 64, 256), (42, 256, 64),
                               (6, 256, 12), (20, 12, 256), (6, 512, 10), (20, 6, 512), (10, 512, 512), (2, 8, 6),
                               (4, 6, 8), (2, 256, 4)])
def test_poisson_fft(oracle, n):
    cfg = cfg_box(*n)
    rhs = np.asfortranarray(np.random.default_rng(51).standard_normal(n))
    with NSSolver(cfg) as s:
        g = s.poisson_fft(rhs)
    r = oracle.poisson_fft(rhs, cfg.dx, cfg.dy, cfg.dz)
    assert rel(g, r) <= 1e-12


# including 64- and 128-long z lines with ragged row groups (nx = 20, 12, 70, 15)
@pytest.mark.parametrize("pbc", [(3, 3, 3), (2, 2, 2), (2, 3, 2)])
@pytest.mark.parametrize("n", [(16, 16, 16), (12, 10, 14), (48, 20, 28), (70, 36, 20), (128, 24, 40), (20, 10, 128),
                               (12, 6, 64), (70, 8, 128), (15, 8, 64)])
def test_dct_poisson(oracle, pbc, n):
    cfg = cfg_box(*n, bc=(2, 2, 2), pbc=pbc)
    rhs = np.asfortranarray(np.random.default_rng(61).standard_normal(n))
    with NSSolver(cfg) as s:
        g = s.DCT_poisson_solver(rhs, 0.37)
    r = oracle.poisson_dct(rhs, pbc, cfg.dx, cfg.dy, cfg.dz, 0.37)
    assert rel(g, r) <= 1e-12


@pytest.mark.parametrize("bc,flags", [((1, 1, 1), 0), ((1, 1, 1), FLAG_CN_CYCLIC), ((2, 2, 2), 0), ((4, 4, 4), 0),
                                      ((2, 1, 4), 0)])
def test_adi_sweep(oracle, bc, flags):
    cfg = cfg_box(12, 10, 8, bc=bc, scheme=AB2_CN, flags=flags).replace(dt=0.05, nu=0.5)
    with NSSolver(cfg) as s:
        for d in (1, 2, 3):
            u, v, w, _ = rand_fields(cfg, 70 + d)
            gu, gv, gw = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
            s.adi_sweep(gu, gv, gw, d)
            oracle.adi_sweep(u, v, w, cfg, d)
            for a, b in ((gu, u), (gv, v), (gw, w)):
                assert rel(a, b) <= 1e-12, (d, rel(a, b))


def run_both(oracle, cfg, u, v, w, p, nsteps, planes=None, first=True, hist=None):
    st = oracle.OracleState(cfg, u, v, w, p, first=first, **(hist or {}))
    with NSSolver(cfg) as s:
        s.upload(u, v, w, p, *(None if hist is None else hist[k] for k in ("hx", "hy", "hz")), first=first)
        if planes is not None:
            s.set_bc_planes(planes)
        s.step(nsteps)
        out = s.download()
        div = s.max_div()
    for _ in range(nsteps):
        st.step(planes)
    return out, st, div


def test_step_periodic_ab2_turbulence(oracle):
    """decaying turbulence (big_periodic_3D path), AB2 + FFT, corner gauge"""
    cfg = NSConfig.cube(32, nu=1e-3, dt=5e-4 * 8, scheme=AB2, flags=FLAG_GAUGE_CORNER)
    u, v, w, p = fields.decaying_turbulence(cfg, seed=2)
    out, st, div = run_both(oracle, cfg, u, v, w, p, 20)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p), ("hx", st.hx)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    assert div <= 1e-12
    assert st.max_div() <= 1e-12


@pytest.mark.parametrize("flags", [0, FLAG_CN_CYCLIC])
def test_step_periodic_ab2cn_tgv(oracle, flags):
    """Taylor-Green (AB2-CN + FFT), reference periodic CN matrix or the cyclic one"""
    cfg = NSConfig.cube(24, nu=1 / 1600, dt=1e-2, scheme=AB2_CN, flags=FLAG_GAUGE_CORNER | flags)
    u, v, w, p = fields.taylor_green(cfg, seed=1)
    out, st, div = run_both(oracle, cfg, u, v, w, p, 8)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    assert div <= 1e-12


@pytest.mark.parametrize("scheme", [EULER, AB2, AB2_CN])
@pytest.mark.parametrize("bc,pbc", [((2, 2, 2), (3, 3, 3)), ((4, 4, 4), (2, 2, 2)), ((2, 4, 2), (3, 2, 3))])
def test_step_walls(oracle, scheme, bc, pbc):
    """wall boundaries with a fixed random plane set, DCT/DST Poisson, dp_flag=1"""
    cfg = cfg_box(12, 10, 14, bc=bc, pbc=pbc, scheme=scheme, dp_flag=1).replace(dt=1e-3)
    u, v, w, p = rand_fields(cfg, 81)
    u, v, w = 0.3 * u, 0.3 * v, 0.3 * w
    planes = rand_planes(cfg, 82, 0.2)
    out, st, div = run_both(oracle, cfg, u, v, w, p, 4, planes=planes)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))


@pytest.mark.parametrize("n", [(126, 12, 10), (12, 126, 10), (12, 10, 126), (250, 6, 5), (6, 250, 5), (6, 5, 250),
                               (320, 5, 4), (480, 4, 3), (4, 480, 3), (4, 3, 480), (700, 4, 3), (4, 700, 3), (5, 4, 700), (2048, 3, 2), (3, 2048, 2), (3, 2, 2048)],
                         ids=["long_x", "long_y", "long_z", "k9_x", "k9_y", "k9_z", "wide_x", "rows40_x", "rows40_y",
                              "rows40_z", "twisted_x", "twisted_y", "twisted_z", "max_x", "max_y", "max_z"])
def test_step_walls_cn_long_lines(oracle, n):
    """AB2-CN on walls with long lines along one axis: the partitioned sweeps
    with 5 blocks per line (126 cells) and 9 (250 cells; the x sweep's LDS
    staging of 64 lines), x lines too long for 64 of them in the LDS (320
    cells: 32 lines per block), lines over 448 rows in blocks of up to 40 rows
    (480 cells), the twisted solve of lines over 640 rows (700 cells; y and z
    lines too long for the LDS columns keep the elimination results in the
    line itself), lines of the longest transform length (2048 cells), and the
    DCT passes on long lines with odd radices, against the oracle after 3
    steps."""
    cfg = cfg_box(*n, bc=(2, 4, 2), pbc=(3, 2, 3), scheme=AB2_CN, dp_flag=1).replace(dt=1e-3)
    u, v, w, p = rand_fields(cfg, 91)
    u, v, w = 0.3 * u, 0.3 * v, 0.3 * w
    planes = rand_planes(cfg, 92, 0.2)
    out, st, div = run_both(oracle, cfg, u, v, w, p, 3, planes=planes)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))


def _smooth(lo, hi):
    out = []
    for n in range(lo, hi + 1):
        m = n
        for f in (2, 3, 5, 7):
            while m % f == 0:
                m //= f
        if m == 1:
            out.append(n)
    return out


def _random_shapes(seed, count):
    """seeded box shapes: one axis drawn from the 7-smooth lengths in
    [60, 720], the other two from [2, 30], the long axis placed at random"""
    rng = np.random.default_rng(seed)
    long_, short = _smooth(60, 720), _smooth(2, 30)
    shapes = []
    for _ in range(count):
        n = [int(rng.choice(short)), int(rng.choice(short)), int(rng.choice(short))]
        n[int(rng.integers(3))] = int(rng.choice(long_))
        shapes.append(tuple(n))
    return shapes


RANDOM_SHAPES = _random_shapes(2026, 10)


@pytest.mark.parametrize("n", RANDOM_SHAPES, ids=["x".join(map(str, n)) for n in RANDOM_SHAPES])
@pytest.mark.parametrize("kind", ["walls", "periodic"])
@pytest.mark.parametrize("scheme", [AB2_CN, AB2], ids=["ab2_cn", "ab2"])
def test_step_random_shapes(oracle, n, kind, scheme):
    """AB2-CN and AB2 steps on seeded box shapes with one long axis (the
    launch shapes of every sweep, fill, predictor tile and transform pass
    depend on the lengths: lines per block, blocks per line, rows per lane,
    LDS staging), walls with mixed codes and planes or triply periodic,
    against the oracle after 3 steps.
    The periodic boxes start from smooth solenoidal fields, as the other
    periodic tests: from white noise the pressure of a 25 x 375 x 2 box came
    to 1.75e-11 of the oracle's (velocity within 1e-11), the rounding of the
    FFT solve on a box whose smallest and largest eigenvalues lie 1e5 apart."""
    if kind == "walls":
        cfg = cfg_box(*n, bc=(2, 4, 2), pbc=(3, 2, 3), scheme=scheme, dp_flag=1).replace(dt=1e-3)
        planes = rand_planes(cfg, 97, 0.2)
        u, v, w, p = rand_fields(cfg, 96)
        u, v, w = 0.3 * u, 0.3 * v, 0.3 * w
    else:
        cfg = cfg_periodic(*n, scheme=scheme, dp_flag=0)
        planes = None
        u, v, w, p = fields.decaying_turbulence(cfg, seed=96)
    out, st, div = run_both(oracle, cfg, u, v, w, p, 3, planes=planes)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    assert div <= 1e-12


@pytest.mark.parametrize("n", [64, 96, 128])
@pytest.mark.parametrize("scheme", [AB2_CN, AB2], ids=["ab2_cn", "ab2"])
def test_step_walls_cube(oracle, n, scheme):
    """re-simulation boxes at their working sizes (n^3 walls, DCT Poisson,
    dp_flag 1): every axis long enough for several partition blocks, predictor
    tiles, transform groups and adjoint chunks at once, against the oracle
    after 3 steps.  The random planes are shifted to carry no net flux
    through the walls (plane_sets.zero_flux), so the Neumann problem is
    solvable and an explicit step leaves max |div| <= 1e-12.  Under AB2-CN
    the y and z sweeps also move the wall-normal faces (they lie on lines of
    those sweeps), so the flux is not exactly the planes' and the divergence
    is its uniform mean: it must be the same in every cell to 1e-12."""
    import plane_sets

    cfg = NSConfig.cube(n, nu=1e-3, dt=5e-4, scheme=scheme, bc=(2, 2, 2), pbc=(3, 3, 3), dp_flag=1)
    u, v, w, p = fields.decaying_turbulence(cfg, seed=21)
    planes = plane_sets.zero_flux(cfg, rand_planes(cfg, 22, 0.05))
    out, st, div = run_both(oracle, cfg, u, v, w, p, 3, planes=planes)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    d = plane_sets.divergence(cfg, out["u"], out["v"], out["w"])
    assert np.abs(d - d.mean()).max() <= 1e-12
    if scheme == AB2:
        assert div <= 1e-12, div


def test_step_walls_256_component_streams(oracle):
    """boxes of 2^24 cells or more take the three velocity components' ADI
    sweeps on three streams (fork and join events on the context stream, also
    inside the graph that replays steady steps): a 256^3 wall box, AB2-CN,
    against the oracle after 3 steps (an eager first step, then a two-step
    graph), with the uniform-divergence check of the cube test"""
    import plane_sets

    cfg = NSConfig.cube(256, nu=1e-3, dt=5e-4, scheme=AB2_CN, bc=(2, 2, 2), pbc=(3, 3, 3), dp_flag=1)
    u, v, w, p = fields.decaying_turbulence(cfg, seed=23)
    planes = plane_sets.zero_flux(cfg, rand_planes(cfg, 24, 0.05))
    out, st, div = run_both(oracle, cfg, u, v, w, p, 3, planes=planes)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    d = plane_sets.divergence(cfg, out["u"], out["v"], out["w"])
    assert np.abs(d - d.mean()).max() <= 1e-12


def cfg_periodic(nx, ny, nz, **kw):
    return NSConfig(nx=nx, ny=ny, nz=nz, dx=2 * np.pi / nx, dy=0.5, dz=0.35, nu=2e-3, dt=2e-3, bc=(1, 1, 1),
                    pbc=(1, 1, 1), flags=FLAG_GAUGE_CORNER, **kw)


@pytest.mark.parametrize("n", [(320, 4, 4), (480, 4, 4), (6, 4, 480), (700, 4, 4), (6, 4, 700), (2048, 3, 2),
                               (4, 3, 2048)],
                         ids=["wide_x", "rows40_x", "rows40_z", "cyclic_x", "cyclic_z", "max_x", "max_z"])
def test_step_periodic_cn_long_lines(oracle, n):
    """AB2-CN on long periodic lines: the partitioned cyclic solve with x
    lines staged 32 per block (320 unknowns) and with blocks of up to 40 rows
    (480), and beyond it (700, and 2048, the longest transform length) the
    one-lane reduced cyclic solve with its Sherman-Morrison correction (the
    separate periodic boundary-row launch), against the oracle after 3 steps"""
    cfg = cfg_periodic(*n, scheme=AB2_CN, dp_flag=0)
    u, v, w, p = fields.decaying_turbulence(cfg, seed=13)
    out, st, div = run_both(oracle, cfg, u, v, w, p, 3)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    assert div <= 1e-12


@pytest.mark.parametrize("scheme,dp_flag", [(AB2, 0), (AB2_CN, 0), (AB2, 1), (EULER, 0)])
def test_step_periodic_multi_tile(oracle, scheme, dp_flag):
    """periodic box spanning several predictor tiles (64 x 4 x 16 cells) with
    partial tiles on every axis: the periodic wraps of u(nx+1), v(ny+1) and
    w(nz+1) and of the lazily applied pressure gradient cross tile edges"""
    cfg = cfg_periodic(72, 10, 20, scheme=scheme, dp_flag=dp_flag)
    u, v, w, p = fields.decaying_turbulence(cfg, seed=7)
    out, st, div = run_both(oracle, cfg, u, v, w, p, 6)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p), ("hx", st.hx), ("hz", st.hz)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    assert div <= 1e-12


def test_step_periodic_unfilled_ghosts(oracle):
    """periodic run from arrays whose ghost faces are not periodic copies
    (u(nx+1) != u(1) etc.): the first step's x/y/z averages read those ghost
    values exactly as avg(bc=1) does (FD_functions.f90:246-301)"""
    cfg = cfg_periodic(70, 9, 18, scheme=AB2, dp_flag=0)
    u, v, w, p = rand_fields(cfg, 83)
    out, st, div = run_both(oracle, cfg, 0.3 * u, 0.3 * v, 0.3 * w, p, 3)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    assert div <= 1e-12


@pytest.mark.parametrize("scheme", [AB2, AB2_CN])
def test_step_walls_lazy_multi_tile(oracle, scheme):
    """walls with dp_flag=0 (projection folded into the next predictor) on a
    box spanning several predictor tiles"""
    cfg = cfg_box(70, 9, 18, bc=(2, 2, 2), pbc=(3, 3, 3), scheme=scheme, dp_flag=0).replace(dt=1e-3)
    u, v, w, p = rand_fields(cfg, 84)
    u, v, w = 0.3 * u, 0.3 * v, 0.3 * w
    planes = rand_planes(cfg, 85, 0.2)
    out, st, div = run_both(oracle, cfg, u, v, w, p, 4, planes=planes)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))


def periodic_ghosts_ok(a, own):
    """every ghost entry of a ghosted component equals its interior periodic
    image (own axis: entry n is entry 1; other axes: 1 is n-1, n is 2)"""
    for ax in range(3):
        b = np.moveaxis(a, ax, 0)
        if ax == own:
            if not np.array_equal(b[-1], b[0]):
                return False
        elif not (np.array_equal(b[0], b[-2]) and np.array_equal(b[-1], b[1])):
            return False
    return True


@pytest.mark.parametrize("n", [(256, 8, 6), (512, 4, 6), (256, 12, 5)])
def test_periodic_rows_of_256_and_512(oracle, n):
    """x forward pass over rows of 256 and 512 cells (row groups of one plane,
    u staged in LDS, the next row's v from the same thread, the group's last
    row reading the row after it, also across the periodic end): lazy AB2
    (ghosts read through their periodic images) and AB2-CN (filled ghosts)"""
    for scheme in (AB2, AB2_CN):
        cfg = cfg_periodic(*n, scheme=scheme, dp_flag=0)
        u, v, w, p = fields.decaying_turbulence(cfg, seed=17)
        out, st, div = run_both(oracle, cfg, u, v, w, p, 4)
        for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
            assert rel(out[k], ref) <= 1e-11, (scheme, k, rel(out[k], ref))
        assert div <= 1e-12


def test_lazy_periodic_steps_skip_the_velocity_fill(oracle):
    """all-periodic lazy AB2 steps leave the ghost entries of u*, v*, w*
    unfilled and read them through their periodic images.  Against the oracle
    (which fills every step): from a fresh start and from a restart whose
    uploaded ghosts are random (the first predictor must read those as they
    are; later steps replay graphs), and the u*, v*, w* that ns_download_work
    returns have every ghost equal to its periodic image"""
    cfg = cfg_periodic(64, 16, 24, scheme=AB2, dp_flag=0)
    u, v, w, p = fields.decaying_turbulence(cfg, seed=13)
    out, st, div = run_both(oracle, cfg, u, v, w, p, 5)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    assert div <= 1e-12
    u2, v2, w2, p2 = (0.3 * x for x in rand_fields(cfg, 90))
    hist = {k: 0.1 * x for k, x in zip(("hx", "hy", "hz"), rand_fields(cfg, 91)[:3])}
    out, st, div = run_both(oracle, cfg, u2, v2, w2, p2, 6, first=False, hist=hist)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    with NSSolver(cfg) as s:
        s.upload(u, v, w, p, first=True)
        s.step(4)
        wk = s.download_work()
    for c, k in enumerate(("u_star", "v_star", "w_star")):
        assert periodic_ghosts_ok(wk[k], c), k


@pytest.mark.parametrize("scheme", [AB2, AB2_CN])
def test_walls_dpflag1_lazy_matches_eager_projection(oracle, scheme):
    """wall boxes with dp_flag = 1 fold the projection of u into the next
    predictor and form p where dp is formed (DCT x inverse pass, pressure
    fill).  Against the oracle, and against the same steps projected every
    step (FLAG_EAGER_PROJECTION): the lazy predictor's projection divides
    with the corrected reciprocal, so the two agree to rounding, not bitwise"""
    from ns_3d_staggered_dp_amd import FLAG_EAGER_PROJECTION

    cfg = cfg_box(70, 20, 18, bc=(2, 4, 2), pbc=(3, 2, 3), scheme=scheme, dp_flag=1).replace(dt=1e-3)
    u, v, w, p = (0.3 * x for x in rand_fields(cfg, 92))
    planes = rand_planes(cfg, 93, 0.2)
    out, st, div = run_both(oracle, cfg, u, v, w, p, 7, planes=planes)
    for k, ref in (("u", st.u), ("v", st.v), ("w", st.w), ("p", st.p)):
        assert rel(out[k], ref) <= 1e-11, (k, rel(out[k], ref))
    with NSSolver(cfg.replace(flags=FLAG_EAGER_PROJECTION)) as s:
        s.upload(u, v, w, p, first=True)
        s.set_bc_planes(planes)
        s.step(7)
        eager = s.download()
    for k in ("u", "v", "w", "p", "hx", "hy", "hz"):
        assert rel(out[k], eager[k]) <= 1e-13, (k, rel(out[k], eager[k]))


@pytest.mark.parametrize("case", ["periodic_ab2_lazy", "periodic_ab2cn", "walls_dct_dpflag"])
def test_graph_replay_matches_eager(case):
    """ns_step replays steady steps from a HIP graph (two steps per replay, one
    graph per parity of the lazy u* buffers).  The same call pattern with
    eager launches (FLAG_NO_GRAPH) must give bitwise the same state after
    every call; host downloads and odd step counts sit between the replays."""
    from ns_3d_staggered_dp_amd import FLAG_NO_GRAPH

    if case == "periodic_ab2_lazy":
        cfg = cfg_periodic(64, 16, 24, scheme=AB2, dp_flag=0)
        u, v, w, p = fields.decaying_turbulence(cfg, seed=9)
        planes = None
    elif case == "periodic_ab2cn":
        cfg = cfg_periodic(32, 20, 16, scheme=AB2_CN, dp_flag=0)
        u, v, w, p = fields.taylor_green(cfg, seed=3)
        planes = None
    else:
        cfg = cfg_box(40, 12, 18, bc=(2, 2, 2), pbc=(3, 3, 3), scheme=AB2, dp_flag=1).replace(dt=1e-3)
        u, v, w, p = rand_fields(cfg, 86)
        u, v, w = 0.3 * u, 0.3 * v, 0.3 * w
        planes = rand_planes(cfg, 87, 0.2)
    pattern = (1, 4, 3, 1, 2, 5)
    runs = []
    for flags in (cfg.flags | FLAG_NO_GRAPH, cfg.flags):
        c = cfg.replace(flags=flags)
        states = []
        with NSSolver(c) as s:
            s.upload(u, v, w, p, first=True)
            if planes is not None:
                s.set_bc_planes(planes)
            for n in pattern:
                s.step(n)
                states.append(s.download())
        runs.append(states)
    for a, b in zip(*runs):
        for k in ("u", "v", "w", "p", "hx", "hy", "hz"):
            assert np.array_equal(a[k], b[k]), k


# ---- boundary branches against independent numpy statements (tests/manufactured.py) ----

def test_manufactured_mixed_poisson_gpu(oracle):
    """ns_pr_bc_staggered_modify_rhs + ns_dct_poisson_solver on the
    reference's manufactured DST-2 / DCT-2 problem (test_subroutine.f90:280-297):
    second-order convergence to the exact solution, and the oracle's solve
    to 1e-12"""
    from manufactured import manufactured_case

    from ns_3d_staggered_dp_amd.grid import pack_planes

    errs = []
    for n in (16, 32, 64):
        h, rhs, pl, phi = manufactured_case(n)
        nz = rhs.shape[2]
        planes = pack_planes(n, n, nz, pl)
        cfg = NSConfig(nx=n, ny=n, nz=nz, dx=h, dy=h, dz=h, bc=(2, 2, 2), pbc=(2, 3, 3), dp_flag=1)
        r = rhs.copy(order="F")
        with NSSolver(cfg) as s:
            s.pr_bc_staggered_modify_rhs(r, planes)
            sol = s.DCT_poisson_solver(r, 0.0)
        ro = rhs.copy(order="F")
        oracle.pr_bc_modify_rhs(ro, planes, (2, 3, 3), h, h, h)
        assert np.array_equal(r, ro)
        assert rel(sol, oracle.poisson_dct(ro, (2, 3, 3), h, h, h, 0.0)) <= 1e-12
        errs.append(np.abs(sol - phi).max())
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 1.95), (errs, orders)


@pytest.mark.parametrize("pbc", [(2, 2, 2), (3, 3, 3), (4, 4, 4), (2, 3, 4)])
def test_modify_rhs_moves_the_fill_terms_gpu(pbc):
    """ns_pr_bc_staggered_modify_rhs adds exactly the planes' part of the
    Laplacian of the field ns_pr_bc_staggered fills"""
    from manufactured import interior_laplacian

    cfg = cfg_box(9, 7, 6, pbc=pbc)
    rng = np.random.default_rng(31)
    p = np.asfortranarray(rng.standard_normal((11, 9, 8)))
    planes = rng.standard_normal(plane_set_length(9, 7, 6))
    a, b = p.copy(order="F"), p.copy(order="F")
    d = (cfg.dx, cfg.dy, cfg.dz)
    with NSSolver(cfg) as s:
        s.pr_bc_staggered(a, planes)
        s.pr_bc_staggered(b, np.zeros_like(planes))
        r = np.asfortranarray(interior_laplacian(a, *d))
        s.pr_bc_staggered_modify_rhs(r, planes)
    ref = interior_laplacian(b, *d)
    assert np.abs(r - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("bc", [2, 4])
@pytest.mark.parametrize("direction", [1, 2, 3])
def test_cn_rows_and_sweep_solve_the_line_equations_gpu(bc, direction):
    """ns_vel_bc_staggered_cn2 (codes 2, 4) writes the planes into the line
    ends and ns_adi_sweep then satisfies the CN line equations stated in numpy"""
    from manufactured import cn2_rows, cn_line_residual

    from ns_3d_staggered_dp_amd.grid import unpack_planes

    cfg = cfg_box(12, 10, 8, bc=(bc, bc, bc), pbc=(3, 3, 3) if bc == 2 else (2, 2, 2), scheme=AB2_CN,
                  dp_flag=1).replace(dt=0.05, nu=0.5)
    u, v, w, _ = rand_fields(cfg, 40 + direction)
    planes = np.random.default_rng(50 + direction).standard_normal(plane_set_length(12, 10, 8))
    r = cn2_rows(u, v, w, unpack_planes(12, 10, 8, planes), direction)
    with NSSolver(cfg) as s:
        s.vel_bc_staggered_CN2(u, v, w, planes, bc, direction)
        for x, ref in zip((u, v, w), r):
            assert np.array_equal(x, ref)
        s.adi_sweep(u, v, w, direction)
    h = (cfg.dx, cfg.dy, cfg.dz)[direction - 1]
    c = 0.5 * cfg.dt * cfg.nu / (h * h)
    for comp, (x, rr) in enumerate(zip((u, v, w), r)):
        res = cn_line_residual(x, rr, direction - 1, c, comp == direction - 1, bc)
        assert res <= 1e-12 * np.abs(rr).max(), (comp, res)
