"""GPU: batched steps of boxes of different sizes (ns_batch), explicit and
AB2-CN, are bitwise the steps of each box on its own (the same kernel bodies
and parameters, one launch for all boxes), including the Euler first step of
fresh boxes next to restarted ones, and graph replay of the steady steps."""
import numpy as np
import pytest

from ns_3d_staggered_dp_amd import AB2, AB2_CN, EULER, FLAG_NO_GRAPH, NSConfig, _abi, fields
from ns_3d_staggered_dp_amd.batch import Batch, box_sizes
from ns_3d_staggered_dp_amd.grid import plane_set_length
from ns_3d_staggered_dp_amd.solver import NSSolver

pytestmark = pytest.mark.gpu

SIZES = [(12, 10, 8), (16, 16, 16), (20, 12, 14), (9, 10, 12), (32, 8, 10), (70, 9, 18)]


def make(nx, ny, nz, scheme, flags=0, pbc=(3, 3, 3), bc=(2, 2, 2)):
    return NSConfig(nx=nx, ny=ny, nz=nz, dx=0.31, dy=0.27, dz=0.23, nu=0.01, dt=1e-3, scheme=scheme, bc=bc, pbc=pbc,
                    dp_flag=1, flags=flags)


def inputs(cfg, seed):
    u, v, w, p = fields.random_ghosted(cfg, seed)
    rng = np.random.default_rng(seed + 100)
    planes = 0.2 * rng.standard_normal(plane_set_length(cfg.nx, cfg.ny, cfg.nz))
    return (0.3 * u, 0.3 * v, 0.3 * w, p), planes


@pytest.mark.parametrize("scheme,flags,bc,pbc", [(AB2, 0, (2, 2, 2), (3, 3, 3)), (EULER, 0, (4, 4, 4), (2, 2, 2)),
                                                (AB2, FLAG_NO_GRAPH, (2, 4, 2), (3, 2, 3)),
                                                (AB2_CN, 0, (2, 2, 2), (3, 3, 3)),
                                                (AB2_CN, FLAG_NO_GRAPH, (4, 2, 4), (2, 3, 2)),
                                                # periodic x velocity lines (cyclic partitioned sweeps) beside walls
                                                (AB2_CN, 0, (1, 2, 4), (3, 3, 2))])
def test_batch_equals_individual_steps(scheme, flags, bc, pbc):
    check_batch([make(*s, scheme, flags, pbc, bc) for s in SIZES])


def test_batch_long_lines():
    """AB2-CN boxes whose lines take different partitioned kernels in one
    batch: x lines staged 32 per block (320, 350 cells), blocks of up to 40
    rows (480 cells) and of 28 rows with the same block count (350 cells:
    13 blocks, as 480), beside a short box"""
    check_batch([make(*s, AB2_CN) for s in [(350, 4, 3), (480, 4, 3), (24, 16, 12), (4, 480, 3), (320, 4, 3)]])


def test_batch_random_shapes():
    """a batch of seeded box shapes, each with one long axis (7-smooth, 60 to
    400) and two short ones (2 to 24): sweeps grouped by blocks per line with
    different rows per lane and x staging, transform passes with different
    lines per block, in one launch each"""
    rng = np.random.default_rng(5)
    smooth = [n for n in range(2, 401) if all(p in (2, 3, 5, 7) for p in _factors(n))]
    long_, short = [n for n in smooth if n >= 60], [n for n in smooth if n <= 24]
    shapes = []
    for _ in range(6):
        n = [int(rng.choice(short)) for _ in range(3)]
        n[int(rng.integers(3))] = int(rng.choice(long_))
        shapes.append(tuple(n))
    check_batch([make(*n, AB2_CN) for n in shapes])


def _factors(n):
    out, f = [], 2
    while n > 1:
        while n % f == 0:
            out.append(f)
            n //= f
        f += 1
    return out


def check_batch(cfgs):
    data = [inputs(c, 60 + i) for i, c in enumerate(cfgs)]
    nsteps = 5
    ref = []
    for i, (c, (st, pl)) in enumerate(zip(cfgs, data)):
        with NSSolver(c) as s:
            s.upload(*st, first=(i % 2 == 0))
            s.set_bc_planes(pl)
            s.step(nsteps)
            ref.append(s.download())
    solvers = [NSSolver(c) for c in cfgs]
    try:
        for i, (s, (st, pl)) in enumerate(zip(solvers, data)):
            s.upload(*st, first=(i % 2 == 0))
            s.set_bc_planes(pl)
        with Batch(solvers) as b:
            b.step(2)
            b.step(nsteps - 2)
        for i, s in enumerate(solvers):
            got = s.download()
            for k in ("u", "v", "w", "p", "hx", "hy", "hz"):
                assert np.array_equal(got[k], ref[i][k]), ((cfgs[i].nx, cfgs[i].ny, cfgs[i].nz), k)
    finally:
        for s in solvers:
            s.close()


def test_batch_rejects_mixed_or_unsupported_boxes():
    a = NSSolver(make(8, 8, 8, AB2))
    b = NSSolver(make(8, 8, 8, EULER))
    per = NSSolver(NSConfig.cube(8, scheme=AB2, dp_flag=1))  # periodic: FFT Poisson, not batched
    try:
        import ctypes

        for pair in ((a, b), (a, per)):
            arr = (ctypes.c_void_p * 2)(pair[0].h, pair[1].h)
            h = ctypes.c_void_p()
            assert _abi.lib().ns_batch_create(arr, 2, ctypes.byref(h)) == _abi.NS_ERR_UNSUPPORTED
    finally:
        for s in (a, b, per):
            s.close()


def test_box_sizes_are_transform_sizes():
    sizes = box_sizes(64, 32, 128, seed=4)
    assert len(sizes) == 64 and all(32 <= n <= 128 for s in sizes for n in s)
    assert box_sizes(64, seed=4) == sizes


def test_batch_groups_equal_individual_steps():
    """24 boxes: the batch deals them into three groups of about equal cell
    counts, each stepped on its own stream; every box still steps bitwise as
    on its own, fresh (Euler first step) and restarted boxes alike"""
    sizes = box_sizes(24, 8, 36, seed=11)
    check_batch([make(*sz, AB2_CN) for sz in sizes])


@pytest.mark.parametrize("sizes", [[(12, 10, 8), (70, 64, 40), (40, 36, 48), (96, 90, 126)],  # three size classes
                                   box_sizes(16, 10, 40, seed=9)],  # 16 boxes: two groups on their own streams
                         ids=["classes", "two_groups"])
def test_tuned_batch_keeps_state_and_equals_individual_steps(sizes):
    """ns_batch_tune times launch shapes on the boxes' own state and restores
    it; the tuned batch still steps bitwise like the boxes on their own."""
    cfgs = [make(*sz, AB2_CN) for sz in sizes]
    data = [inputs(c, 80 + i) for i, c in enumerate(cfgs)]
    ref = []
    for c, (st, pl) in zip(cfgs, data):
        with NSSolver(c) as s:
            s.upload(*st, first=False)
            s.set_bc_planes(pl)
            s.step(3)
            ref.append(s.download())
    solvers = [NSSolver(c) for c in cfgs]
    try:
        for s, (st, pl) in zip(solvers, data):
            s.upload(*st, first=False)
            s.set_bc_planes(pl)
        before = [s.download() for s in solvers]
        with Batch(solvers) as b:
            knobs = b.tune(trial_steps=1)
            assert set(knobs) == {"small", "medium", "large"}
            for s, st0 in zip(solvers, before):
                got = s.download()
                for k in ("u", "v", "w", "p", "hx", "hy", "hz"):
                    assert np.array_equal(got[k], st0[k]), k
            b.step(3)
        for i, s in enumerate(solvers):
            got = s.download()
            for k in ("u", "v", "w", "p", "hx", "hy", "hz"):
                assert np.array_equal(got[k], ref[i][k]), (sizes[i], k)
    finally:
        for s in solvers:
            s.close()
