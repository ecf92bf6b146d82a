"""GPU: the discrete adjoint (ns_adjoint_step) against finite differences.

One forward step is a polynomial of degree two in the state and the plane
set (the convection is quadratic, everything else affine), so the central
difference (J(x+ed) - J(x-ed))/2e of J = <y, step(x)> is the exact
directional derivative up to rounding: the adjoint gradient must match it to
~1e-10 relative for random x, y, d that touch every input (u, v, w, p, the
AB2 history, all planes and mean_dp).  Over several steps (checkpoint and
recompute) the bar is the 1e-6 of central differences with a small step."""
import ctypes

import numpy as np
import pytest

from ns_3d_staggered_dp_amd import (AB2, AB2_CN, EULER, FLAG_CN_CYCLIC, FLAG_EAGER_PROJECTION, FLAG_GAUGE_CORNER,
                                    NSConfig, _abi, fields)
from ns_3d_staggered_dp_amd.grid import plane_set_length
from ns_3d_staggered_dp_amd.solver import NSSolver, _ptr

pytestmark = pytest.mark.gpu

KEYS = ("u", "v", "w", "p", "hx", "hy", "hz")


def box(nx, ny, nz, **kw):
    return NSConfig(nx=nx, ny=ny, nz=nz, dx=0.31, dy=0.27, dz=0.23, nu=0.05, dt=2e-3, **kw)


def periodic(nx, ny, nz, **kw):
    return NSConfig(nx=nx, ny=ny, nz=nz, dx=2 * np.pi / nx, dy=0.4, dz=0.5, nu=0.02, dt=2e-3, bc=(1, 1, 1),
                    pbc=(1, 1, 1), **kw)


def rand_state(cfg, seed, amp=0.5):
    rng = np.random.default_rng(seed)
    u, v, w, p = fields.random_ghosted(cfg, seed)
    st = {"u": amp * u, "v": amp * v, "w": amp * w, "p": p}
    for k, a in zip(("hx", "hy", "hz"), fields.random_ghosted(cfg, seed + 1)[:3]):
        st[k] = 0.3 * a
    return st, rng


def step_once(s, st, planes, first):
    s.upload(*(st[k] for k in KEYS), first=first)
    s.set_bc_planes(planes)
    s.step(1)
    return s.download()


CASES = [
    ("walls_ab2", lambda: box(12, 10, 8, bc=(2, 2, 2), pbc=(3, 3, 3), scheme=AB2, dp_flag=1), False),
    ("walls_ab2_first", lambda: box(12, 10, 8, bc=(2, 2, 2), pbc=(3, 3, 3), scheme=AB2, dp_flag=1), True),
    ("walls_euler_dst", lambda: box(10, 9, 7, bc=(4, 4, 4), pbc=(2, 2, 2), scheme=EULER, dp_flag=1), False),
    ("walls_mixed", lambda: box(10, 8, 9, bc=(2, 4, 2), pbc=(3, 2, 3), scheme=AB2, dp_flag=1), False),
    ("periodic_fft", lambda: periodic(16, 12, 10, scheme=AB2, dp_flag=1), False),
    ("periodic_gauge_eager", lambda: periodic(16, 12, 10, scheme=AB2, dp_flag=0,
                                              flags=FLAG_EAGER_PROJECTION | FLAG_GAUGE_CORNER), False),
    # lazily projected contexts (the periodic parent's default: u*, dp stored between steps)
    ("periodic_lazy", lambda: periodic(16, 12, 10, scheme=AB2, dp_flag=0), False),
    ("periodic_lazy_gauge_first", lambda: periodic(16, 12, 10, scheme=AB2, dp_flag=0, flags=FLAG_GAUGE_CORNER), True),
    ("periodic_lazy_cn", lambda: periodic(16, 12, 10, scheme=AB2_CN, dp_flag=0, flags=FLAG_GAUGE_CORNER), False),
    ("walls_cn", lambda: box(12, 10, 8, bc=(2, 2, 2), pbc=(3, 3, 3), scheme=AB2_CN, dp_flag=1), False),
    ("walls_cn_first", lambda: box(12, 10, 8, bc=(2, 2, 2), pbc=(3, 3, 3), scheme=AB2_CN, dp_flag=1), True),
    ("walls_cn_mixed_dst", lambda: box(10, 8, 9, bc=(2, 4, 4), pbc=(3, 2, 2), scheme=AB2_CN, dp_flag=1), False),
    ("walls_cn_long_lines", lambda: box(40, 6, 36, bc=(4, 2, 2), pbc=(3, 3, 3), scheme=AB2_CN, dp_flag=1), False),
    # 9 partitioned blocks per x line; x lines staged 32 per block; y lines in blocks of up to 40 rows;
    # y lines over 640 rows (twisted solve)
    ("walls_cn_k9_lines", lambda: box(250, 5, 4, bc=(2, 2, 4), pbc=(3, 3, 2), scheme=AB2_CN, dp_flag=1), False),
    ("walls_cn_wide_x", lambda: box(320, 4, 3, bc=(2, 2, 4), pbc=(3, 3, 2), scheme=AB2_CN, dp_flag=1), False),
    ("walls_cn_rows40_lines", lambda: box(4, 480, 3, bc=(2, 4, 2), pbc=(3, 2, 3), scheme=AB2_CN, dp_flag=1), False),
    ("walls_cn_twisted_lines", lambda: box(4, 700, 3, bc=(2, 4, 2), pbc=(3, 2, 3), scheme=AB2_CN, dp_flag=1), False),
    # periodic CN lines: the reference's row-1 coupling (default) and the cyclic operator
    ("periodic_cn", lambda: periodic(16, 12, 10, scheme=AB2_CN, dp_flag=1), False),
    ("periodic_cn_first", lambda: periodic(16, 12, 10, scheme=AB2_CN, dp_flag=1), True),
    ("periodic_cn_cyclic_gauge", lambda: periodic(16, 12, 10, scheme=AB2_CN, dp_flag=0,
                                                  flags=FLAG_CN_CYCLIC | FLAG_EAGER_PROJECTION | FLAG_GAUGE_CORNER),
     False),
    ("cn_periodic_x_walls_yz", lambda: box(12, 10, 8, bc=(1, 2, 4), pbc=(3, 3, 2), scheme=AB2_CN, dp_flag=1), False),
    # long periodic lines: x staged 32 per block (320), blocks of up to 40 rows (480), one-lane cyclic (700)
    ("periodic_cn_wide_x", lambda: periodic(320, 4, 3, scheme=AB2_CN, dp_flag=1), False),
    ("periodic_cn_rows40_x", lambda: periodic(480, 4, 3, scheme=AB2_CN, dp_flag=1), False),
    ("periodic_cn_cyclic_x", lambda: periodic(700, 4, 3, scheme=AB2_CN, dp_flag=1), False),
]


@pytest.mark.parametrize("name,mk,first", CASES, ids=[c[0] for c in CASES])
def test_adjoint_step_exact_on_quadratic_map(name, mk, first):
    check_exact(mk(), first)


def _shapes(seed, count):
    """seeded shapes: one axis of 7-smooth length in [60, 400], two in [2, 24]"""
    def smooth(n):
        for f in (2, 3, 5, 7):
            while n % f == 0:
                n //= f
        return n == 1
    long_ = [n for n in range(60, 401) if smooth(n)]
    short = [n for n in range(2, 25) if smooth(n)]
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = [int(rng.choice(short)) for _ in range(3)]
        n[int(rng.integers(3))] = int(rng.choice(long_))
        out.append(tuple(n))
    return out


SHAPES = _shapes(77, 4) + [(64, 64, 64), (128, 128, 128)]


@pytest.mark.parametrize("n", SHAPES, ids=["x".join(map(str, n)) for n in SHAPES])
@pytest.mark.parametrize("kind", ["walls_cn", "periodic_cn"])
def test_adjoint_step_exact_random_shapes(n, kind):
    """the exactness check on seeded box shapes with one long axis, whose
    launch shapes (the adjoint z-march's chunks, the sweeps' blocks per line
    and LDS staging, the transform passes' lines per block) differ"""
    if kind == "walls_cn":
        cfg = box(*n, bc=(2, 4, 2), pbc=(3, 2, 3), scheme=AB2_CN, dp_flag=1)
    else:
        cfg = periodic(*n, scheme=AB2_CN, dp_flag=1)
    check_exact(cfg, False)


def check_exact(cfg, first):
    st, rng = rand_state(cfg, 31)
    L = plane_set_length(cfg.nx, cfg.ny, cfg.nz)
    planes = 0.2 * rng.standard_normal(L)
    y = {k: np.asfortranarray(rng.standard_normal(st[k].shape)) for k in KEYS}
    d = {k: np.asfortranarray(rng.standard_normal(st[k].shape)) for k in KEYS}
    dpl = rng.standard_normal(L)
    with NSSolver(cfg) as s:
        def J(eps):
            xs = {k: np.asfortranarray(st[k] + eps * d[k]) for k in KEYS}
            out = step_once(s, xs, planes + eps * dpl, first)
            return sum(float(np.sum(y[k] * out[k])) for k in KEYS)

        eps = 1e-2
        fd = (J(eps) - J(-eps)) / (2 * eps)
        s.upload(*(st[k] for k in KEYS), first=first)
        _abi.call("ns_adjoint_set_state", s.h, *(_ptr(y[k]) for k in KEYS))
        pbar = np.zeros(L)
        _abi.call("ns_adjoint_step", s.h, _ptr(st["u"]), _ptr(st["v"]), _ptr(st["w"]), int(first), _ptr(pbar))
        g = {k: np.zeros(st[k].shape, order="F") for k in KEYS}
        _abi.call("ns_adjoint_get_state", s.h, *(_ptr(g[k]) for k in KEYS))
    ad = sum(float(np.sum(g[k] * d[k])) for k in KEYS) + float(pbar @ dpl)
    scale = sum(float(np.sum(np.abs(g[k] * d[k]))) for k in KEYS) + float(np.abs(pbar * dpl).sum())
    assert abs(fd - ad) <= 1e-10 * scale, (fd, ad, scale)


def test_adjoint_step_needs_state():
    """ns_adjoint_step before ns_adjoint_set_state is an invalid call"""
    cfg = periodic(8, 8, 8, scheme=AB2, dp_flag=0)
    st, _ = rand_state(cfg, 3)
    with NSSolver(cfg) as s:
        rc = _abi.lib().ns_adjoint_step(s.h, _ptr(st["u"]), _ptr(st["v"]), _ptr(st["w"]), 0, None)
        assert rc == _abi.NS_ERR_INVALID


@pytest.mark.parametrize("scheme,lazy", [(AB2, False), (AB2_CN, False), (AB2, True)],
                         ids=["ab2", "ab2_cn", "periodic_lazy_ab2"])
def test_rollout_gradient_matches_central_differences(scheme, lazy):
    """torch.autograd through 6 steps (checkpoint every 2, recompute), walls
    with per-step planes (or a lazily projected periodic box): d/d(u0, v0, w0,
    planes) of a sub-domain mismatch, checked in random directions against
    central differences (1e-6)."""
    import torch

    from ns_3d_staggered_dp_amd.adjoint import Rollout, make_rollout_function

    if lazy:
        cfg = periodic(12, 10, 8, scheme=scheme, dp_flag=0, flags=FLAG_GAUGE_CORNER)
    else:
        cfg = box(12, 10, 8, bc=(2, 2, 2), pbc=(3, 3, 3), scheme=scheme, dp_flag=1)
    nsteps = 6
    st, rng = rand_state(cfg, 41, amp=0.3)
    L = plane_set_length(cfg.nx, cfg.ny, cfg.nz)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).cuda()
    u0, v0, w0 = dev(st["u"]), dev(st["v"]), dev(st["w"])
    planes = dev(0.1 * rng.standard_normal(nsteps * L))
    target = {k: dev(0.3 * rng.standard_normal(st[k].shape)) for k in ("u", "v", "w")}
    mask = {k: dev((rng.uniform(size=st[k].shape) < 0.5).astype(np.float64)) for k in ("u", "v", "w")}
    ro = Rollout(cfg, nsteps, checkpoint_every=2)
    F = make_rollout_function()

    def objective(a, b, c, pl):
        uN, vN, wN = F.apply(a, b, c, pl, ro)
        return 0.5 * sum(((x - target[k]) ** 2 * mask[k]).sum() for k, x in (("u", uN), ("v", vN), ("w", wN)))

    xs = [t.clone().requires_grad_(True) for t in (u0, v0, w0, planes)]
    J0 = objective(*xs)
    grads = torch.autograd.grad(J0, xs)
    for trial in range(3):
        d = [torch.from_numpy(np.random.default_rng(50 + trial).standard_normal(t.numel())).cuda() for t in xs]
        eps = 1e-6
        with torch.no_grad():
            jp = objective(*(t + eps * dd for t, dd in zip(xs, d))).item()
            jm = objective(*(t - eps * dd for t, dd in zip(xs, d))).item()
        fd = (jp - jm) / (2 * eps)
        ad = sum(float((g * dd).sum()) for g, dd in zip(grads, d))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (trial, fd, ad)
    ro.close()


def _rollout_inputs(cfg, nsteps, seed):
    import torch

    st, rng = rand_state(cfg, seed, amp=0.3)
    L = plane_set_length(cfg.nx, cfg.ny, cfg.nz)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).cuda()
    x = [dev(st["u"]), dev(st["v"]), dev(st["w"]), dev(0.1 * rng.standard_normal(nsteps * L))]
    target = [dev(0.3 * rng.standard_normal(st[k].shape)) for k in ("u", "v", "w")]
    return x, target


def test_rollout_two_applications_before_backward():
    """two applications on one Rollout, then one backward of a loss that
    uses both: each application reverses from its own checkpoints and planes,
    so the gradients equal those of each application differentiated alone"""
    import torch

    from ns_3d_staggered_dp_amd.adjoint import Rollout, make_rollout_function

    cfg = box(12, 10, 8, bc=(2, 2, 2), pbc=(3, 3, 3), scheme=AB2_CN, dp_flag=1)
    nsteps = 4
    (xa, ta), (xb, tb) = _rollout_inputs(cfg, nsteps, 61), _rollout_inputs(cfg, nsteps, 62)
    ro = Rollout(cfg, nsteps, checkpoint_every=2)
    F = make_rollout_function()

    def mismatch(xs, target):
        return 0.5 * sum(((o - t) ** 2).sum() for o, t in zip(F.apply(*xs, ro), target))

    alone = []
    for x, t in ((xa, ta), (xb, tb)):
        xs = [a.clone().requires_grad_(True) for a in x]
        alone.append(torch.autograd.grad(mismatch(xs, t), xs))
    xsa = [a.clone().requires_grad_(True) for a in xa]
    xsb = [a.clone().requires_grad_(True) for a in xb]
    together = torch.autograd.grad(mismatch(xsa, ta) + mismatch(xsb, tb), xsa + xsb)
    ro.close()
    for g, ref in zip(together, alone[0] + alone[1]):
        assert torch.equal(g, ref)


def test_rollout_without_planes_steps_with_zero_planes():
    """planes=None means zero plane sets, whatever set the context held
    before; the planes slot of the gradient is then None"""
    import torch

    from ns_3d_staggered_dp_amd.adjoint import Rollout, make_rollout_function

    cfg = box(12, 10, 8, bc=(2, 2, 2), pbc=(3, 3, 3), scheme=AB2, dp_flag=1)
    nsteps = 4
    x, target = _rollout_inputs(cfg, nsteps, 63)
    ro = Rollout(cfg, nsteps, checkpoint_every=2)
    F = make_rollout_function()
    with torch.no_grad():
        F.apply(*x, ro)  # leaves the last of x's non-zero plane sets in the context
    zeros = torch.zeros_like(x[3])
    results = []
    for pl in (None, zeros):
        xs = [a.clone().requires_grad_(True) for a in x[:3]]
        out = F.apply(*xs, pl, ro)
        loss = 0.5 * sum(((o - t) ** 2).sum() for o, t in zip(out, target))
        results.append(([o.detach() for o in out], torch.autograd.grad(loss, xs)))
    ro.close()
    for a, b in zip(results[0][0] + list(results[0][1]), results[1][0] + list(results[1][1])):
        assert torch.equal(a, b)
