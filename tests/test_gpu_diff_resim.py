"""GPU: gradients of a differentiable re-simulation at its full size.

A 128^3 wall box (velocity code 2, pressure code 3, dp_flag 1, AB2-CN, DCT
Poisson) advanced 200 sub-steps with per-step boundary planes, checkpoints
every 20 steps and recompute; the loss is the L2 mismatch of the final
velocity with a target (the unperturbed final velocity plus 1% solenoidal
noise).  The adjoint's gradient with respect to the initial velocity and
every step's plane set is checked against central differences of the solver
itself in 16 seeded random directions, and, at 32^3, against central
differences of the FP64 oracle (an independent CPU implementation of the
forward map).  Inputs are synthetic (no parent run): a smooth solenoidal
initial field of the decaying-turbulence spectrum and small random planes."""
import numpy as np
import pytest

from ns_3d_staggered_dp_amd import AB2_CN, NSConfig, fields
from ns_3d_staggered_dp_amd.grid import field_shapes, plane_set_length

pytestmark = pytest.mark.gpu

NSTEPS, CKPT = 200, 20


def sub_box(n):
    h = 2 * np.pi / 512  # a sub-box of a 512^3 parent on [0, 2 pi]^3
    return NSConfig(nx=n, ny=n, nz=n, dx=h, dy=h, dz=h, nu=1e-3, dt=5e-4 / 4, scheme=AB2_CN, bc=(2, 2, 2),
                    pbc=(3, 3, 3), dp_flag=1)


def setup(n, seed):
    import torch

    from ns_3d_staggered_dp_amd.adjoint import Rollout, make_rollout_function

    cfg = sub_box(n)
    u, v, w, _ = fields.decaying_turbulence(cfg, seed=seed)
    rng = np.random.default_rng(seed + 1)
    L = plane_set_length(n, n, n)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).cuda()
    x0 = [dev(u), dev(v), dev(w), dev(0.05 * rng.standard_normal(NSTEPS * L))]
    ro = Rollout(cfg, NSTEPS, checkpoint_every=CKPT)
    F = make_rollout_function()
    with torch.no_grad():
        fin = F.apply(*x0, ro)
    noise = [dev(a) for a in fields.decaying_turbulence(cfg, seed=seed + 2)[:3]]
    target = [f + 0.01 * float(f.std()) / float(z.std()) * z for f, z in zip(fin, noise)]

    def loss(*xs):
        out = F.apply(*xs, ro)
        return 0.5 * sum(((o - t) ** 2).sum() for o, t in zip(out, target))

    return cfg, x0, ro, loss, target


def directions(x0, seed, count):
    import torch

    rng = np.random.default_rng(seed)
    return [[torch.from_numpy(rng.standard_normal(t.numel())).cuda() for t in x0] for _ in range(count)]


def test_gradient_128_matches_central_differences():
    import torch

    cfg, x0, ro, loss, _ = setup(128, 3)
    xs = [t.clone().requires_grad_(True) for t in x0]
    grads = torch.autograd.grad(loss(*xs), xs)
    eps = 1e-5
    errs = []
    for d in directions(x0, 6, 16):
        with torch.no_grad():
            jp = loss(*(t + eps * dd for t, dd in zip(x0, d))).item()
            jm = loss(*(t - eps * dd for t, dd in zip(x0, d))).item()
        fd = (jp - jm) / (2 * eps)
        ad = sum(float((g * dd).sum()) for g, dd in zip(grads, d))
        errs.append(abs(fd - ad) / abs(ad))
    ro.close()
    # measured: 2e-10 to 1.5e-8 over the 16 directions
    assert max(errs) <= 1e-6, errs


def test_gradient_32_matches_oracle_differences(oracle):
    import torch

    cfg, x0, ro, loss, target = setup(32, 8)
    xs = [t.clone().requires_grad_(True) for t in x0]
    grads = torch.autograd.grad(loss(*xs), xs)
    L = plane_set_length(32, 32, 32)
    tgt = [t.cpu().numpy() for t in target]

    def oracle_loss(xs_):
        u, v, w, pl = (t.cpu().numpy() for t in xs_)
        sh = field_shapes(32, 32, 32)
        f = lambda a, k: np.asfortranarray(a.reshape(sh[k], order="F"))
        st = oracle.OracleState(cfg, f(u, "u"), f(v, "v"), f(w, "w"), np.zeros(sh["p"], order="F"), first=True)
        for n in range(NSTEPS):
            st.step(pl[n * L:(n + 1) * L])
        return 0.5 * sum(float(((np.asarray(o).ravel(order="F") - t) ** 2).sum()) for o, t in zip((st.u, st.v, st.w), tgt))

    eps = 1e-5
    for i, d in enumerate(directions(x0, 9, 3)):
        jp = oracle_loss([t + eps * dd for t, dd in zip(x0, d)])
        jm = oracle_loss([t - eps * dd for t, dd in zip(x0, d)])
        fd = (jp - jm) / (2 * eps)
        ad = sum(float((g * dd).sum()) for g, dd in zip(grads, d))
        # measured: 3e-11 to 1.1e-9
        assert abs(fd - ad) <= 1e-6 * abs(ad), (i, fd, ad)
    ro.close()
