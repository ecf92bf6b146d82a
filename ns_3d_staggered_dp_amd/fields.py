"""Synthetic initial fields on the staggered grid (host-side input preparation).

Coordinates follow the reference (big_periodic_3D.f90:14-18):
``xu=(i)*dx, i=0..nx`` (faces), ``yu=((i+0.5)*dy, i=-1..ny)`` (centres with
one ghost each side), and the same pattern for v, w, p.

Random fields are built as the discrete staggered curl of a random vector
potential, so they are discretely divergence-free to round-off without a
Poisson solve.  Ghost layers are filled with the periodic copies of
``vel_bc_staggered`` (NS_functions.f90:236-241,257-262,278-283), so the
arrays are exactly what a periodic step would leave behind.
"""
from __future__ import annotations

import numpy as np

from .config import NSConfig, PI
from .grid import field_shapes


def fill_periodic_velocity(u, v, w):
    """Periodic ghost fill in x, then y, then z (NS_functions.f90:236-283)."""
    u[-1, :, :] = u[0, :, :]
    v[0, :, :] = v[-2, :, :]; v[-1, :, :] = v[1, :, :]
    w[0, :, :] = w[-2, :, :]; w[-1, :, :] = w[1, :, :]
    u[:, 0, :] = u[:, -2, :]; u[:, -1, :] = u[:, 1, :]
    v[:, -1, :] = v[:, 0, :]
    w[:, 0, :] = w[:, -2, :]; w[:, -1, :] = w[:, 1, :]
    u[:, :, 0] = u[:, :, -2]; u[:, :, -1] = u[:, :, 1]
    v[:, :, 0] = v[:, :, -2]; v[:, :, -1] = v[:, :, 1]
    w[:, :, -1] = w[:, :, 0]


def fill_periodic_pressure(p):
    """pr_bc_staggered with pbc=1 (NS_functions.f90:1419-1421,1437-1439,1455-1457)."""
    p[0, :, :] = p[-2, :, :]; p[-1, :, :] = p[1, :, :]
    p[:, 0, :] = p[:, -2, :]; p[:, -1, :] = p[:, 1, :]
    p[:, :, 0] = p[:, :, -2]; p[:, :, -1] = p[:, :, 1]


def _alloc(cfg: NSConfig):
    sh = field_shapes(cfg.nx, cfg.ny, cfg.nz)
    return tuple(np.zeros(sh[c], dtype=np.float64, order="F") for c in ("u", "v", "w", "p"))


def staggered_curl(ax, ay, az, dx, dy, dz):
    """Velocity on the owned faces from an edge vector potential (periodic).

    ax at (x-cell, y-face, z-face), ay at (x-face, y-cell, z-face),
    az at (x-face, y-face, z-cell); all arrays (nx, ny, nz).  Returns the
    owned-face arrays u(nx,ny,nz) (faces 1..nx), v, w."""
    r = np.roll
    u = (r(az, -1, 1) - az) / dy - (r(ay, -1, 2) - ay) / dz
    v = (r(ax, -1, 2) - ax) / dz - (r(az, -1, 0) - az) / dx
    w = (r(ay, -1, 0) - ay) / dx - (r(ax, -1, 1) - ax) / dy
    return u, v, w


def place_owned(cfg: NSConfig, uo, vo, wo):
    """Put owned periodic values into ghosted arrays and fill the ghosts."""
    u, v, w, p = _alloc(cfg)
    u[0:cfg.nx, 1:cfg.ny + 1, 1:cfg.nz + 1] = uo
    v[1:cfg.nx + 1, 0:cfg.ny, 1:cfg.nz + 1] = vo
    w[1:cfg.nx + 1, 1:cfg.ny + 1, 0:cfg.nz] = wo
    fill_periodic_velocity(u, v, w)
    return u, v, w, p


def _random_potential(cfg: NSConfig, amp_of_k, rng):
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    kx = np.fft.fftfreq(nx, 1.0 / nx)[:, None, None]
    ky = np.fft.fftfreq(ny, 1.0 / ny)[None, :, None]
    kz = np.fft.rfftfreq(nz, 1.0 / nz)[None, None, :]
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    amp = amp_of_k(kk)
    amp[0, 0, 0] = 0.0
    comps = []
    for _ in range(3):
        phase = rng.uniform(0.0, 2.0 * np.pi, size=amp.shape)
        spec = amp * np.exp(1j * phase)
        comps.append(np.fft.irfftn(spec, s=(nx, ny, nz), axes=(0, 1, 2)))
    return comps


def taylor_green(cfg: NSConfig, seed: int = 1, perturbation: float = 1e-3, kmax: float = 4.0):
    """u=sin x cos y cos z, v=-cos x sin y cos z, w=0 at the staggered
    positions, plus a seeded solenoidal perturbation in modes |k|<=kmax
    scaled to max|u'|=perturbation."""
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    xf = np.arange(nx) * cfg.dx
    xc = (np.arange(nx) + 0.5) * cfg.dx
    yf = np.arange(ny) * cfg.dy
    yc = (np.arange(ny) + 0.5) * cfg.dy
    zc = (np.arange(nz) + 0.5) * cfg.dz
    uo = np.sin(xf)[:, None, None] * np.cos(yc)[None, :, None] * np.cos(zc)[None, None, :]
    vo = -np.cos(xc)[:, None, None] * np.sin(yf)[None, :, None] * np.cos(zc)[None, None, :]
    wo = np.zeros((nx, ny, nz))
    if perturbation:
        rng = np.random.default_rng(seed)
        ax, ay, az = _random_potential(cfg, lambda k: np.where(k <= kmax, 1.0, 0.0), rng)
        pu, pv, pw = staggered_curl(ax, ay, az, cfg.dx, cfg.dy, cfg.dz)
        s = perturbation / max(np.abs(pu).max(), np.abs(pv).max(), np.abs(pw).max())
        uo, vo, wo = uo + s * pu, vo + s * pv, wo + s * pw
    return place_owned(cfg, uo, vo, wo)


def decaying_turbulence(cfg: NSConfig, seed: int = 2, k0: float = 4.0, energy: float = 0.5):
    """Random solenoidal field with E(k) ~ k^4 exp(-2 (k/k0)^2), uniform
    phases, normalised so that 0.5*<u.u> = energy over the owned faces."""

    def amp(k):
        # |u_hat| ~ sqrt(E(k)/(4 pi k^2)); the curl multiplies |A_hat| by ~k
        with np.errstate(divide="ignore", invalid="ignore"):
            e = k ** 4 * np.exp(-2.0 * (k / k0) ** 2)
            a = np.sqrt(e / (4.0 * np.pi * np.maximum(k, 1e-30) ** 2)) / np.maximum(k, 1e-30)
        return np.where(k > 0, a, 0.0)

    rng = np.random.default_rng(seed)
    ax, ay, az = _random_potential(cfg, amp, rng)
    uo, vo, wo = staggered_curl(ax, ay, az, cfg.dx, cfg.dy, cfg.dz)
    e = 0.5 * (np.mean(uo * uo) + np.mean(vo * vo) + np.mean(wo * wo))
    s = np.sqrt(energy / e)
    return place_owned(cfg, uo * s, vo * s, wo * s)


def random_ghosted(cfg: NSConfig, seed: int, scale: float = 1.0):
    """Unstructured random ghosted arrays (for operator-level parity tests)."""
    rng = np.random.default_rng(seed)
    sh = field_shapes(cfg.nx, cfg.ny, cfg.nz)
    return tuple(np.asfortranarray(scale * rng.standard_normal(sh[c])) for c in ("u", "v", "w", "p"))
