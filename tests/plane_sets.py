"""Boundary-plane sets for wall tests (host-side helpers, no GPU).

A wall box (bc 2, pbc 3) ends a step with the wall-normal velocity
u(1) = bx_u_1 - dt*bx_p_1 and u(nx+1) = bx_u_nx - dt*bx_p_nx on its x walls
(vel_bc_staggered bc=2 sets the face, NS_functions.f90:242-248, and the
projection subtracts dt times the code-3 gradient that pr_bc_staggered puts
across the wall, NS_functions.f90:1403-1464), and likewise on the y and z
walls.  The Neumann pressure problem is solvable only when those velocities
carry no net flux; otherwise the transform solve drops the mean of its
right-hand side and the divergence after the step is that mean, the same in
every cell.
"""
from __future__ import annotations

import numpy as np

from ns_3d_staggered_dp_amd.grid import unpack_planes


def net_wall_flux(cfg, planes) -> float:
    """Net outflow through the box walls of the velocity an explicit step
    (Euler, AB2) leaves on them, for bc = (2,2,2), pbc = (3,3,3)."""
    assert tuple(cfg.bc) == (2, 2, 2) and tuple(cfg.pbc) == (3, 3, 3)
    d = unpack_planes(cfg.nx, cfg.ny, cfg.nz, np.asarray(planes))
    dt = cfg.dt
    area = {"x": cfg.dy * cfg.dz, "y": cfg.dx * cfg.dz, "z": cfg.dx * cfg.dy}
    f = 0.0
    for ax, comp in (("x", "u"), ("y", "v"), ("z", "w")):
        hi = d[f"b{ax}_{comp}_n{ax}"] - dt * d[f"b{ax}_p_n{ax}"]
        lo = d[f"b{ax}_{comp}_1"] - dt * d[f"b{ax}_p_1"]
        # interior part of each plane: the faces of the box's own cells
        f += float(np.sum(hi[1:-1, 1:-1] - lo[1:-1, 1:-1])) * area[ax]
    return f


def zero_flux(cfg, planes) -> np.ndarray:
    """planes with a constant added to the interior of bx_u_nx so that
    net_wall_flux is zero (to rounding)."""
    out = np.array(planes, dtype=np.float64, copy=True)
    d = unpack_planes(cfg.nx, cfg.ny, cfg.nz, out)  # views into out
    f = net_wall_flux(cfg, out)
    d["bx_u_nx"][1:-1, 1:-1] -= f / (cfg.ny * cfg.nz * cfg.dy * cfg.dz)
    return out


def divergence(cfg, u, v, w) -> np.ndarray:
    """div u over the interior cells (big_periodic_3D.f90:576-579), numpy"""
    return ((u[1:, 1:-1, 1:-1] - u[:-1, 1:-1, 1:-1]) / cfg.dx
            + (v[1:-1, 1:, 1:-1] - v[1:-1, :-1, 1:-1]) / cfg.dy
            + (w[1:-1, 1:-1, 1:] - w[1:-1, 1:-1, :-1]) / cfg.dz)
