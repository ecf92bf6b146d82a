"""Independent numpy statements used to pin the boundary branches of the
pressure solve and of the CN sweeps (no oracle or library code inside).

- A manufactured mixed Dirichlet / Neumann Poisson problem, the one of the
  reference's transform test (test_subroutine.f90:65-67,280-297): on
  [0,1]^2 with cell centres (i-0.5)h, phi = (1-x^2)(2y^3-3y^2+1) has
  lap phi = -2(2y^3-3y^2+1) + 6(1-x^2)(2y-1); x walls are Dirichlet on the
  cell wall (pbc 2: phi(0,y) = 2y^3-3y^2+1, phi(1,y) = 0), y walls Neumann
  on the cell wall (pbc 3: dphi/dy = 0 at both ends).  The reference adds
  +1/dy at the low y wall (:296), which does not match this phi
  (dphi/dy(x,0) = 0); the consistent value 0 is used.  Here the problem is
  extruded in z with Neumann walls (phi independent of z), so a 3-D solve
  must converge to phi at second order.
- The discrete Laplacian of a ghost-filled pressure field (7 points over the
  interior cells), so that the boundary terms pr_bc_staggered_modify_rhs adds
  can be checked against the ghost values pr_bc_staggered writes.
- The CN line equations of mat_CN_tridiagonal (NS_functions.f90:1297-1342):
  interior rows (1 + 2c) x_i - c (x_{i-1} + x_{i+1}) = r_i with
  c = dt*nu/(2 h^2); end rows x = r (faces of codes 2 and 4, centres of
  code 4) or (x_1 + x_2)/2 = r (centres of code 2, the wall average).
"""
from __future__ import annotations

import numpy as np


def manufactured_phi(x, y):
    return (1.0 - x * x) * (2.0 * y ** 3 - 3.0 * y ** 2 + 1.0)


def manufactured_case(n, nz=4):
    """(h, rhs(n,n,nz), plane dict, exact phi(n,n,nz)) on [0,1]^2 x [0, nz h]"""
    h = 1.0 / n
    xc = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    lap = -2.0 * (2.0 * Y ** 3 - 3.0 * Y ** 2 + 1.0) + 6.0 * (1.0 - X * X) * (2.0 * Y - 1.0)
    rhs = np.asfortranarray(np.repeat(lap[:, :, None], nz, axis=2))
    phi = np.repeat(manufactured_phi(X, Y)[:, :, None], nz, axis=2)
    # pressure planes (ny+2, nz+2) etc.; only their interior (the wall faces of the cells) enters
    yg = (np.arange(n + 2) - 0.5) * h
    bx1 = np.repeat((2.0 * yg ** 3 - 3.0 * yg ** 2 + 1.0)[:, None], nz + 2, axis=1)
    planes = {"bx_p_1": bx1, "bx_p_nx": np.zeros((n + 2, nz + 2)),
              "by_p_1": np.zeros((n + 2, nz + 2)), "by_p_ny": np.zeros((n + 2, nz + 2)),
              "bz_p_1": np.zeros((n + 2, n + 2)), "bz_p_nz": np.zeros((n + 2, n + 2))}
    return h, rhs, planes, phi


def interior_laplacian(p, dx, dy, dz):
    """7-point Laplacian over the interior cells of a ghosted (nx+2,ny+2,nz+2) array"""
    c = p[1:-1, 1:-1, 1:-1]
    return ((p[2:, 1:-1, 1:-1] - 2.0 * c + p[:-2, 1:-1, 1:-1]) / (dx * dx)
            + (p[1:-1, 2:, 1:-1] - 2.0 * c + p[1:-1, :-2, 1:-1]) / (dy * dy)
            + (p[1:-1, 1:-1, 2:] - 2.0 * c + p[1:-1, 1:-1, :-2]) / (dz * dz))


def cn_line_residual(x, r, axis, c, face, bc):
    """max |A x - r| over the lines of x along `axis` (0-based) for the CN
    matrix of a component on faces (face=True) or centres along that axis"""
    xl = np.moveaxis(x, axis, 0)
    rl = np.moveaxis(r, axis, 0)
    res = np.zeros_like(xl)
    res[1:-1] = (1.0 + 2.0 * c) * xl[1:-1] - c * (xl[:-2] + xl[2:]) - rl[1:-1]
    if face or bc == 4:
        res[0] = xl[0] - rl[0]
        res[-1] = xl[-1] - rl[-1]
    else:  # code 2 on cell centres: the wall average
        res[0] = 0.5 * (xl[0] + xl[1]) - rl[0]
        res[-1] = 0.5 * (xl[-1] + xl[-2]) - rl[-1]
    return float(np.abs(res).max())


def cn2_rows(u, v, w, planes_by_name, direction):
    """u*, v*, w* with the line ends along `direction` (1, 2, 3) replaced by
    the planes, as vel_bc_staggered_CN2 with code 2 or 4 does (NS_functions.f90:337-405)"""
    ax = "xyz"[direction - 1]
    end = "n" + ax
    out = []
    for comp, a in (("u", u), ("v", v), ("w", w)):
        b = np.array(a, order="F", copy=True)
        lo, hi = planes_by_name[f"b{ax}_{comp}_1"], planes_by_name[f"b{ax}_{comp}_{end}"]
        idx_lo = [slice(None)] * 3
        idx_hi = [slice(None)] * 3
        idx_lo[direction - 1] = 0
        idx_hi[direction - 1] = -1
        b[tuple(idx_lo)] = lo
        b[tuple(idx_hi)] = hi
        out.append(b)
    return out
