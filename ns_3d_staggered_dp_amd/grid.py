"""Staggered MAC grid shapes and the packed boundary-plane format.

Storage follows the reference exactly (Fortran column-major, one ghost layer):
``u(nx+1,ny+2,nz+2)``, ``v(nx+2,ny+1,nz+2)``, ``w(nx+2,ny+2,nz+1)``,
``p,dp(nx+2,ny+2,nz+2)`` (big_periodic_3D.f90:14-18,27-28).  In numpy these
are arrays of those shapes in Fortran order.

A boundary-plane set is one contiguous float64 vector holding the 18
velocity planes in the argument order of ``vel_bc_staggered``
(NS_functions.f90:213-217), the 6 pressure planes of ``pr_bc_staggered``
(NS_functions.f90:1403-1406), and one trailing scalar ``mean_dp`` (the
Neumann gauge of ``DCT_poisson_solver``, NS_functions.f90:1577).  Each plane
is the reference's 2-D array, column-major: ``bx_u_1(ny+2,nz+2)``,
``by_u_1(nx+1,nz+2)``, ``bz_u_1(nx+1,ny+2)`` ... (big_periodic_3D.f90:53-56).
One vector makes time interpolation a single fused update.
"""
from __future__ import annotations

import numpy as np

PLANE_NAMES = [
    f"b{ax}_{c}_{end}"
    for c in ("u", "v", "w", "p")
    for ax in ("x", "y", "z")
    for end in ("1", "n" + ax)
]


def field_shapes(nx: int, ny: int, nz: int) -> dict:
    return {
        "u": (nx + 1, ny + 2, nz + 2),
        "v": (nx + 2, ny + 1, nz + 2),
        "w": (nx + 2, ny + 2, nz + 1),
        "p": (nx + 2, ny + 2, nz + 2),
    }


def plane_layout(nx: int, ny: int, nz: int):
    """Return ([(name, shape, offset)], total_length) of the packed set."""
    sh = field_shapes(nx, ny, nz)
    out = []
    off = 0
    for c in ("u", "v", "w", "p"):
        s = sh[c]
        for ax, pshape in (("x", (s[1], s[2])), ("y", (s[0], s[2])), ("z", (s[0], s[1]))):
            for end in ("1", "n" + ax):
                out.append((f"b{ax}_{c}_{end}", pshape, off))
                off += pshape[0] * pshape[1]
    return out, off + 1


def plane_set_length(nx: int, ny: int, nz: int) -> int:
    return plane_layout(nx, ny, nz)[1]


def pack_planes(nx: int, ny: int, nz: int, planes: dict | None = None, mean_dp: float = 0.0) -> np.ndarray:
    lay, n = plane_layout(nx, ny, nz)
    buf = np.zeros(n, dtype=np.float64)
    planes = planes or {}
    for name, shape, off in lay:
        if name in planes:
            a = np.asarray(planes[name], dtype=np.float64)
            if a.shape != shape:
                raise ValueError(f"plane {name}: shape {a.shape} != {shape}")
            buf[off:off + shape[0] * shape[1]] = a.ravel(order="F")
    buf[-1] = mean_dp
    return buf


def unpack_planes(nx: int, ny: int, nz: int, buf: np.ndarray) -> dict:
    lay, n = plane_layout(nx, ny, nz)
    if buf.shape != (n,):
        raise ValueError(f"plane set length {buf.shape} != ({n},)")
    d = {name: buf[off:off + shape[0] * shape[1]].reshape(shape, order="F") for name, shape, off in lay}
    d["mean_dp"] = float(buf[-1])
    return d


def interior_count(nx: int, ny: int, nz: int) -> int:
    return nx * ny * nz


def subbox(a: np.ndarray, comp: str, x0: int, y0: int, z0: int, nx0: int, ny0: int, nz0: int) -> np.ndarray:
    """Parent-field slice of a sub-box at 1-based offset (x0,y0,z0), as the
    reference's idx_* index vectors select it (big_periodic_3D.f90:21-25):
    face axes take x0..x0+n0, centred axes x0..x0+n0+1 (ghosts included)."""
    ext = {"u": (nx0 + 1, ny0 + 2, nz0 + 2), "v": (nx0 + 2, ny0 + 1, nz0 + 2),
           "w": (nx0 + 2, ny0 + 2, nz0 + 1), "p": (nx0 + 2, ny0 + 2, nz0 + 2)}[comp]
    o = (x0 - 1, y0 - 1, z0 - 1)
    return np.asfortranarray(a[o[0]:o[0] + ext[0], o[1]:o[1] + ext[1], o[2]:o[2] + ext[2]])
