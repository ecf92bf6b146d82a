"""Run configuration shared by the C-ABI (``ns_config`` in include/ns_hip.h).

Mirrors the compile-time ``parameter`` block of the reference drivers
(big_periodic_3D.f90:9-21,104-108; re_simulation_sub_tstep_interBC_combine.f90:100-112):
grid, viscosity, time step, time scheme and boundary codes.  Codes keep the
reference numbering (NS_functions.f90:219-220): 1 periodic, 2 Dirichlet on
the cell wall, 3 Neumann on the cell wall, 4 Dirichlet on the ghost cell.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

EULER, AB2, AB2_CN = 0, 1, 2
SCHEMES = {"Euler": EULER, "AB2": AB2, "AB2-CN": AB2_CN}

# dp = dp - dp(1,1,1) after the pressure solve (big_periodic_3D.f90:562)
FLAG_GAUGE_CORNER = 1
# periodic CN rows use the wrap stencil instead of mat_CN's row 1
# (NS_functions.f90:1273-1278 couples row 1 to column nx+1)
FLAG_CN_CYCLIC = 2
# every step as individual launches (no HIP graph replay of steady steps)
FLAG_NO_GRAPH = 4
# project every step instead of folding the projection of u into the next
# predictor (the fold: dp_flag = 0, or dp_flag = 1 with DCT codes and no gauge)
FLAG_EAGER_PROJECTION = 8

PI = 3.1415926535897932  # big_periodic_3D.f90:6


class CConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("nu", ctypes.c_double), ("dt", ctypes.c_double),
        ("scheme", ctypes.c_int),
        ("bc", ctypes.c_int * 3), ("pbc", ctypes.c_int * 3),
        ("dp_flag", ctypes.c_int),
        ("flags", ctypes.c_int),
    ]


@dataclass
class NSConfig:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    nu: float = 1e-3
    dt: float = 1e-3
    scheme: int = AB2
    bc: tuple = (1, 1, 1)
    pbc: tuple = (1, 1, 1)
    dp_flag: int = 0
    flags: int = 0

    @classmethod
    def cube(cls, n: int, length: float = 2.0 * PI, **kw) -> "NSConfig":
        """n^3 grid on [0, length]^3 (dx = lx/nx as at big_periodic_3D.f90:14)."""
        d = length / n
        return cls(nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, **kw)

    @property
    def all_periodic(self) -> bool:
        return tuple(self.bc) == (1, 1, 1) and tuple(self.pbc) == (1, 1, 1)

    def to_c(self) -> CConfig:
        c = CConfig()
        c.nx, c.ny, c.nz = self.nx, self.ny, self.nz
        c.dx, c.dy, c.dz = self.dx, self.dy, self.dz
        c.nu, c.dt = self.nu, self.dt
        c.scheme = self.scheme
        for i in range(3):
            c.bc[i] = int(self.bc[i])
            c.pbc[i] = int(self.pbc[i])
        c.dp_flag = int(self.dp_flag)
        c.flags = int(self.flags)
        return c

    def replace(self, **kw) -> "NSConfig":
        d = dict(self.__dict__)
        d.update(kw)
        return NSConfig(**d)
