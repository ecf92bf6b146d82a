"""MI355X-native (gfx950, HIP, FP64) NS_3D_Staggered_dp time step.

The compute path is libns_hip.so (hand-written HIP kernels behind the C ABI
of include/ns_hip.h); this package holds the host-side mirror of the
reference's interfaces, grid/plane formats and synthetic inputs.
"""
from .config import (AB2, AB2_CN, EULER, FLAG_CN_CYCLIC, FLAG_EAGER_PROJECTION, FLAG_GAUGE_CORNER,  # noqa: F401
                     FLAG_NO_GRAPH, SCHEMES, NSConfig)
from .grid import field_shapes, pack_planes, plane_layout, plane_set_length, unpack_planes  # noqa: F401

__all__ = [
    "NSConfig", "EULER", "AB2", "AB2_CN", "SCHEMES", "FLAG_GAUGE_CORNER", "FLAG_CN_CYCLIC", "FLAG_NO_GRAPH", "FLAG_EAGER_PROJECTION",
    "field_shapes", "plane_layout", "plane_set_length", "pack_planes", "unpack_planes",
]


def solver(cfg):
    from .solver import NSSolver

    return NSSolver(cfg)
