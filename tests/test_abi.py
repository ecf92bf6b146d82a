"""CPU-side checks of the C ABI library: it loads and exports every entry
point that include/ns_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os

from ns_3d_staggered_dp_amd import _abi
from ns_3d_staggered_dp_amd.config import CConfig
from ns_3d_staggered_dp_amd.grid import plane_set_length

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ns_hip.h")


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = _abi.declared_symbols(HEADER)
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    assert set(names) <= set(_abi.SIGNATURES), set(names) - set(_abi.SIGNATURES)


def test_plane_set_length_matches_python_layout():
    lib = _abi.lib()
    for n in [(4, 5, 6), (16, 16, 16), (33, 20, 7)]:
        assert lib.ns_plane_set_length(*n) == plane_set_length(*n)


def test_config_struct_layout_matches_header():
    # int nx,ny,nz; double dx,dy,dz,nu,dt; int scheme, bc[3], pbc[3], dp_flag, flags
    assert CConfig.dx.offset == 16
    assert CConfig.scheme.offset == 56
    assert CConfig.flags.offset == 88
    assert ctypes.sizeof(CConfig) == 96


def test_invalid_config_rejected_without_gpu_work():
    lib = _abi.lib()
    h = ctypes.c_void_p()
    assert lib.ns_create(None, ctypes.byref(h)) == _abi.NS_ERR_INVALID
    c = CConfig()
    c.nx = c.ny = c.nz = 8
    c.dx = c.dy = c.dz = 0.1
    c.nu, c.dt = 0.01, 1e-3
    c.scheme = 7  # unknown time scheme
    for i in range(3):
        c.bc[i] = c.pbc[i] = 1
    assert lib.ns_create(ctypes.byref(c), ctypes.byref(h)) == _abi.NS_ERR_INVALID
    c.scheme = 1
    c.bc[0] = 3  # velocity bc 3 has no fill in the reference (NS_functions.f90:249)
    assert lib.ns_create(ctypes.byref(c), ctypes.byref(h)) == _abi.NS_ERR_INVALID


def test_h5_library_exports_every_declared_symbol():
    from ns_3d_staggered_dp_amd import h5io

    if not os.path.exists(h5io.LIB_PATH):
        import pytest

        pytest.skip("HDF5 not installed: libns_h5.so not built")
    lib = ctypes.CDLL(h5io.LIB_PATH)
    names = _abi.declared_symbols(os.path.join(ROOT, "include", "ns_h5.h"))
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    h5io.lib()  # every declared symbol gets its ctypes signature


def test_spline_weights_host_entry_needs_no_gpu():
    from ns_3d_staggered_dp_amd.resim import spline_weights

    assert abs(spline_weights([0.0, 1.0, 2.0, 3.0], 1.5).sum() - 1.0) < 1e-15
