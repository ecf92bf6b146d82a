"""The Fortran host path: ns_hip_mod.f90 (ISO_C_BINDING wrappers with the
reference's subroutine names) compiles with flang, and a Fortran driver in
the shape of big_periodic_3D.f90 reproduces the oracle through it."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from ns_3d_staggered_dp_amd import AB2, FLAG_GAUGE_CORNER, NSConfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "ns_3d_staggered_dp_amd", "fortran")
DRIVER = os.path.join(FDIR, "_build", "ns_hip_driver")


def _flang():
    return shutil.which("flang") or ("/opt/rocm/lib/llvm/bin/flang" if os.path.exists("/opt/rocm/lib/llvm/bin/flang") else None)


def test_fortran_module_compiles():
    fl = _flang()
    if fl is None:
        pytest.skip("flang not installed")
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([fl, "-c", os.path.join(FDIR, "ns_hip_mod.f90"), "-module-dir", d, "-o",
                            os.path.join(d, "m.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def _read(path, shape):
    return np.fromfile(path, dtype=np.float64).reshape(shape, order="F")


@pytest.mark.gpu
def test_fortran_driver_matches_oracle(oracle, tmp_path):
    assert os.path.exists(DRIVER), "build the Fortran driver first (__graft_entry__.build())"
    n, steps = 16, 5
    r = subprocess.run([DRIVER, str(n), str(steps), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    u = _read(tmp_path / "init_u.bin", (n + 1, n + 2, n + 2))
    v = _read(tmp_path / "init_v.bin", (n + 2, n + 1, n + 2))
    w = _read(tmp_path / "init_w.bin", (n + 2, n + 2, n + 1))
    dx = 2 * 3.1415926535897932 / n
    cx, cy, _ = oracle.cal_conv(u, v, w, (1, 1, 1), dx, dx, dx)
    _, _, fz = oracle.cal_diff(u, v, w, (1, 1, 1), dx * dx, dx * dx, dx * dx)
    assert np.array_equal(_read(tmp_path / "conv_x.bin", (n, n, n)), cx)
    assert np.array_equal(_read(tmp_path / "conv_y.bin", (n, n, n)), cy)
    assert np.array_equal(_read(tmp_path / "diff_z.bin", (n, n, n)), fz)
    cfg = NSConfig.cube(n, nu=1.0 / 1600.0, dt=1e-3, scheme=AB2, flags=FLAG_GAUGE_CORNER)
    st = oracle.OracleState(cfg, u, v, w)
    for _ in range(steps):
        st.step()
    for name, ref, shape in (("u", st.u, u.shape), ("v", st.v, v.shape), ("w", st.w, w.shape),
                             ("p", st.p, (n + 2,) * 3)):
        got = _read(tmp_path / f"final_{name}.bin", shape)
        den = max(np.abs(ref).max(), 1e-300)
        assert np.abs(got - ref).max() / den <= 1e-11, name
