"""Builds the Fortran binding module and the Fortran host driver with flang
(AMD flang, ISO_C_BINDING) against libns_hip.so, in-tree:
ns_3d_staggered_dp_amd/fortran/_build/{ns_hip_mod.o, ns_hip_mod.mod, ns_hip_driver}."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"


def build():
    os.makedirs(OUT, exist_ok=True)
    mod = os.path.join(HERE, "ns_hip_mod.f90")
    obj = os.path.join(OUT, "ns_hip_mod.o")
    subprocess.run([FLANG, "-O2", "-c", mod, "-module-dir", OUT, "-o", obj], check=True)
    exe = os.path.join(OUT, "ns_hip_driver")
    subprocess.run([FLANG, "-O2", "-I", OUT, os.path.join(HERE, "ns_hip_driver.f90"), obj, "-o", exe,
                    "-L", PKG, "-lns_hip", "-Wl,-rpath,$ORIGIN/../..", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return exe


if __name__ == "__main__":
    print(build())
    sys.exit(0)
