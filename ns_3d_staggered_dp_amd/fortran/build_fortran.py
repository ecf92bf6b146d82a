"""Builds the Fortran binding module and the Fortran host drivers with flang
(AMD flang, ISO_C_BINDING) against libns_hip.so, in-tree:
ns_3d_staggered_dp_amd/fortran/_build/{ns_hip_mod.o, ns_hip_mod.mod,
ns_hip_driver, ns_hip_resim_driver}.  -ffp-contract=off keeps a*b+c unfused,
as the library and the reference's build round each operation."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"


DRIVERS = ("ns_hip_driver", "ns_hip_resim_driver")
FFLAGS = ["-O2", "-ffp-contract=off"]


def build():
    os.makedirs(OUT, exist_ok=True)
    mod = os.path.join(HERE, "ns_hip_mod.f90")
    obj = os.path.join(OUT, "ns_hip_mod.o")
    subprocess.run([FLANG, *FFLAGS, "-c", mod, "-module-dir", OUT, "-o", obj], check=True)
    exes = []
    for name in DRIVERS:
        exe = os.path.join(OUT, name)
        subprocess.run([FLANG, *FFLAGS, "-I", OUT, os.path.join(HERE, name + ".f90"), obj, "-o", exe,
                        "-L", PKG, "-lns_hip", "-Wl,-rpath,$ORIGIN/../..", "-Wl,-rpath,/opt/rocm/lib"], check=True)
        exes.append(exe)
    return exes


if __name__ == "__main__":
    print(build())
    sys.exit(0)
