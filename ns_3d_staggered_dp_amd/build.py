"""Build libns_hip.so (gfx950, FP64) in-tree with hipcc.

Run as ``python -m ns_3d_staggered_dp_amd.build``.  Objects go to
``ns_3d_staggered_dp_amd/_build/``; the library to
``ns_3d_staggered_dp_amd/libns_hip.so`` so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libns_hip.so")
SOURCES = ["ns_stencil.hip", "ns_predict.hip", "ns_poisson.hip", "ns_adi.hip", "ns_metrics.hip", "ns_adjoint.hip",
           "ns_abi.hip"]
EXTRA_SOURCES: list[str] = []

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # each operation rounds on its own, like the reference's build
    "-ffp-contract=off",
    # keep the compiler from emitting packed FP32 ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
]


def _compile(src: str) -> str:
    stem = os.path.splitext(os.path.basename(src))[0]
    obj = os.path.join(BUILD, stem + ".o")
    # the compiler's per-kernel resource report (registers, scratch, occupancy)
    cmd = [HIPCC, *FLAGS, "-Rpass-analysis=kernel-resource-usage", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    with open(os.path.join(BUILD, stem + ".resources.txt"), "w") as f:
        f.write(r.stderr)
    return obj


def kernel_resources() -> dict[str, dict[str, int]]:
    """Per-kernel VGPRs, scratch bytes per lane and occupancy (waves per SIMD)
    from the resource reports of the last build (empty if there is none)."""
    out: dict[str, dict[str, int]] = {}
    keys = {"VGPRs": "vgprs", "ScratchSize [bytes/lane]": "scratch", "Occupancy [waves/SIMD]": "occupancy"}
    for src in SOURCES + EXTRA_SOURCES:
        path = os.path.join(BUILD, os.path.splitext(os.path.basename(src))[0] + ".resources.txt")
        if not os.path.exists(path):
            continue
        name = None
        for line in open(path):
            if "remark:" not in line:
                continue
            text = line.split("remark:", 1)[1].split("[-Rpass", 1)[0].strip()
            if text.startswith("Function Name:"):
                name = text.split(":", 1)[1].strip()
                out[name] = {}
            elif name is not None:
                for k, v in keys.items():
                    if text.startswith(k + ":"):
                        out[name][v] = int(text.split(":", 1)[1])
    return out


HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")
LIB_H5 = os.path.join(PKG, "libns_h5.so")


def build_h5(verbose: bool = False) -> str | None:
    """libns_h5.so: the reference's HDF5 files from C (host only, g++ against
    the HDF5 C + high-level libraries).  Skipped when HDF5 is not installed."""
    inc = os.path.join(HDF5_DIR, "include")
    libdir = os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5_hl.h")):
        if verbose:
            print(f"HDF5 not found under {HDF5_DIR}: libns_h5.so not built")
        return None
    tmp = LIB_H5 + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"), "-I", inc,
           os.path.join(CSRC, "ns_h5.cpp"), "-L", libdir, "-lhdf5_hl", "-lhdf5", f"-Wl,-rpath,{libdir}", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libns_h5 build failed:\n{r.stderr[-4000:]}")
    shutil.move(tmp, LIB_H5)
    if verbose:
        print(f"built {LIB_H5}")
    return LIB_H5


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = SOURCES + EXTRA_SOURCES
    with cf.ThreadPoolExecutor(max_workers=min(len(srcs), 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    shutil.move(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_h5(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
