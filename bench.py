#!/usr/bin/env python3
"""Benchmark: FP64 Mcell-steps/s of the decaying-turbulence time step
(big_periodic_3D path: AB2 convection/diffusion + FFT pressure Poisson +
projection, triply periodic, 256^3, nu=1/1000, dt=5e-4) on one MI355X.

A step is one full time step of the whole 256^3 field; inputs are resident
in HBM before the timed region.  The per-kernel roofline numbers come from
HIP events recorded on the solver's own stream (ns_step_profiled) over the
same number of steps.  The CPU baseline times the oracle (the repository's
C restatement of the reference step, OpenMP) on a short sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic FP64 arrays streamed per interior cell by each phase (one read or
# write of one full array = 8 B/cell; the r2c half spectrum counts as one real array).
# With dp_flag = 0 the projection is folded into the next predictor, which then
# reads u*, v*, w*, dp and the three AB2 histories and writes u*, v*, w* and the
# histories (13 arrays); no separate projection pass runs inside the step.
PHASES = ["predictor", "vel_bc", "div_fft_x_fwd", "fft_y_fwd", "fft_z_solve", "fft_y_inv", "fft_x_inv_dp",
          "pr_bc", "project"]
ARRAYS_PER_CELL = {"predictor": 13, "vel_bc": 0, "div_fft_x_fwd": 4, "fft_y_fwd": 2, "fft_z_solve": 2,
                   "fft_y_inv": 2, "fft_x_inv_dp": 2, "pr_bc": 0, "project": 0}
# of which read
ARRAYS_READ = {"predictor": 7, "div_fft_x_fwd": 3, "fft_y_fwd": 1, "fft_z_solve": 1, "fft_y_inv": 1, "fft_x_inv_dp": 1}
# the reference pipeline's budget (explicit projection pass): 32 arrays = 256 B/cell-step
REFERENCE_BUDGET_B = 256
KERNEL_NAMES = {"predictor": "k_predict_tiled<true,64>", "div_fft_x_fwd": "k_fft_x_fwd_rs<false,512,1>",
                "fft_y_fwd": "k_fft_sq<1,0,16,8>", "fft_z_solve": "k_fft_sq<2,2,16,8>",
                "fft_y_inv": "k_fft_sq<1,1,16,8>", "fft_x_inv_dp": "k_fft_x_inv<false,1024,true>", "project": "k_project"}
COPY_CAL_BYTES = 2 ** 30  # bytes each way of the streaming copy (ns_stream_copy) that scales the traffic counters
CAL_KERNEL = "k_stream_copy"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic_256.json"))
    ap.add_argument("--adjoint-n", type=int, default=128,
                    help="grid of the adjoint/forward cost measurement (walls, AB2, DCT, dp_flag=1); 0 skips it")
    ap.add_argument("--resim-scheme", choices=["AB2-CN", "AB2"], default="AB2-CN",
                    help="time scheme of the adjoint and batched re-simulation measurements (walls)")
    ap.add_argument("--batch-boxes", type=int, default=64,
                    help="sub-boxes of the batched re-simulation measurement (sizes 32..128, seed 4); 0 skips it")
    ap.add_argument("--resim-only", action="store_true",
                    help="run only the adjoint / batch measurements (for profiling them) and print them as one JSON line")
    ap.add_argument("--pipeline", action="store_true",
                    help="run the re-simulation pipeline end to end (a 512^3 periodic AB2 parent capturing a 128^3 "
                         "sub-box every 4th step for 100 steps, then the sub-box re-simulated with walls, AB2-CN, "
                         "dt/4, 400 sub-steps, planes linearly interpolated), print its timings and errors as one "
                         "JSON line and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device-to-device copy (scale for the FETCH_SIZE/WRITE_SIZE counters) and exit")
    ap.add_argument("--summarize-profile", metavar="DIR",
                    help="(no GPU) turn rocprofv3 output under DIR (kt/, FETCH_SIZE/, WRITE_SIZE/, cal_FETCH_SIZE/, "
                         "cal_WRITE_SIZE/) into markdown tables (the 256^3 step's kernels against their algorithmic "
                         "bytes, written to the traffic JSON when present; every kernel's counter traffic), then exit")
    return ap.parse_args()


def short_kernel_name(name):
    """'void ns::(anonymous namespace)::k_fft_yz<1, -1, 0, false>(ns::...)' -> 'k_fft_yz<1,-1,0,false>'"""
    s = name.strip('"')
    if s.startswith("void "):
        s = s[5:]
    s = s.replace("ns::(anonymous namespace)::", "")
    depth, out = 0, []
    for ch in s:
        if ch == "(" and depth == 0:
            break
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        out.append(ch)
    return "".join(out).replace(" ", "")


def summarize_profile(d, traffic_json):
    """Per-kernel traffic between the L2s and the memory fabric (Infinity Cache hits
    included) from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; KiB per dispatch),
    each scaled so that the library's streaming copy of 1 GiB reads and writes 1 GiB."""
    import csv
    import glob

    def counter(sub, cname):
        per = {}
        for f in glob.glob(os.path.join(d, sub, "*counter_collection.csv")):
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] == cname:
                    per.setdefault(short_kernel_name(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in per.items()}

    fetch, write = counter("FETCH_SIZE", "FETCH_SIZE"), counter("WRITE_SIZE", "WRITE_SIZE")
    cf, cw = counter("cal_FETCH_SIZE", "FETCH_SIZE"), counter("cal_WRITE_SIZE", "WRITE_SIZE")
    f_fac = (COPY_CAL_BYTES / 1024) / cf[CAL_KERNEL]
    w_fac = (COPY_CAL_BYTES / 1024) / cw[CAL_KERNEL]
    stats = {}
    for f in glob.glob(os.path.join(d, "kt", "*kernel_stats.csv")):
        for r in csv.DictReader(open(f)):
            stats[short_kernel_name(r["Name"])] = (int(r["Calls"]), float(r["AverageNs"]) / 1e3)
    n3 = 256 ** 3
    traffic, rows = {}, []
    # the algorithmic table is the 256^3 headline step's (its lazy predictor is in the run)
    headline = KERNEL_NAMES["predictor"] in fetch
    for phase, kname in KERNEL_NAMES.items():
        if not headline or kname not in fetch or kname not in write:
            continue
        rd, wr = fetch[kname] * f_fac * 1024, write[kname] * w_fac * 1024
        traffic[kname] = int(rd + wr)
        alg = ARRAYS_PER_CELL.get(phase, 0) * 8 * n3 if phase != "project" else 8 * 8 * n3
        calls, us = stats.get(kname, (0, float("nan")))
        nr, nw = ARRAYS_READ.get(phase, 0), ARRAYS_PER_CELL.get(phase, 0) - ARRAYS_READ.get(phase, 0)
        if phase == "project":
            nr, nw = 4, 4
        ar, aw = nr * 8 * n3, nw * 8 * n3
        rows.append(f"| {kname} | {calls} | {us:.1f} | {alg / 1e6:.0f} | {rd / 1e6:.0f} | {wr / 1e6:.0f} | "
                    f"{rd / ar:.2f} | {wr / aw:.2f} | {alg / (us * 1e-6) / 1e9:.0f} | {(rd + wr) / (us * 1e-6) / 1e9:.0f} |")
    if rows:
        json.dump(traffic, open(traffic_json, "w"), indent=1)
    print(f"FETCH_SIZE factor {f_fac:.3f}, WRITE_SIZE factor {w_fac:.3f} ({CAL_KERNEL}, 1 GiB each way)")
    if rows:
        print("| kernel | calls | avg us | alg MB/launch | L2-fabric read MB | L2-fabric write MB | read/alg | "
              "write/alg | alg GB/s | L2-fabric GB/s |")
        print("|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
        print("\n".join(rows))
        print()
    # every kernel of the run: counter traffic per launch and the rate it implies
    print("| kernel (all launches of the run) | calls | avg us | L2-fabric read MB | L2-fabric write MB | "
          "L2-fabric GB/s |")
    print("|---|---:|---:|---:|---:|---:|")
    allk = []
    for kname in set(fetch) & set(write):
        if kname == CAL_KERNEL or "fillBuffer" in kname or kname not in stats:
            continue
        rd, wr = fetch[kname] * f_fac * 1024, write[kname] * w_fac * 1024
        calls, us = stats[kname]
        allk.append((calls * us, f"| {kname} | {calls} | {us:.1f} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | "
                                 f"{(rd + wr) / (us * 1e-6) / 1e9:.0f} |"))
    print("\n".join(r for _, r in sorted(allk, reverse=True)))
    return 0


def adjoint_cost(n, scheme_name="AB2-CN", steps=10):
    """Adjoint/forward cost ratio on the re-simulation setting (walls bc=2,
    Neumann pressure pbc=3, AB2-CN or AB2, dp_flag=1, n^3): mean ms of one forward step
    (ns_step) and of one reverse step (ns_adjoint_step, forward velocity
    already on the device); the checkpointed reverse sweep adds one forward
    recompute per step on top of that."""
    import ctypes

    import torch

    from ns_3d_staggered_dp_amd import SCHEMES, NSConfig, fields
    from ns_3d_staggered_dp_amd import _abi
    from ns_3d_staggered_dp_amd.grid import plane_set_length
    from ns_3d_staggered_dp_amd.solver import NSSolver

    cfg = NSConfig.cube(n, nu=1e-3, dt=5e-4, scheme=SCHEMES[scheme_name], bc=(2, 2, 2), pbc=(3, 3, 3), dp_flag=1)
    u, v, w, p = fields.decaying_turbulence(cfg, seed=5)
    rng = np.random.default_rng(6)
    L = plane_set_length(n, n, n)
    with NSSolver(cfg) as s:
        s.upload(u, v, w, p, first=True)
        s.set_bc_planes(0.01 * rng.standard_normal(L))
        s.step(3)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        fwd = (time.perf_counter() - t0) / steps * 1e3
        us, vs, ws = (torch.empty(int(x), dtype=torch.float64, device="cuda") for x in
                      ((n + 1) * (n + 2) ** 2, (n + 2) * (n + 1) * (n + 2), (n + 2) ** 2 * (n + 1)))
        _abi.call("ns_download_state", s.h, *(ctypes.c_void_p(t.data_ptr()) for t in (us, vs, ws)), None, None, None,
                  None)
        ub = torch.from_numpy(np.asfortranarray(u).ravel(order="F")).cuda()
        _abi.call("ns_adjoint_set_state", s.h, ctypes.c_void_p(ub.data_ptr()), None, None, None, None, None, None)
        pb = torch.zeros(L, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()  # the context stream does not wait for torch's
        args = [ctypes.c_void_p(t.data_ptr()) for t in (us, vs, ws)]
        _abi.call("ns_adjoint_step", s.h, *args, 0, ctypes.c_void_p(pb.data_ptr()))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            _abi.call("ns_adjoint_step", s.h, *args, 0, ctypes.c_void_p(pb.data_ptr()))
        torch.cuda.synchronize()
        adj = (time.perf_counter() - t0) / steps * 1e3
    return {"config": f"{n}^3 walls bc=2 pbc=3, {scheme_name}, DCT Poisson, dp_flag=1", "forward_ms": round(fwd, 4),
            "adjoint_ms": round(adj, 4), "ratio": round(adj / fwd, 3),
            "ratio_with_recompute": round((adj + fwd) / fwd, 3)}


def batch_throughput(nbox, scheme_name="AB2-CN", steps=20):
    """Batched re-simulation of nbox sub-boxes of uneven sizes (walls bc=2,
    pbc=3, AB2-CN or AB2, DCT, dp_flag=1; sizes drawn per axis from the 7-smooth lengths
    in [32, 128], seed 4): Mcell-steps/s of ns_batch_step over all boxes, and
    of the same boxes stepped one after another, each with its own launches
    (eager: every kernel launched per step; graph: steady steps replayed from
    the box's own HIP graph).  The batched rate is measured with the default
    launch shapes and again after ns_batch_tune."""
    from ns_3d_staggered_dp_amd import FLAG_NO_GRAPH, SCHEMES, NSConfig, fields
    from ns_3d_staggered_dp_amd.batch import Batch, box_sizes
    from ns_3d_staggered_dp_amd.grid import plane_set_length
    from ns_3d_staggered_dp_amd.solver import NSSolver

    sizes = box_sizes(nbox, 32, 128, seed=4)
    cells = sum(nx * ny * nz for nx, ny, nz in sizes)

    def make(flags):
        out = []
        for i, (nx, ny, nz) in enumerate(sizes):
            h = 2 * np.pi / 256
            cfg = NSConfig(nx=nx, ny=ny, nz=nz, dx=h, dy=h, dz=h, nu=1e-3, dt=5e-4, scheme=SCHEMES[scheme_name],
                           bc=(2, 2, 2),
                           pbc=(3, 3, 3), dp_flag=1, flags=flags)
            s = NSSolver(cfg)
            out.append(s)
            u, v, w, p = fields.random_ghosted(cfg, 400 + i)
            s.upload(0.1 * u, 0.1 * v, 0.1 * w, p, first=True)
            s.set_bc_planes(0.01 * np.random.default_rng(500 + i).standard_normal(plane_set_length(nx, ny, nz)))
        return out

    def sequential(solvers):
        for s in solvers:
            s.step(3)
            s.synchronize()
        t0 = time.perf_counter()
        for s in solvers:
            s.step(steps)
            s.synchronize()
        return time.perf_counter() - t0

    res, knobs = {}, None
    for name, flags in (("eager", FLAG_NO_GRAPH), ("graph", 0)):
        solvers = []
        try:
            solvers = make(flags)
            res[name] = sequential(solvers)
            if flags == 0:
                with Batch(solvers) as b:
                    b.step(3)
                    t0 = time.perf_counter()
                    b.step(steps)
                    res["batch_default"] = time.perf_counter() - t0
                    knobs = b.tune()
                    b.step(1)
                    t0 = time.perf_counter()
                    b.step(steps)
                    res["batch"] = time.perf_counter() - t0
        finally:
            for s in solvers:
                s.close()
    rate = {k: round(cells * steps / v / 1e6, 1) for k, v in res.items()}
    return {"config": f"{nbox} sub-boxes, sizes 32..128 per axis (7-smooth, seed 4), walls bc=2 pbc=3, "
                      f"{scheme_name}, DCT, dp_flag=1", "cells": cells, "steps": steps,
            "batched_mcell_steps_s": rate["batch"], "batched_default_shapes_mcell_steps_s": rate["batch_default"],
            "tuned_shapes": knobs, "sequential_graph_mcell_steps_s": rate["graph"],
            "sequential_eager_mcell_steps_s": rate["eager"], "speedup_vs_graph": round(res["graph"] / res["batch"], 3),
            "speedup_vs_eager": round(res["eager"] / res["batch"], 3)}


def resim_pipeline(N=512, sub=(101, 201, 151, 128, 128, 128), dt0=5e-4, every=4, parent_steps=100):
    """The re-simulation workload end to end on one device: the parent
    (decaying turbulence, big_periodic_3D.f90:107's dp_flag 0, corner gauge)
    captures the sub-box's plane set every `every` steps
    (big_periodic_3D.f90:433-489,664-712); the sub-box starts from the parent's
    state at the first capture and runs `every` * 4 sub-steps per capture
    interval at dt0/4 with the planes interpolated linearly in time
    (re_simulation_sub_tstep_interBC_combine.f90:388-396), and is compared
    with the parent's sub-box at the end (the reference's error table,
    :826-833)."""
    import ctypes
    import math

    import torch

    from ns_3d_staggered_dp_amd import AB2, AB2_CN, FLAG_GAUGE_CORNER, NSConfig, _abi, fields
    from ns_3d_staggered_dp_amd.grid import plane_set_length
    from ns_3d_staggered_dp_amd.solver import NSSolver

    x0, y0, z0, a, b, c = sub
    h = 2 * np.pi / N
    cfg = NSConfig(nx=N, ny=N, nz=N, dx=h, dy=h, dz=h, nu=1e-3, dt=dt0, scheme=AB2, bc=(1, 1, 1), pbc=(1, 1, 1),
                   dp_flag=0, flags=FLAG_GAUGE_CORNER)
    L = plane_set_length(a, b, c)
    nsets = parent_steps // every + 1
    buf = torch.zeros(nsets * L, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    keys = ("u", "v", "w", "p", "hx", "hy", "hz")
    u, v, w, p = fields.decaying_turbulence(cfg, seed=2)
    with NSSolver(cfg) as s:
        s.upload(u, v, w, p, first=True)
        del u, v, w, p
        s.step(3)
        sb = _abi.SubBox(x0, y0, z0, a, b, c)
        _abi.call("ns_capture_begin", s.h, ctypes.byref(sb), ctypes.c_void_p(buf.data_ptr()), nsets, every)
        s.step(every)  # the first capture: the sub-box's start
        start = {k: s.extract_subbox(x0, y0, z0, a, b, c, k) for k in keys}
        s.synchronize()
        t0 = time.perf_counter()
        s.step(parent_steps)
        s.synchronize()
        tpar = time.perf_counter() - t0
        cnt = ctypes.c_int()
        _abi.call("ns_capture_end", s.h, ctypes.byref(cnt))
        end = {k: s.extract_subbox(x0, y0, z0, a, b, c, k) for k in ("u", "v", "w", "p")}
    snaps = buf.view(nsets, L)
    scfg = NSConfig(nx=a, ny=b, nz=c, dx=h, dy=h, dz=h, nu=cfg.nu, dt=dt0 / 4, scheme=AB2_CN, bc=(2, 2, 2),
                    pbc=(3, 3, 3), dp_flag=1)
    nsub = parent_steps * 4
    with NSSolver(scfg) as r:
        r.upload(*(start[k] for k in keys), first=False)
        r.synchronize()
        t0 = time.perf_counter()
        for i in range(1, nsub + 1):
            t = i * dt0 / 4
            k = min(int(math.floor(t / (every * dt0))), nsets - 2)
            r.bc_interp(snaps[k], snaps[k + 1], (t - k * every * dt0) / (every * dt0))
            r.step(1)
        r.synchronize()
        tre = time.perf_counter() - t0
        err, _ = r.error_metrics(end["u"], end["v"], end["w"], end["p"])
    return {"parent": f"{N}^3 periodic decaying turbulence, AB2, FFT, dp_flag 0, dt={dt0}, capture every {every}",
            "parent_steps": parent_steps, "parent_ms_per_step": round(tpar / parent_steps * 1e3, 3),
            "parent_mcell_steps_s": round(N ** 3 * parent_steps / tpar / 1e6, 1), "captured_sets": cnt.value,
            "sub_box": f"{a}x{b}x{c} at ({x0},{y0},{z0}), walls bc=2 pbc=3, AB2-CN, DCT, dp_flag 1, dt={dt0 / 4}",
            "sub_steps": nsub, "sub_ms_per_step": round(tre / nsub * 1e3, 4),
            "sub_mcell_steps_s": round(a * b * c * nsub / tre / 1e6, 1), "sub_total_ms": round(tre * 1e3, 1),
            "err_max_rel_uvw": [float(x) for x in err[:3]], "err_mean_rel_uvw": [float(x) for x in err[4:7]]}


def copy_calibration():
    """three launches of the library's streaming copy (k_stream_copy) of 1 GiB:
    a hand-written kernel whose traffic is known, so the FETCH_SIZE and
    WRITE_SIZE counters it reads give their scale"""
    import ctypes

    import torch

    from ns_3d_staggered_dp_amd import NSConfig, _abi
    from ns_3d_staggered_dp_amd.solver import NSSolver

    a = torch.ones(COPY_CAL_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    with NSSolver(NSConfig.cube(8)) as s:
        for _ in range(3):
            _abi.call("ns_stream_copy", s.h, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), a.numel())
    torch.cuda.synchronize()
    print("copy calibration done", float(b[0]), flush=True)
    return 0


def host_cpu():
    """model name and core counts of this host (lscpu; /proc/cpuinfo as a fallback)"""
    info = {"logical_cpus": os.cpu_count()}
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        sockets = int(kv.get("Socket(s)", "1"))
        cores = int(kv.get("Core(s) per socket", "0"))
        info.update(sockets=sockets, cores_per_socket=cores, threads_per_core=int(kv.get("Thread(s) per core", "1")),
                    physical_cores=sockets * cores)
    except Exception:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return info


def cpu_baseline(cfg, u, v, w, p, nsteps):
    from oracle import oracle as O

    st = O.OracleState(cfg, u, v, w, p, first=True)
    st.step()  # untimed Euler first step
    t0 = time.perf_counter()
    for _ in range(nsteps):
        st.step()
    dt = time.perf_counter() - t0
    n3 = cfg.nx * cfg.ny * cfg.nz
    return {
        "value": n3 * nsteps / dt / 1e6,
        "unit": "Mcell-steps/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": f"{nsteps} AB2 steps of the same {cfg.nx}^3 field after 1 untimed Euler step "
                  f"({dt:.1f} s, oracle/ns_oracle.c, OpenMP)",
        "host": host_cpu(),
    }


def main():
    a = parse()
    if a.summarize_profile:
        return summarize_profile(a.summarize_profile, a.traffic_json)
    if a.copy_calibration:
        return copy_calibration()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): the multi-GPU decomposition (slabs with "
              "halo exchange and transform all-to-alls, whole sub-boxes per GPU) is not built.", file=sys.stderr)
        return 2
    import torch

    from ns_3d_staggered_dp_amd import AB2, FLAG_GAUGE_CORNER, NSConfig, fields
    from ns_3d_staggered_dp_amd import _abi
    from ns_3d_staggered_dp_amd.solver import NSSolver

    if a.pipeline:
        print(json.dumps({"pipeline": resim_pipeline()}), flush=True)
        return 0
    if a.resim_only:
        out = {}
        if a.adjoint_n > 0:
            out["adjoint"] = adjoint_cost(a.adjoint_n, a.resim_scheme)
        if a.batch_boxes > 0:
            out["batch"] = batch_throughput(a.batch_boxes, a.resim_scheme)
        print(json.dumps(out), flush=True)
        return 0

    n = a.n
    cfg = NSConfig.cube(n, nu=1e-3, dt=5e-4, scheme=AB2, flags=FLAG_GAUGE_CORNER)
    u, v, w, p = fields.decaying_turbulence(cfg, seed=2)
    s = NSSolver(cfg)
    s.upload(u, v, w, p, first=True)
    s.step(a.warmup)
    s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(a.steps)
    s.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    n3 = n ** 3
    ms_step = elapsed / a.steps * 1e3
    value = n3 * a.steps / elapsed / 1e6
    finite = s.check_finite()
    maxdiv = s.max_div()

    # per-phase GPU time on the solver stream, same number of steps
    import ctypes

    ph = (ctypes.c_double * 9)()
    _abi.call("ns_step_profiled", s.h, a.steps, ph)
    phase_ms = {k: ph[i] / a.steps for i, k in enumerate(PHASES)}
    dom = max((k for k in PHASES if ARRAYS_PER_CELL[k] > 0), key=lambda k: phase_ms[k])
    bytes_launch = ARRAYS_PER_CELL[dom] * 8 * n3
    achieved = bytes_launch / (phase_ms[dom] * 1e-3) / 1e9
    traffic = None
    if os.path.exists(a.traffic_json):
        try:
            tj = json.load(open(a.traffic_json))
            traffic = tj.get(KERNEL_NAMES.get(dom, dom))
        except Exception:
            traffic = None
    total_alg = sum(ARRAYS_PER_CELL.values()) * 8
    step_gbs = total_alg * n3 / (ms_step * 1e-3) / 1e9
    s.close()

    line = {
        "metric": "FP64 Mcell-steps/s at 256^3 + % HBM roofline; adjoint/forward cost ratio",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: random solenoidal field, E(k)~k^4 exp(-2(k/4)^2), seed 2, energy 0.5",
        "config": {"workload": f"decaying turbulence {n}^3 periodic, AB2 + FFT Poisson, nu=1e-3, dt=5e-4",
                   "grid": [n, n, n], "scheme": "AB2", "poisson": "FFT", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": bytes_launch, "avg_ms": round(phase_ms[dom], 4)},
        "step_roofline": {"alg_bytes_per_cell_step": total_alg, "achieved_GBs": round(step_gbs, 1),
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4),
                          "frac_of_256B_pipeline_bound": round(REFERENCE_BUDGET_B * value * 1e6 / (HBM_PEAK_GBS * 1e9), 4)},
        "phase_ms": {k: round(v, 4) for k, v in phase_ms.items()},
        # each pass's algorithmic bytes (arrays x 8 B x n^3) over its HIP-event time, as a fraction of 8 TB/s
        "phase_frac": {k: round(ARRAYS_PER_CELL[k] * 8 * n3 / (phase_ms[k] * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)
                       for k in PHASES if ARRAYS_PER_CELL.get(k, 0) > 0 and phase_ms[k] > 0},
        "checks": {"finite": bool(finite), "max_div": maxdiv},
    }
    if a.adjoint_n > 0:
        line["adjoint"] = adjoint_cost(a.adjoint_n, a.resim_scheme)
    if a.batch_boxes > 0:
        line["batch"] = batch_throughput(a.batch_boxes, a.resim_scheme)
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cfg, u, v, w, p, a.cpu_steps)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
