#!/usr/bin/env python3
"""Benchmark of the MI355X explicit finite-volume step.

Metric: FP64 cell-updates per second per RK stage (interior cells x RK stages
x time steps / wall time of the stepping loop), plus Riemann main-loop
iterations per interface face.  One "step" is one time step of the driver loop
(Main.cpp:653-712): the CFL time step, all RK stages (interface Riemann
solves, fused residual/RK kernels, level-set stages) and the end-of-step
material update, with the state resident in HBM.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1]
Prints one JSON line.  Only one GPU is used; --gpus N > 1 exits.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from m2c_amd import abi, configs  # noqa: E402

HBM_PEAK_GBPS = 8000.0


def stage_bytes(integrator, multimat):
    """Algorithmic HBM bytes per interior cell for each launch of the fused
    stage kernel: stage 1 reads V (40 B) and writes Un, U1, V1 (120 B);
    intermediate RK3 stage reads V, U1, Un and writes U1, V2 (200 B); final
    stages read V, U1, Un and write V (160 B).  +1 B/cell for the material ID
    when more than one material is present."""
    idb = 1 if multimat else 0
    if integrator == abi.TI_RK3:
        per = [160 + idb, 200 + idb, 160 + idb]
    elif integrator == abi.TI_RK2:
        per = [160 + idb, 160 + idb]
    else:
        per = [120 + idb]
    return per


def build_config(name):
    if name == "c1":
        return configs.c1_sod(), "3D Sod shock tube 128x32x32, SG gamma 1.4, MUSCL van Albada primitive, HLLC, RK2, CFL 0.5"
    if name == "c2":
        return configs.c2_shock_bubble(256), ("two-material shock-bubble 256^3: SG water (6.12, 3.43e8) + SG air "
                                              "bubble r=0.15L, Mach 1.5 shock, exact interface Riemann tol 1e-10, "
                                              "HLLC, van Albada, RK3, CFL 0.4, level set FD3")
    if name == "c2_single":
        g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(256)
        # same grid and state, bubble replaced by water (single material)
        water = ID == 1
        V[water] = [1000.0, 0.0, 0.0, 0.0, 1.0e5]
        ID[:] = 0
        s.n_levelsets = 0
        mats1 = abi.materials(mats[0])
        return (g, mats1, 1, s, V, ID, None), "single-material 256^3 shock in SG water, HLLC, van Albada, RK3"
    raise SystemExit("unknown config " + name)


def cpu_baseline(name, budget_s=20.0):
    """Time the CPU oracle (this project's restatement of the reference's
    step, single thread) on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    if name == "c1":
        (g, mats, nmat, s, V, ID, Phi), _ = build_config("c1")
        sample = "config 1 (128x32x32), first %d steps"
    else:
        g, mats, nmat, s, V, ID, Phi = configs.c2_shock_bubble(64)
        if name == "c2_single":
            V[ID == 1] = [1000.0, 0.0, 0.0, 0.0, 1.0e5]
            ID[:] = 0
            s.n_levelsets = 0
            mats = abi.materials(mats[0])
            nmat = 1
            Phi = None
        sample = "same configuration at 64^3 cells, first %d steps"
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, Phi)
    stages = {abi.TI_FE: 1, abi.TI_RK2: 2, abi.TI_RK3: 3}[s.integrator]
    cells = g.nx * g.ny * g.nz
    t0 = time.perf_counter()
    steps = 0
    while True:
        o.run_steps(1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 100:
            break
    value = cells * stages * steps / el
    return {"value": value, "unit": "cell-stage-updates/s", "cores": 1, "kind": "port",
            "sample": sample % steps, "cpu": platform.processor() or platform.machine()}


def measured_traffic(workload):
    """HBM bytes per k_stage launch measured by rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE) for this workload, scaled by the same counters on a copy of
    known size; kept as a `bytes_per_launch: N` line in
    profiles/k_stage_<workload>_traffic.txt.  None when no such measurement
    exists."""
    path = os.path.join(ROOT, "profiles", "k_stage_%s_traffic.txt" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def copy_probe(nbytes):
    """A plain device copy of `nbytes` (read nbytes, write nbytes), the
    calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default=os.environ.get("M2C_BENCH_CONFIG", "c2"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build does not run the step across GPUs "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    from m2c_amd import runtime
    steps = args.steps if args.steps is not None else (100 if args.config == "c1" else 20)
    (g, mats, nmat, s, V, ID, Phi), desc = build_config(args.config)
    print("# bench: config=%s cells=%d steps=%d warmup=%d" % (args.config, g.nx * g.ny * g.nz, steps, args.warmup),
          file=sys.stderr, flush=True)
    solver = runtime.Solver(g, mats, nmat, s)
    solver.set_state(V, ID, Phi)
    del V
    stages = {abi.TI_FE: 1, abi.TI_RK2: 2, abi.TI_RK3: 3}[s.integrator]
    cells = g.nx * g.ny * g.nz
    solver.run_steps(args.warmup)
    solver.synchronize()
    solver.reset_counters()
    solver.profile(True)
    t0 = time.perf_counter()
    solver.run_steps(steps)
    solver.synchronize()
    el = time.perf_counter() - t0
    cnt = solver.counters()
    stage_ms, stage_n = solver.kernel_time("stage")
    riem_ms, riem_n = solver.kernel_time("riemann")
    per = stage_bytes(s.integrator, nmat > 1)
    avg_bytes = cells * sum(per) / len(per)
    avg_stage_ms = stage_ms / max(stage_n, 1)
    achieved = avg_bytes / (avg_stage_ms * 1e-3) / 1e9
    value = cells * stages * steps / el
    iters_per_face = cnt["riemann_main_iters"] / max(cnt["riemann_calls"], 1)
    fevals_per_face = cnt["riemann_fevals"] / max(cnt["riemann_calls"], 1)
    out = {
        "metric": "FP64 cell-updates/s per RK stage + Riemann Newton iters/face, learned vs plain",
        "value": value,
        "unit": "cell-stage-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.config, "description": desc, "cells": cells, "rk_stages": stages,
                   "riemann_guess": "acoustic (plain)"},
        "riemann": {"faces": cnt["riemann_calls"], "trivial": cnt["riemann_trivial"],
                    "main_iters_per_face": iters_per_face, "fevals_per_face": fevals_per_face,
                    "errors": cnt["riemann_errors"], "riemann_ms_per_step": riem_ms / steps},
        "roofline": {"bound": "hbm", "kernel": "k_stage (fused residual + RK stage)",
                     "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": measured_traffic(args.config),
                     "bytes_per_cell_per_launch": per, "avg_launch_ms": avg_stage_ms},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.config)
    solver.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
