#!/usr/bin/env python3
"""Benchmark of the MI355X explicit finite-volume step.

Metric: FP64 cell-updates per second per RK stage (interior cells x RK stages
x time steps / wall time of the stepping loop), plus Riemann main-loop
iterations per interface face, for the plain (acoustic) and the learned
initial guess.  One "step" is one time step of the driver loop
(Main.cpp:653-712): the CFL time step, all RK stages (interface Riemann
solves, fused residual/RK kernels, level-set stages) and the end-of-step
material update, with the state resident in HBM.

Workloads (--config):
  c2          256^3 two-material shock-bubble, RK3 (default; the headline)
  c1          128x32x32 Sod tube, RK2
  c2_single   the c2 grid with water only (single-material residual)
  c4          uneven ensemble of shock-bubble runs (edges 64..192, --instances)
              advanced in lockstep with one captured graph per step
  c5          bubble cloud: 512x256x256, NASG water, 200 air bubbles, Mach-2
              shock (initialised on the device; no level-set reinitialisation)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]
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
METRIC = "FP64 cell-updates/s per RK stage + Riemann Newton iters/face, learned vs plain"
GUESS_PARAMS = os.path.join(ROOT, "m2c_amd", "data", "guess_params.npy")


def stage_bytes(integrator, multimat):
    """Algorithmic HBM bytes per interior cell for each launch of the fused
    stage kernel: stage 1 reads V (40 B) and writes Un, U1, V1 (120 B);
    intermediate RK3 stage reads V, U1, Un and writes U1, V2 (200 B); final
    stages read V, U1, Un and write V (160 B).  +1 B/cell for the material ID
    when more than one material is present."""
    idb = 1 if multimat else 0
    if integrator == abi.TI_RK3:
        return [160 + idb, 200 + idb, 160 + idb]
    if integrator == abi.TI_RK2:
        return [160 + idb, 160 + idb]
    return [120 + idb]


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
    if name == "c5":
        return configs.c5_bubble_cloud(), ("bubble cloud 512x256x256: NASG water (1.19, 6.2178e8, 6.7212e-4, "
                                           "-1.177788e6) with 200 air bubbles r 0.005-0.03 L, Mach-2 shock, exact "
                                           "interface Riemann tol 1e-10, HLLC, van Albada, RK3, CFL 0.4, level set "
                                           "FD3 reinitialised every 10 steps (full domain, Fixed first layer, "
                                           "30 iterations at most)")
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
        sample = ("the same configuration at 64^3 cells, first %d steps" if name != "c5" else
                  "the two-material shock-bubble configuration at 64^3 cells, first %d steps")
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


def _quantile(hist, q):
    """Smallest count k with at least a fraction q of the faces at <= k."""
    c = np.cumsum(hist.astype(np.float64))
    if c[-1] == 0:
        return 0
    return int(np.searchsorted(c, q * c[-1]))


def riemann_stats(cnt, hist=None):
    calls = max(cnt["riemann_calls"], 1)
    out = {"faces": cnt["riemann_calls"], "trivial": cnt["riemann_trivial"],
           "main_iters_per_face": cnt["riemann_main_iters"] / calls,
           "fevals_per_face": cnt["riemann_fevals"] / calls,
           "errors": cnt["riemann_errors"], "guess_fallbacks": cnt["riemann_guess_fallbacks"]}
    if hist is not None:
        it, fe = hist
        out.update({"main_iters_p50": _quantile(it, 0.5), "main_iters_p99": _quantile(it, 0.99),
                    "fevals_p50": _quantile(fe, 0.5), "fevals_p99": _quantile(fe, 0.99)})
    return out


def run_single(cfg, steps, warmup, learned_params=None):
    """Time `steps` steps of one solver; returns (result dict, solver)."""
    from m2c_amd import runtime
    (g, mats, nmat, s, V, ID, Phi) = cfg
    if learned_params is not None:
        s.riem_guess = abi.GUESS_MLP
    solver = runtime.Solver(g, mats, nmat, s)
    if learned_params is not None:
        solver.mlp_set_params(learned_params)
    if isinstance(V, tuple):  # (default state, default id, geometries): device initialisation
        solver.initialize(V, ID, Phi)
    else:
        solver.set_state(V, ID, Phi)
    stages = {abi.TI_FE: 1, abi.TI_RK2: 2, abi.TI_RK3: 3}[s.integrator]
    cells = g.nx * g.ny * g.nz
    solver.run_steps(warmup)
    solver.synchronize()
    solver.reset_counters()
    solver.profile(True)
    t0 = time.perf_counter()
    solver.run_steps(steps)
    solver.synchronize()
    el = time.perf_counter() - t0
    cnt = solver.counters()
    hist = solver.riemann_histogram()
    stage_ms, stage_n = solver.kernel_time("stage")
    # with interfaces the stage kernel runs in two launches per stage (tiles
    # without interface faces during the Riemann solve, then the others)
    sif_ms, _ = solver.kernel_time("stage_if")
    stage_ms += sif_ms
    standalone = None
    if nmat > 1:
        # during the timed steps the first stage launch shares the CUs with the
        # Riemann kernel; the kernel's own launch time comes from a few more
        # steps with the overlap off (one launch per stage after the solve)
        solver.set_overlap(False)
        ms0, n0 = solver.kernel_time("stage")
        solver.run_steps(min(10, steps))
        ms1, n1 = solver.kernel_time("stage")
        solver.set_overlap(True)
        standalone = (ms1 - ms0, n1 - n0)
    riem_ms, _ = solver.kernel_time("riemann")
    mlp_ms, _ = solver.kernel_time("mlp")
    reinit_ms, _ = solver.kernel_time("reinit")
    res = {"value": cells * stages * steps / el, "ms_per_step": el * 1e3 / steps, "cells": cells,
           "stages": stages, "stage_ms": stage_ms, "stage_launches": stage_n, "stage_standalone": standalone,
           "riemann": dict(riemann_stats(cnt, hist), riemann_ms_per_step=riem_ms / steps,
                           mlp_ms_per_step=mlp_ms / steps)}
    if cnt["reinit_calls"]:
        res["reinit"] = {"calls": cnt["reinit_calls"], "iterations": cnt["reinit_iterations"],
                         "ms_per_call": reinit_ms / cnt["reinit_calls"], "ms_per_step": reinit_ms / steps}
    solver.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default=os.environ.get("M2C_BENCH_CONFIG", "c2"))
    ap.add_argument("--instances", type=int, default=64, help="ensemble size for --config c4")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
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
    if args.config == "c4":
        return bench_ensemble(args)
    # the configurations' own run lengths: 100 steps (c1), 200 (c2, c2_single); c5's 300 are cut to 50
    steps = args.steps if args.steps is not None else {"c1": 100, "c2": 200, "c2_single": 200}.get(args.config, 50)
    cfg, desc = build_config(args.config)
    g, mats, nmat, s = cfg[:4]
    print("# bench: config=%s cells=%d steps=%d warmup=%d" % (args.config, g.nx * g.ny * g.nz, steps, args.warmup),
          file=sys.stderr, flush=True)
    plain = run_single(cfg, steps, args.warmup)
    per = stage_bytes(s.integrator, nmat > 1)
    cells = plain["cells"]
    avg_bytes = cells * sum(per) / len(per)
    live_stage_ms = plain["stage_ms"] / max(plain["stage_launches"], 1)
    sa = plain.get("stage_standalone")
    if sa and sa[1] > 0:
        avg_stage_ms = sa[0] / sa[1]
        launch_note = ("k_stage alone: %d launches after the timed steps with the overlap off; in the timed steps "
                       "its tiles without interface faces share the CUs with the Riemann kernel "
                       "(both launches of a stage: %.3f ms)" % (sa[1], live_stage_ms))
    else:
        avg_stage_ms = live_stage_ms
        launch_note = "k_stage launches of the timed steps"
    achieved = avg_bytes / (avg_stage_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": plain["value"],
        "unit": "cell-stage-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.config, "description": desc, "cells": cells, "rk_stages": plain["stages"],
                   "riemann_guess": "acoustic (plain)"},
        "riemann": plain["riemann"],
        "reinit": plain.get("reinit"),
        "roofline": {"bound": "hbm", "kernel": "k_stage (fused residual + RK stage)",
                     "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": measured_traffic(args.config),
                     "bytes_per_cell_per_launch": per, "avg_launch_ms": avg_stage_ms, "launch_time": launch_note},
    }
    if nmat > 1 and not args.no_learned and os.path.exists(GUESS_PARAMS):
        theta = np.load(GUESS_PARAMS)
        cfg2, _ = build_config(args.config)
        lr = run_single(cfg2, steps, args.warmup, learned_params=theta)
        out["learned"] = {"value": lr["value"], "ms_per_step": lr["ms_per_step"], "riemann": lr["riemann"],
                          "speedup_vs_plain": lr["value"] / plain["value"], "params": "m2c_amd/data/guess_params.npy"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.config)
    print(json.dumps(out))


def bench_ensemble(args):
    """config 4: uneven shock-bubble ensemble, lockstep with graph replay."""
    from m2c_amd import runtime
    steps = args.steps if args.steps is not None else 100  # the configuration: 100 steps each
    cases = configs.c4_instances(count=args.instances, seed=4)
    print("# bench: ensemble of %d, cells %d, steps %d" % (len(cases), sum(c[0][0] ** 3 for c in cases), steps),
          file=sys.stderr, flush=True)
    learned = (not args.no_learned) and os.path.exists(GUESS_PARAMS)
    theta = np.load(GUESS_PARAMS) if learned else None
    members, cells = [], 0
    for (n, r, m), (g, mats, nmat, s, V, ID, Phi) in cases:
        if learned:
            s.riem_guess = abi.GUESS_MLP
        sol = runtime.Solver(g, mats, nmat, s)
        if learned:
            sol.mlp_set_params(theta)
        sol.set_state(V, ID, Phi)
        del V, ID, Phi
        members.append(sol)
        cells += g.nx * g.ny * g.nz
    ens = runtime.Ensemble(members)
    ens.run_steps(args.warmup)
    for sol in members:
        sol.reset_counters()
    t0 = time.perf_counter()
    ens.run_steps(steps)
    el = time.perf_counter() - t0
    calls = iters = fev = 0
    hit = hfe = None
    for sol in members:
        c = sol.counters()
        calls += c["riemann_calls"]
        iters += c["riemann_main_iters"]
        fev += c["riemann_fevals"]
        a, b = sol.riemann_histogram()
        hit = a if hit is None else hit + a
        hfe = b if hfe is None else hfe + b
    out = {"metric": METRIC, "value": cells * 3 * steps / el, "unit": "cell-stage-updates/s", "n_gpus": 1,
           "steps": steps, "warmup": args.warmup, "ms_per_step": el * 1e3 / steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "c4", "instances": len(cases), "cells": cells, "rk_stages": 3,
                      "riemann_guess": "learned (MLP)" if learned else "acoustic (plain)",
                      "description": "uneven ensemble of shock-bubble runs (edges 64..192, r 0.05-0.25 L, "
                                     "Mach 1.2-3.0), lockstep, one graph per step"},
           "riemann": {"faces": calls, "main_iters_per_face": iters / max(calls, 1),
                       "fevals_per_face": fev / max(calls, 1), "main_iters_p50": _quantile(hit, 0.5),
                       "main_iters_p99": _quantile(hit, 0.99), "fevals_p50": _quantile(hfe, 0.5),
                       "fevals_p99": _quantile(hfe, 0.99)}}
    ens.close()
    for sol in members:
        sol.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
