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
  c5          bubble cloud (default; the headline): 512x256x256, NASG water,
              200 air bubbles, Mach-2 shock, level set reinitialised every 10
              steps; built on the device
  c2          256^3 two-material shock-bubble, RK3
  c1          128x32x32 Sod tube, RK2
  c2_single   the c2 grid with water only (single-material residual)
  c4          uneven ensemble of shock-bubble runs alone (edges 64..192,
              --instances), lockstep, one captured graph per step

The default run also measures the c4 ensemble (learned guess, batched)
against the same members run one after another with the plain guess, under
the "ensemble_c4" key, and the CPU oracle on all host cores ("cpu_baseline",
measured first, before the GPU is touched).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]
Prints one JSON line.  Only one GPU is used; --gpus N > 1 exits.
"""
import argparse
import json
import os
import platform
import subprocess
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


def _brick_setup(name, index, count):
    """Brick `index` of a `count`-way split of the workload's grid, with its
    initial state built on the host as the device builds it (Outlet faces
    where the brick was cut: each brick is stepped on its own)."""
    if name == "c1":
        (g, mats, nmat, s, V, ID, Phi), _ = build_config("c1")
        n = (g.nx, g.ny, g.nz)
        p = configs.brick_split(n, count)
        b = (n[0] // p[0], n[1] // p[1], n[2] // p[2])
        bi = (index % p[0], (index // p[0]) % p[1], index // (p[0] * p[1]))
        lo = [bi[a] * b[a] for a in range(3)]
        h = [1.0 / n[0], 1.0 / n[1], 1.0 / n[2]]
        bg = abi.grid(b, tuple(lo[a] * h[a] for a in range(3)), tuple((lo[a] + b[a]) * h[a] for a in range(3)),
                      (abi.BC_OUTLET,) * 6)
        Vb = np.ascontiguousarray(V[lo[2]:lo[2] + b[2] + 2, lo[1]:lo[1] + b[1] + 2, lo[0]:lo[0] + b[0] + 2])
        IDb = np.ascontiguousarray(ID[lo[2]:lo[2] + b[2] + 2, lo[1]:lo[1] + b[1] + 2, lo[0]:lo[0] + b[0] + 2])
        configs.fill_ghosts(Vb, IDb, None, bg)
        return bg, mats, nmat, s, Vb, IDb, None
    if name == "c5":
        g, mats, nmat, s, st, sid, geoms = configs.c5_bubble_cloud()
    else:  # c2, c2_single, and the shock-bubble members of c4
        g, mats, nmat, s, st, sid, geoms = configs.c2_geometries(256)
        if name == "c2_single":
            geoms = geoms[:1]  # the shock plane: water everywhere
            s.n_levelsets = 0
            mats = abi.materials(mats[0])
            nmat = 1
    n = (g.nx, g.ny, g.nz)
    p = configs.brick_split(n, count)
    b = (n[0] // p[0], n[1] // p[1], n[2] // p[2])
    bi = (index % p[0], (index // p[0]) % p[1], index // (p[0] * p[1]))
    lo = tuple(bi[a] * b[a] for a in range(3))
    bg, V, ID, Phi = configs.initialize_host(g, s, st, sid, geoms, lo, b)
    return bg, mats, nmat, s, V, ID, Phi


def cpu_brick_worker(spec):
    """One process of the CPU baseline: the oracle (single thread) on one
    brick, stepped for about `budget` seconds; prints a JSON line."""
    name, index, count, budget = spec.split(":")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    g, mats, nmat, s, V, ID, Phi = _brick_setup(name, int(index), int(count))
    o = ol.Oracle(g, mats, nmat, s)
    o.set_state(V, ID, Phi)
    del V, ID, Phi
    stages = {abi.TI_FE: 1, abi.TI_RK2: 2, abi.TI_RK3: 3}[s.integrator]
    t0 = time.perf_counter()
    steps = 0
    while True:
        o.run_steps(1)
        steps += 1
        el = time.perf_counter() - t0
        if el > float(budget) or steps >= 100:
            break
    print(json.dumps({"cells": g.nx * g.ny * g.nz, "steps": steps, "seconds": el, "stages": stages,
                      "brick": [g.nx, g.ny, g.nz]}))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_cores():
    """Cores this process may use (its CPU affinity), capped by
    OMP_NUM_THREADS where the host sets it (the GPU box's share)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_baseline(name, workers, budget_s=20.0):
    """The CPU oracle (this project's single-threaded restatement of the
    reference's step) run the way the reference runs: one process per core on
    a brick split of the same grid, all at once; the value is the sum of the
    processes' cell-stage-update rates.  Bricks are stepped independently (no
    halo exchange, which the reference's MPI run would add).  Started before
    the GPU is touched."""
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-brick",
                               "%s:%d:%d:%g" % (name, b, workers, budget_s)],
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for b in range(workers)]
    res = []
    for p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("CPU baseline worker failed: " + err[-2000:])
        res.append(json.loads(out.strip().splitlines()[-1]))
    value = sum(r["cells"] * r["stages"] * r["steps"] / r["seconds"] for r in res)
    steps = sorted(r["steps"] for r in res)
    nsteps = str(steps[0]) if steps[0] == steps[-1] else "%d to %d" % (steps[0], steps[-1])
    return {"value": value, "unit": "cell-stage-updates/s", "cores": workers, "kind": "port",
            "sample": ("the same configuration split into %d bricks of %s cells, one single-threaded oracle "
                       "process per brick and core, all running at once, each for its first %s steps from the "
                       "initial state (about %g s); bricks do not exchange halos"
                       % (workers, "x".join(str(x) for x in res[0]["brick"]), nsteps, budget_s)),
            "cpu": _cpu_model(), "processes": workers}


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


def riemann_stats(cnt, hist=None, maxits=None):
    calls = max(cnt["riemann_calls"], 1)
    out = {"faces": cnt["riemann_calls"], "trivial": cnt["riemann_trivial"],
           "main_iters_per_face": cnt["riemann_main_iters"] / calls,
           "fevals_per_face": cnt["riemann_fevals"] / calls,
           "errors": cnt["riemann_errors"], "guess_fallbacks": cnt["riemann_guess_fallbacks"]}
    if hist is not None:
        it, fe = hist
        out.update({"main_iters_p50": _quantile(it, 0.5), "main_iters_p99": _quantile(it, 0.99),
                    "fevals_p50": _quantile(fe, 0.5), "fevals_p99": _quantile(fe, 0.99)})
        if maxits:
            # faces whose secant loop ran out of iterations (then the
            # non-adaptive pass runs as well): their count and their share of
            # all main-loop iterations (the last histogram bin holds counts
            # at or above its index, taken at that index)
            k = np.arange(len(it), dtype=np.float64)
            tot = float((k * it).sum())
            out.update({"maxits_faces": int(it[maxits:].sum()),
                        "maxits_face_fraction": float(it[maxits:].sum()) / calls,
                        "maxits_share_of_iterations": float((k[maxits:] * it[maxits:]).sum()) / tot if tot else 0.0})
    return out


FP64_PEAK_TFLOPS = 78.6


def _profile_figures(name):
    """`key: value` lines of a summary under profiles/ (numbers only)."""
    out = {}
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                key, sep, val = line.partition(":")
                if sep and key.strip() and " " not in key.strip():
                    try:
                        out[key.strip()] = float(val.split()[0])
                    except (ValueError, IndexError):
                        pass
    return out


def riemann_kernel_figures(cnt, riem_ms, steps):
    """The exact-Riemann kernel's own rates over the timed steps (its launch
    time from HIP events on its stream; with the overlap it shares the CUs
    with stage tiles).  The FP64 FLOP per f(p) evaluation that a counter
    profile of the kernel measured (profiles/k_riemann_c5_valu.txt: FP64
    instructions x 64 lanes x the active-lane fraction, per evaluation) turns
    the evaluation rate into FLOP/s."""
    sec = riem_ms * 1e-3
    out = {"ms_per_step": riem_ms / steps, "faces_per_s": cnt["riemann_calls"] / sec,
           "main_iters_per_s": cnt["riemann_main_iters"] / sec, "fevals_per_s": cnt["riemann_fevals"] / sec}
    prof = _profile_figures("k_riemann_c5_valu.txt")
    if "fp64_flop_per_feval" in prof:
        fpe = prof["fp64_flop_per_feval"]
        out["fp64_flop_per_feval"] = fpe
        out["fp64_tflops"] = fpe * out["fevals_per_s"] / 1e12
        out["fp64_peak_tflops"] = FP64_PEAK_TFLOPS
        if "active_lane_fraction" in prof:
            out["active_lane_fraction"] = prof["active_lane_fraction"]
    return out


def riemann_roofline(rk):
    """The exact-Riemann kernel against the FP64 vector peak: it is bound by
    the serial latency of its slowest faces, not by a throughput, so the
    fraction is far below one by construction and says how much of the
    chip's FP64 issue the latency-bound solve uses."""
    if not rk or "fp64_tflops" not in rk:
        return None
    return {"kernel": "k_riemann (exact two-material Riemann solver, one lane per interface face)",
            "bound": "fp64-latency", "achieved": rk["fp64_tflops"], "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": rk["fp64_tflops"] / FP64_PEAK_TFLOPS, "flop_per_feval": rk["fp64_flop_per_feval"],
            "active_lane_fraction": rk.get("active_lane_fraction")}


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
    # the other kernels' times of the timed steps, before any further step
    riem_ms, _ = solver.kernel_time("riemann")
    mlp_ms, _ = solver.kernel_time("mlp")
    reinit_ms, _ = solver.kernel_time("reinit")
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
    res = {"value": cells * stages * steps / el, "ms_per_step": el * 1e3 / steps, "cells": cells,
           "stages": stages, "stage_ms": stage_ms, "stage_launches": stage_n, "stage_standalone": standalone,
           "riemann": dict(riemann_stats(cnt, hist, s.riem_maxits_main), riemann_ms_per_step=riem_ms / steps,
                           mlp_ms_per_step=mlp_ms / steps)}
    if nmat > 1 and riem_ms > 0:
        res["riemann_kernel"] = riemann_kernel_figures(cnt, riem_ms, steps)
    if cnt["reinit_calls"]:
        res["reinit"] = {"calls": cnt["reinit_calls"], "iterations": cnt["reinit_iterations"],
                         "ms_per_call": reinit_ms / cnt["reinit_calls"], "ms_per_step": reinit_ms / steps}
    solver.close()
    return res


def ensemble_case(count, learned, seed=4):
    """The c4 members (device-built shock-bubbles of uneven size) as solvers."""
    from m2c_amd import runtime
    theta = np.load(GUESS_PARAMS) if learned else None
    members, cells = [], 0
    for (n, r, m), (g, mats, nmat, s, st, sid, geoms) in configs.c4_geometries(count=count, seed=seed):
        if learned:
            s.riem_guess = abi.GUESS_MLP
        sol = runtime.Solver(g, mats, nmat, s)
        if learned:
            sol.mlp_set_params(theta)
        sol.initialize(st, sid, geoms)
        members.append(sol)
        cells += g.nx * g.ny * g.nz
    return members, cells


def run_ensemble(count, steps, warmup, learned):
    """All members in lockstep: one captured graph per step."""
    from m2c_amd import runtime
    members, cells = ensemble_case(count, learned)
    ens = runtime.Ensemble(members)
    ens.run_steps(warmup)
    for sol in members:
        sol.synchronize()
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
    ens.close()
    for sol in members:
        sol.close()
    return {"value": cells * 3 * steps / el, "ms_per_step": el * 1e3 / steps, "cells": cells, "instances": count,
            "riemann": {"faces": calls, "main_iters_per_face": iters / max(calls, 1),
                        "fevals_per_face": fev / max(calls, 1), "main_iters_p50": _quantile(hit, 0.5),
                        "main_iters_p99": _quantile(hit, 0.99), "fevals_p50": _quantile(hfe, 0.5),
                        "fevals_p99": _quantile(hfe, 0.99)}}


def run_solo(count, steps, warmup):
    """The same members one after another, each alone with the plain guess
    (m2cx_run_steps); the stepping-loop times are summed."""
    members, cells = ensemble_case(count, False)
    total = 0.0
    for sol in members:
        sol.run_steps(warmup)
        sol.synchronize()
        t0 = time.perf_counter()
        sol.run_steps(steps)
        sol.synchronize()
        total += time.perf_counter() - t0
        sol.close()
    return {"value": cells * 3 * steps / total, "ms_per_step": total * 1e3 / steps, "cells": cells}


def c4_summary(count, steps, warmup):
    """The batched + learned ensemble against the plain method on the same
    members (run alone, one after another)."""
    learned = os.path.exists(GUESS_PARAMS)
    ens = run_ensemble(count, steps, warmup, learned)
    solo = run_solo(count, steps, warmup)
    return {"workload": "c4", "instances": count, "cells": ens["cells"], "steps": steps, "warmup": warmup,
            "riemann_guess": "learned (MLP)" if learned else "acoustic (plain)",
            "value": ens["value"], "ms_per_step": ens["ms_per_step"], "riemann": ens["riemann"],
            "plain_alone_value": solo["value"], "plain_alone_ms_per_step": solo["ms_per_step"],
            "speedup_vs_plain_alone": ens["value"] / solo["value"],
            "description": "uneven ensemble of shock-bubble runs (edges 64..192, r 0.05-0.25 L, Mach 1.2-3.0), "
                           "lockstep, one graph per step, against the plain guess with each member run alone"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default=os.environ.get("M2C_BENCH_CONFIG", "c5"))
    ap.add_argument("--instances", type=int, default=64, help="ensemble size for c4")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the c4 ensemble of the default run")
    ap.add_argument("--cpu-workers", type=int, default=0, help="CPU baseline processes (default: host cores)")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds each CPU baseline process steps")
    ap.add_argument("--cpu-brick", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.cpu_brick:
        cpu_brick_worker(args.cpu_brick)
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build does not run the step across GPUs "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    # the CPU baseline first, before anything touches the GPU
    cpu = None
    if not args.no_cpu_baseline:
        workers = args.cpu_workers or host_cores()
        print("# bench: CPU baseline, %d processes" % workers, file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.config, workers, args.cpu_budget)
    if args.config == "c4":
        steps = args.steps if args.steps is not None else 100  # the configuration: 100 steps each
        summ = c4_summary(args.instances, steps, args.warmup)
        out = {"metric": METRIC, "value": summ["value"], "unit": "cell-stage-updates/s", "n_gpus": 1,
               "steps": steps, "warmup": args.warmup, "ms_per_step": summ["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
               "config": {k: summ[k] for k in ("workload", "instances", "cells", "riemann_guess", "description")},
               "riemann": summ["riemann"], "ensemble_c4": summ}
        if cpu:
            out["cpu_baseline"] = cpu
        print(json.dumps(out))
        return
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
        "riemann_kernel": plain.get("riemann_kernel"),
        "roofline_riemann": riemann_roofline(plain.get("riemann_kernel")),
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
                          "riemann_kernel": lr.get("riemann_kernel"),
                          "speedup_vs_plain": lr["value"] / plain["value"], "params": "m2c_amd/data/guess_params.npy"}
    if args.config == "c5" and not args.no_ensemble:
        print("# bench: c4 ensemble of %d" % args.instances, file=sys.stderr, flush=True)
        # its first 10 steps: the members run alone with the plain guess take
        # about 3.5 s a step together
        out["ensemble_c4"] = c4_summary(args.instances, min(steps, 10), args.warmup)
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
