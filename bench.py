#!/usr/bin/env python3
"""Benchmark of the HyPE-2D finite-volume RHS on one MI355X.

A "step" is one full SSP-RK time step (all stages) of the workload, with the
state resident in HBM.  Default workload: the Kapila 5-equation air-helium
shock-bubble on 1024x512 cells, FP64, 500 steps (WENO3 + Rusanov + SSPRK(4,3),
CFL 0.5).  Prints one JSON line with the throughput in Mcell-updates/s, the
HBM roofline of the stage kernel and a CPU baseline (the FP64 C oracle on the
host cores, bounded sample).

    python bench.py [--steps K] [--warmup W] [--config kapila|euler|bn]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hype-parallel_amd"))

import numpy as np  # noqa: E402

import hype2d as H  # noqa: E402

METRIC = "Mcell-updates/s + % HBM peak; served p50/p99 latency vs one-at-a-time"
HBM_PEAK_GBS = 8000.0


def workload(name):
    """(problem, initial state, steps, description, recons path)."""
    if name == "kapila":
        nx, ny = 1024, 512
        h = 2.0 / nx
        p = H.make_problem(H.KAPILA, nx, ny, order=2, prim=True, domain=(0.0, 2.0, 0.0, 1.0), cfl=0.5,
                           bc=(H.BC_TRANSMISSIVE, H.BC_TRANSMISSIVE, H.BC_REFLECTIVE, H.BC_REFLECTIVE),
                           gamma1=1.4, gamma2=1.648, pi1=0.0, pi2=0.0)
        params = [1.22, 0.5, 1.0, 0.5, 0.25, 1e-6, 2.0 * h, 1.4, 0.25463, 1.0e5]
        U0 = H.initial_state(p, H.IC_SHOCK_BUBBLE, params)
        return p, U0, 500, "kapila_shock_bubble_1024x512"
    if name == "euler":
        nx = ny = 256
        p = H.make_problem(H.EULER, nx, ny, order=2, prim=True, domain=(0.0, 1.0, 0.0, 1.0), cfl=0.5,
                           bc=(H.BC_TRANSMISSIVE,) * 4, gamma1=1.4)
        params = [0.5, 0.5, 1.5, 0.0, 0.0, 1.5, 0.5323, 1.206, 0.0, 0.3, 0.138, 1.206, 1.206, 0.029,
                  0.5323, 0.0, 1.206, 0.3]
        U0 = H.initial_state(p, H.IC_QUADRANTS, params)
        return p, U0, 200, "euler_quadrants_256x256"
    if name == "bn":
        nx, ny = 2048, 1024
        h = 2.0 / nx
        p = H.make_problem(H.BAER_NUNZIATO, nx, ny, order=2, prim=False, domain=(0.0, 2.0, 0.0, 1.0),
                           cfl=0.5, bc=(H.BC_TRANSMISSIVE,) * 4, gamma1=4.4, gamma2=1.4, pi1=6.0e8, pi2=0.0)
        params = [0.5, 1e-6, 2.0 * h, 1.0, 0.3, 0.1, 1.0e9, 1.0e5, 1000.0, 1.0, 8, 0.005, 2.0]
        U0 = H.initial_state(p, H.IC_WATER_AIR, params, seed=2024)
        return p, U0, 300, "baer_nunziato_water_air_2048x1024"
    raise SystemExit(f"unknown config {name}")


def bytes_per_cell_step(nv, order):
    # fused stage kernels: read the stage input, write the stage output,
    # plus one extra read of U^n in the combination stage of SSPRK(4,3)
    if order == 2:
        return 9 * nv * 8
    return 5 * nv * 8


def cpu_baseline(p, U0, nv, budget_s=12.0):
    """The FP64 C oracle (OpenMP build) on the same workload, bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as orc

    op = orc.problem(p.system, p.nx, p.ny, N=p.order, prim=bool(p.recons_prim),
                     domain=(p.x_min, p.x_max, p.y_min, p.y_max), cfl=p.cfl, bc=tuple(p.bc),
                     g1=p.gamma1, g2=p.gamma2, p1=p.pi1, p2=p.pi2, quirks=p.quirks)
    t0 = time.perf_counter()
    orc.solve(op, U0, 1, threaded=True)
    t1 = time.perf_counter() - t0
    k = int(max(1, min(60, round(budget_s / max(t1, 1e-3)))))
    t0 = time.perf_counter()
    orc.solve(op, U0, k + 1, threaded=True)
    tk = time.perf_counter() - t0
    rate = p.nx * p.ny * k / max(tk - t1, 1e-9) / 1e6
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(rate, 4), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{k} SSP-RK steps of the same {p.nx}x{p.ny} workload from its initial state "
                      f"(C oracle, OpenMP, {threads} threads; priming and allocation cancelled by "
                      f"differencing a 1-step and a {k + 1}-step run)"}


def calibrate_copy(mib):
    """A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE counters
    per dispatch calibrate the stage kernel's traffic figures."""
    import torch

    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_mib": mib, "bytes_per_copy": 2 * n * 8, "ms": ms,
                      "GBps": 2 * n * 8 / ms / 1e6}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="kapila")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--hrows", type=int, default=0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design only "
                                   "(see DESIGN.md), so --gpus must be 1"}))
        sys.exit(2)

    p, U0, steps_default, wname = workload(args.config)
    steps = args.steps if args.steps is not None else steps_default
    nv = H.NVAR[p.system]
    cells = p.nx * p.ny
    flags = H.TS_CLIP | (H.TS_GRAPH if args.graph else 0)
    with H.Context(p) as ctx:
        if args.hrows:
            ctx.set_tiling(args.hrows, 0)
        # warmup (untimed) from the initial state, then restart from it
        ctx.set_state(U0)
        ctx.begin(0.0, float("inf"), max(args.warmup, 1), flags)
        ctx.advance(max(args.warmup, 1))
        ctx.sync()
        ctx.set_state(U0)
        ctx.begin(0.0, float("inf"), steps, flags)
        ctx.sync()
        ms = ctx.advance_timed(steps)
        st = ctx.sync()
        hrows, ntiles, grid = ctx.tiling()
    if st.steps != steps:
        raise SystemExit(f"only {st.steps} of {steps} steps ran")
    sec = ms / 1e3
    mcups = cells * steps / sec / 1e6
    nstage = 4 if p.order == 2 else (2 if p.order == 1 else 1)
    launches = steps * nstage
    avg_launch_s = sec / launches
    bytes_launch = cells * bytes_per_cell_step(nv, p.order) / nstage
    achieved = bytes_launch / avg_launch_s / 1e9
    out = {
        "metric": METRIC,
        "value": round(mcups, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wname, "cells": cells, "nvar": nv, "order": p.order,
                   "path": "RHSFunctionPrim" if p.recons_prim else "RHSFunction",
                   "scheme": "WENO3+Rusanov+SSPRK(4,3)" if p.order == 2 else f"order{p.order}",
                   "cfl": p.cfl, "tile_rows": hrows, "tiles": ntiles, "graph": bool(args.graph)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "stage_kernel", "avg_launch_us": round(avg_launch_s * 1e6, 3),
                     "algorithmic_bytes_per_launch": int(bytes_launch)},
        "final_time": st.t,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, U0, nv)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
