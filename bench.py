#!/usr/bin/env python3
"""Benchmark of the HyPE-2D finite-volume RHS on one MI355X.

A "step" is one full SSP-RK time step (all stages) of the workload, with the
states resident in HBM.  The default run (``--config all``) times the largest
configuration that fits one GPU, the batched set: 64 independent Kapila
air-helium shock-bubbles with sides log-uniform on [128, 1024], FP64, WENO3 +
Rusanov + SSPRK(4,3), 400 steps each, advanced together by one persistent
launch per RK stage (hype_batch).  Its JSON line carries the throughput in
Mcell-updates/s, the HBM roofline of the stage kernel, the same instances
solved one at a time, and a CPU baseline (the FP64 C oracle with MPI ranks
on the host cores, bounded sample).  The three fixed grids are timed in the
same run and reported under "configs", each with its own roofline and CPU
baseline: Kapila 1024x512 (500 steps), Euler 256^2 (200), Baer-Nunziato
2048x1024 (300); so is the served stream's first 64 requests (latency and
throughput against one request at a time).

    python bench.py [--steps K] [--warmup W]            # batched + the three grids
    python bench.py --config kapila|euler|bn|batched     # one configuration
    python bench.py --config served [--requests N]       # Poisson request stream
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hype-parallel_amd"))

import numpy as np  # noqa: E402

import hype2d as H  # noqa: E402
import workloads as WL  # noqa: E402

METRIC = "Mcell-updates/s + % HBM peak; served p50/p99 latency vs one-at-a-time"
HBM_PEAK_GBS = 8000.0


def bytes_per_cell_step(nv, order):
    # fused stage kernels: read the stage input, write the stage output,
    # plus one extra read of U^n in the combination stage of SSPRK(4,3)
    if order == 2:
        return 9 * nv * 8
    return 5 * nv * 8


def mpi_exec():
    """mpiexec of the MPI installation the oracle's MPI program was built with."""
    for cand in (shutil.which("mpiexec"), "/opt/conda/bin/mpiexec"):
        if cand and os.path.exists(cand):
            return cand
    return None


def oracle_mpi(p, U0, ranks, budget_s):
    """The FP64 C oracle with MPI ranks (oracle/orc_mpi_baseline.c: row
    blocks, width-3 halo exchange per RHS call, Allreduce(MIN) of dt), as the
    reference runs its PETSc path.  Returns (ranks, steps, seconds) of the
    timed steps, or None when it cannot run here."""
    exe = os.path.join(ROOT, "oracle", "lib", "orc_mpi_baseline")
    mpiexec = mpi_exec()
    if not os.path.exists(exe) or mpiexec is None or ranks < 2 or p.ny < 3 * ranks:
        return None
    if H.BC_PERIODIC in (p.bc[2], p.bc[3]):
        return None
    h = (p.x_max - p.x_min) / p.nx
    fd, path = tempfile.mkstemp(suffix=".bin")
    try:
        with os.fdopen(fd, "wb") as f:
            np.ascontiguousarray(U0, dtype=np.float64).tofile(f)
        args = [mpiexec, "-n", str(ranks), exe, path, str(p.system), str(p.order), str(p.recons_prim),
                str(p.nx), str(p.ny), repr(p.x_min), repr(p.y_min), repr(h), repr(p.cfl)]
        args += [str(p.bc[k]) for k in range(4)]
        args += [repr(p.gamma1), repr(p.gamma2), repr(p.pi1), repr(p.pi2), str(p.quirks), repr(budget_s)]
        r = subprocess.run(args, capture_output=True, text=True, timeout=240)
        words = r.stdout.split()
        if r.returncode != 0 or len(words) < 6 or words[0] != "ranks":
            return None
        return int(words[1]), int(words[3]), float(words[5])
    except (OSError, subprocess.SubprocessError, ValueError):
        return None
    finally:
        os.unlink(path)


MPI_NOTE = "C oracle, {r} MPI ranks: row blocks with a width-3 halo exchange per RHS call and an Allreduce(MIN) of dt"


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def oracle_openmp(p, U0, budget_s):
    """The OpenMP build of the oracle (where no MPI installation is found):
    (threads, steps, seconds), priming and allocation cancelled by
    differencing a 1-step and a (k+1)-step run."""
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
    return host_threads(), k, max(tk - t1, 1e-9)


def cpu_baseline(p, U0, nv, budget_s=12.0):
    """The FP64 C oracle on the same workload, bounded sample: with MPI ranks
    on the host cores where MPI is installed, else the OpenMP build."""
    threads = host_threads()
    res = oracle_mpi(p, U0, threads, budget_s)
    if res is not None:
        nr, k, sec = res
        return {"value": round(p.nx * p.ny * k / sec / 1e6, 4), "unit": "Mcell-updates/s", "cores": nr,
                "kind": "port",
                "sample": f"{k} SSP-RK steps of the same {p.nx}x{p.ny} workload after its priming call and "
                          f"first step ({MPI_NOTE.format(r=nr)})"}
    nt, k, sec = oracle_openmp(p, U0, budget_s)
    return {"value": round(p.nx * p.ny * k / sec / 1e6, 4), "unit": "Mcell-updates/s", "cores": nt,
            "kind": "port",
            "sample": f"{k} SSP-RK steps of the same {p.nx}x{p.ny} workload from its initial state "
                      f"(C oracle, OpenMP, {nt} threads; priming and allocation cancelled by "
                      f"differencing a 1-step and a {k + 1}-step run)"}


def cpu_baseline_batched(insts, budget_s=4.0, nsample=4):
    """The oracle on a sample of the batched set: the instances at evenly
    spaced ranks of the size order (smallest and largest included), each
    solved alone on all host cores for about budget_s seconds, as the
    reference would run them one program after another.  Rate = sampled
    cell-updates over the sampled time."""
    order = sorted(range(len(insts)), key=lambda i: insts[i][0].nx * insts[i][0].ny)
    pick = sorted({order[round(k * (len(order) - 1) / (nsample - 1))] for k in range(nsample)})
    threads = host_threads()
    cell_steps, sec, cores, how, sides = 0, 0.0, threads, None, []
    for i in pick:
        p, U0 = insts[i]
        res = oracle_mpi(p, U0, threads, budget_s)
        if res is not None:
            cores, k, s = res
            how = MPI_NOTE.format(r=cores)
        else:
            cores, k, s = oracle_openmp(p, U0, budget_s)
            how = f"C oracle, OpenMP, {cores} threads"
        cell_steps += p.nx * p.ny * k
        sec += s
        sides.append(f"#{i} {p.nx}^2 x {k} steps")
    return {"value": round(cell_steps / sec / 1e6, 4), "unit": "Mcell-updates/s", "cores": cores, "kind": "port",
            "sample": f"{len(pick)} of the {len(insts)} instances, each alone for about {budget_s:g} s of "
                      f"SSP-RK steps after its priming call and first step ({', '.join(sides)}; {how})"}


def profile_figure(workload_name, key):
    """A figure measured by rocprofv3 for this workload's stage kernel, read
    from the committed profile summary (a `key: value` line), or None."""
    path = os.path.join(ROOT, "profiles", f"stage_kernel_{workload_name}.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def load_traffic(workload_name):
    """HBM bytes per stage-kernel launch measured by rocprofv3 (FETCH_SIZE and
    WRITE_SIZE, each scaled by the same counter of a plain copy of known
    size)."""
    return profile_figure(workload_name, "traffic_bytes_per_launch")


CU_COUNT, CLOCK_GHZ = 256, 2.4  # MI355X


def valu_roofline(workload_name, launch_s):
    """FP64 VALU issue, the roof that binds the stage kernel: wave-instruction
    issue slots per launch (SQ_ACTIVE_INST_VALU, transcendentals counted at
    their quarter rate) over the launch time, against one wave instruction
    per CU per cycle."""
    slots = profile_figure(workload_name, "valu_issue_slots_per_launch")
    if slots is None:
        return None
    achieved = slots / launch_s / 1e9
    peak = CU_COUNT * CLOCK_GHZ
    return {"bound": "fp64-valu-issue", "achieved": round(achieved, 2), "peak": peak,
            "unit": "G wave-instructions/s", "frac": round(achieved / peak, 4), "issue_slots_per_launch": slots}


def aggregate_roofline(cell_steps_by_variant, seconds, workload_name=None):
    """Algorithmic HBM bytes of all stage launches over the timed region;
    traffic (HBM bytes per launch, rocprofv3) where the workload has a
    committed profile."""
    total = sum(n * bytes_per_cell_step(nv, order) for (nv, order), n in cell_steps_by_variant.items())
    achieved = total / seconds / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": load_traffic(workload_name) if workload_name else None,
            "kernel": "stage_kernel (all variants)", "algorithmic_bytes": int(total)}


def per_launch(roofline, launches):
    """Add the algorithmic bytes per stage launch (the unit of `traffic`)."""
    roofline["algorithmic_bytes_per_launch"] = int(roofline["algorithmic_bytes"] / launches)
    return roofline


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_batched(args, with_cpu=True):
    """The batched set advanced together (hype_batch), against the same
    instances solved one after another (hype_ctx, device-resident, graphs).
    Returns the JSON record."""
    count = args.instances
    steps = args.steps if args.steps is not None else WL.BATCHED_STEPS
    insts = WL.batched(args.seed, count)
    wname = WL.batched_name(count)
    cells = sum(p.nx * p.ny for p, _ in insts)
    progress(f"batched: {count} instances, {cells} cells, {steps} steps")
    with H.Batch([p for p, _ in insts]) as b:
        # tile shape and layout measured on the live states around the
        # tile-shape model's height (untimed; the states are set again below)
        model_rows = b.tiling()[0]
        cands = sorted({max(1, model_rows * f // 4) for f in (2, 3, 4, 5, 6)})
        for i, (_, U0) in enumerate(insts):
            b.set_state(i, U0)
        tuned_rows, tuned_ms = b.autotune(cands, steps=4)
        tuned = {"candidates": cands, "model_rows": model_rows, "rows": tuned_rows, "layout": b.layout(),
                 "ms_per_step": round(tuned_ms / 4, 4)}
        progress(f"batched: autotune {tuned}")
        for i, (_, U0) in enumerate(insts):
            b.set_state(i, U0)
        b.begin(max_steps=min(args.warmup, steps) or 1)
        b.advance(max(args.warmup, 1))
        b.sync()
        for i, (_, U0) in enumerate(insts):
            b.set_state(i, U0)
        b.begin(max_steps=steps)
        b.sync()
        ms = b.advance_timed(steps)
        st = b.sync()
        rows, ntiles, groups = b.tiling()
        check = [b.get_state(i) for i in range(min(count, args.check))]
    if any(s.steps != steps for s in st):
        raise SystemExit("batched: not every instance ran all steps")
    progress(f"batched: {ms:.1f} ms; now one at a time")
    solo_ms = 0.0
    identical = True
    for i, (p, U0) in enumerate([] if args.no_solo else insts):
        with H.Context(p) as ctx:
            ctx.set_state(U0)
            ctx.begin(0.0, float("inf"), steps, H.TS_CLIP | H.TS_GRAPH)
            ctx.sync()
            solo_ms += ctx.advance_timed(steps)
            ctx.sync()
            if i < len(check):
                identical &= bool(np.array_equal(ctx.get_state(), check[i]))
    sec = ms / 1e3
    value = cells * steps / sec / 1e6
    roof = per_launch(aggregate_roofline({(6, 2): cells * steps}, sec, wname), 4 * steps)
    roof["kernel"] = "stage_kernel (Kapila, prim, N=2, split tiles; persistent queue, tallest tiles first)"
    roof["avg_launch_us"] = round(sec / (4 * steps) * 1e6, 3)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(ms / steps, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": wname, "instances": count, "cells": cells, "nvar": 6, "order": 2,
                   "path": "RHSFunctionPrim", "scheme": "WENO3+Rusanov+SSPRK(4,3)", "cfl": 0.5,
                   "seed": args.seed, "tile_rows": rows, "tiles": ntiles, "variants": groups,
                   "autotune": tuned, "graph": True, "kernel_build": "exact"},
        "roofline": roof,
        "valu_roofline": valu_roofline(wname, sec / (4 * steps)),
    }
    if not args.no_solo:
        out["one_at_a_time"] = {"value": round(cells * steps / (solo_ms / 1e3) / 1e6, 3),
                                "unit": "Mcell-updates/s", "ms": round(solo_ms, 3)}
        out["speedup_vs_one_at_a_time"] = round(solo_ms / ms, 3)
        out["identical_to_solo"] = identical
        out["checked_instances"] = len(check)
    if with_cpu and not args.no_cpu:
        progress("batched: CPU baseline")
        out["cpu_baseline"] = cpu_baseline_batched(insts)
    return out


class RequestBuilder:
    """Initial states of the stream's requests, built ahead of their arrival
    on host threads (the IC quadrature runs in C and releases the GIL), at
    most `ahead` requests in advance so host memory stays bounded."""

    def __init__(self, draws, ahead=48, workers=8):
        from concurrent.futures import ThreadPoolExecutor

        self.draws = draws
        self.ahead = ahead
        self.pool = ThreadPoolExecutor(max_workers=workers)
        self.fut = {}
        self.next = 0

    def _fill(self, k):
        while self.next < len(self.draws) and self.next < k + self.ahead:
            _, system, n, steps, seed = self.draws[self.next]
            self.fut[self.next] = self.pool.submit(WL.served_request, system, n, steps, seed)
            self.next += 1

    def ready(self, k):
        """Whether request k has been built (schedules the next ones)."""
        self._fill(k)
        return self.fut[k].done()

    def get(self, k):
        """(problem, U0, t_end, cap) of request k (blocks until built)."""
        self._fill(k)
        return self.fut.pop(k).result()

    def close(self):
        self.pool.shutdown(wait=True)


def digest(U):
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(U, dtype=np.float64).tobytes()).hexdigest()


def serve_trace(draws, policy=H.SERVE_FIFO, active_cells=0):
    """Replay a stream through one hype_server in real time.  Returns per
    request (problem, t_end, cap, info, result digest) in submission order,
    and the makespan; results are hashed and released as they complete."""
    # the states of a stream of up to 1024 requests are built before the
    # clock starts; a longer stream builds ahead of its arrivals, and the
    # loop never waits for a build (it keeps polling, so the device never
    # runs dry while the host computes initial states)
    prebuild = len(draws) <= 1024
    builder = RequestBuilder(draws, ahead=len(draws) if prebuild else 256,
                             workers=min(16, os.cpu_count() or 1))
    if prebuild:
        builder.ready(0)
        for f in builder.fut.values():
            f.result()
    out = [None] * len(draws)
    with H.Server(policy=policy, max_active_cells=active_cells) as srv:
        # warm-up: a few tiny requests (kernel loading, graph capture)
        for k in range(min(3, len(draws))):
            _, system, n, steps, seed = draws[k]
            p, U0, _, _ = WL.served_request(system, min(n, 64), 2, seed)
            srv.submit(p, U0, max_steps=2)
        srv.drain()
        base = srv.now()
        pending = {}
        k = 0
        last = 0.0
        while k < len(draws) or pending:
            now = srv.now() - base
            if now - last > 10.0:
                last = now
                progress(f"served: t={now:.1f}s submitted={k} running/queued={len(pending)}")
            while k < len(draws) and draws[k][0] <= now and builder.ready(k):
                p, U0, t_end, cap = builder.get(k)
                pending[srv.submit(p, U0, t_end=t_end, max_steps=cap)] = (k, p, t_end, cap)
                k += 1
            a, q = srv.poll()
            for rid in [r for r in pending if srv.info(r).state == H.REQ_DONE]:
                j, p, t_end, cap = pending.pop(rid)
                U, inf = srv.result(rid)
                out[j] = (p, t_end, cap, inf, digest(U))
            if (a, q) == (0, 0) and k < len(draws):
                time.sleep(min(max(draws[k][0] - (srv.now() - base), 0.0), 0.01))
        makespan = srv.now() - base
    builder.close()
    return out, base, makespan


POLICIES = {"fifo": H.SERVE_FIFO, "shortest": H.SERVE_SHORTEST}


def run_served(args):
    """Requests arrive as a Poisson stream and are served with continuous
    batching (hype_server); time-to-solution p50/p99 and aggregate throughput,
    against the same requests served one at a time in arrival order."""
    rng = np.random.default_rng(args.seed)
    draws = WL.served_draws(rng, args.requests, args.rate)
    progress(f"served: stream of {len(draws)} requests drawn")
    res, base, makespan = serve_trace(draws, POLICIES[args.policy], args.active_cells)
    tts = np.array([inf.t_done - base - d[0] for (_, _, _, inf, _), d in zip(res, draws)])
    cell_steps = sum(inf.status.steps * p.nx * p.ny for p, _, _, inf, _ in res)
    variants = {}
    for p, _, _, inf, _ in res:
        key = (H.NVAR[p.system], 2)
        variants[key] = variants.get(key, 0) + inf.status.steps * p.nx * p.ny
    out = {
        "metric": METRIC, "value": round(cell_steps / makespan / 1e6, 3), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": int(sum(inf.status.steps for _, _, _, inf, _ in res)), "warmup": 3, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": WL.served_name(args.rate, args.requests),
                   "requests": args.requests, "rate_per_s": args.rate, "seed": args.seed,
                   "policy": args.policy, "max_active_cells": args.active_cells,
                   "cell_steps": int(cell_steps)},
        "roofline": aggregate_roofline(variants, makespan),
        "latency_s": {"p50": round(float(np.percentile(tts, 50)), 4), "p99": round(float(np.percentile(tts, 99)), 4),
                      "makespan": round(makespan, 3)},
        "inadmissible_requests": int(sum(1 for _, _, _, inf, _ in res if inf.status.err)),
        "inadmissible_ids": [j for j, (_, _, _, inf, _) in enumerate(res) if inf.status.err],
    }
    if args.no_solo:
        return out
    # one at a time, FIFO in arrival order: each request alone on its own
    # context (upload, solve, download), queueing simulated from its duration
    progress(f"served: makespan {makespan:.2f}s; now one at a time")
    builder = RequestBuilder(draws)
    identical = True
    finish, solo_tts, busy = 0.0, [], 0.0
    for j, ((_, _, _, inf, dig), d) in enumerate(zip(res, draws)):
        if j % 16 == 0:
            progress(f"served: one at a time {j}/{len(draws)}")
        p, U0, t_end, cap = builder.get(j)
        t0 = time.perf_counter()
        with H.Context(p) as ctx:
            ctx.set_state(U0)
            ctx.begin(0.0, t_end, cap, H.TS_CLIP | H.TS_GRAPH)
            ctx.advance(cap // 2)  # the requested step count; more if t_end is not reached
            # a request may end on an inadmissible state (the reference would
            # abort there); the server reports it, and so must the solo run
            st = ctx.sync(allow_inadmissible=True)
            while not st.done:
                ctx.advance(32)
                st = ctx.sync(allow_inadmissible=True)
            Us = ctx.get_state()
        dur = time.perf_counter() - t0
        busy += dur
        identical &= digest(Us) == dig and st.steps == inf.status.steps and st.err == inf.status.err
        finish = max(finish, d[0]) + dur
        solo_tts.append(finish - d[0])
    builder.close()
    solo_tts = np.array(solo_tts)
    out["one_at_a_time"] = {"value": round(cell_steps / busy / 1e6, 3), "unit": "Mcell-updates/s",
                            "p50": round(float(np.percentile(solo_tts, 50)), 4),
                            "p99": round(float(np.percentile(solo_tts, 99)), 4),
                            "makespan": round(float(finish), 3)}
    out["speedup_vs_one_at_a_time"] = round(busy / makespan, 3)
    out["identical_to_solo"] = identical
    return out


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


MOMENTA = {H.EULER: [(1, 2)], H.KAPILA: [(2, 3)], H.BAER_NUNZIATO: [(1, 2), (5, 6)]}


def rel_dev(a, b, system):
    """max over components of ||a_c - b_c||_inf / ||b_c||_inf, each momentum
    pair scaled by the maximum of its vector magnitude (a transverse momentum
    that is physically still zero holds only rounding noise)"""
    scale = [max(float(np.max(np.abs(b[..., c]))), 1e-300) for c in range(b.shape[-1])]
    for (cu, cv) in MOMENTA[system]:
        scale[cu] = scale[cv] = max(float(np.max(np.hypot(b[..., cu], b[..., cv]))), 1e-300)
    return max(float(np.max(np.abs(a[..., c] - b[..., c]))) / scale[c] for c in range(b.shape[-1]))


def timed_run(p, U0, steps, args, pcie=True):
    """Warm up, then time `steps` device steps from U0 with HIP events on the
    context's stream.  Returns (ms, status, final state, tiling, seconds per
    host-buffer RHS call)."""
    flags = H.TS_CLIP | (H.TS_GRAPH if args.graph else 0)
    rhs_s = None
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
        U_end = ctx.get_state()
        tiling = ctx.tiling()
        if pcie:
            # the drop-in callback with host buffers (TSRHSFunction on host
            # Vecs): every call uploads U and downloads the RHS over PCIe
            ctx.rhs(U0)
            nrhs = 10
            t0 = time.perf_counter()
            for _ in range(nrhs):
                ctx.rhs(U0)
            rhs_s = (time.perf_counter() - t0) / nrhs
    return ms, st, U_end, tiling, rhs_s


def run_single(name, args, steps=None, fma=True, pcie=True, with_cpu=True):
    """One fixed grid: device-resident SSP-RK steps timed with HIP events on
    the context's stream.  Returns the JSON record."""
    p, U0, steps_default, wname = WL.single(name)
    steps = steps if steps is not None else steps_default
    nv = H.NVAR[p.system]
    cells = p.nx * p.ny
    progress(f"{wname}: {steps} steps")
    ms, st, U_end, (hrows, ntiles, grid), rhs_s = timed_run(p, U0, steps, args, pcie=pcie and not args.no_pcie)
    if st.steps != steps:
        raise SystemExit(f"{wname}: only {st.steps} of {steps} steps ran")
    fma_variant = None
    if fma and not args.no_fma:
        # the opt-in FMA-contracted build (HYPE_OPT_FMA) on the same run,
        # reported beside the exact build, which is the headline
        p_fma = H.Problem.from_buffer_copy(p)
        p_fma.options = H.OPT_FMA
        ms_f, st_f, U_f, _, _ = timed_run(p_fma, U0, steps, args, pcie=False)
        fma_variant = {"value": round(cells * steps / (ms_f / 1e3) / 1e6, 3),
                       "max_rel_dev_vs_exact": rel_dev(U_f, U_end, p.system),
                       "final_time_rel_dev": abs(st_f.t - st.t) / st.t,
                       "note": "HYPE_OPT_FMA: same kernels built with FMA contraction; not bitwise"}
    sec = ms / 1e3
    mcups = cells * steps / sec / 1e6
    nstage = {0: 1, 1: 2, 2: 4, 3: 10}[p.order]
    launches = steps * nstage
    avg_launch_s = sec / launches
    bytes_launch = cells * bytes_per_cell_step(nv, p.order) / nstage
    achieved = bytes_launch / avg_launch_s / 1e9
    out = {
        "metric": METRIC, "value": round(mcups, 3), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(ms / steps, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": wname, "cells": cells, "nvar": nv, "order": p.order,
                   "path": "RHSFunctionPrim" if p.recons_prim else "RHSFunction",
                   "scheme": "WENO3+Rusanov+SSPRK(4,3)" if p.order == 2 else f"order{p.order}",
                   "cfl": p.cfl, "tile_rows": hrows, "tiles": ntiles, "graph": bool(args.graph),
                   "kernel_build": "exact"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": load_traffic(wname),
                     "kernel": "stage_kernel", "avg_launch_us": round(avg_launch_s * 1e6, 3),
                     "algorithmic_bytes_per_launch": int(bytes_launch)},
        "valu_roofline": valu_roofline(wname, avg_launch_s),
        "final_time": st.t,
    }
    if rhs_s is not None:
        out["pcie_inclusive"] = {"rhs_call_ms": round(rhs_s * 1e3, 4), "nstage": nstage,
                                 "equivalent_value": round(cells / (rhs_s * nstage) / 1e6, 3),
                                 "note": "hype_rhs on host buffers (upload U, download RHS), per RK stage"}
    if fma_variant is not None:
        out["fma_variant"] = fma_variant
    if with_cpu and not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, U0, nv)
    return out


SUB_KEYS = ("value", "unit", "steps", "ms_per_step", "config", "roofline", "valu_roofline", "final_time",
            "pcie_inclusive", "fma_variant", "cpu_baseline")
SERVED_KEYS = ("value", "unit", "config", "roofline", "latency_s", "one_at_a_time", "speedup_vs_one_at_a_time",
               "identical_to_solo", "inadmissible_requests")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps of the headline configuration (default: its own step count)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="all", choices=["all", "batched", "kapila", "euler", "bn", "served"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--hrows", type=int, default=0)
    ap.add_argument("--instances", type=int, default=64, help="batched: number of instances")
    ap.add_argument("--check", type=int, default=64, help="batched: instances compared with solo runs")
    ap.add_argument("--requests", type=int, default=64, help="served: number of requests of the trace")
    ap.add_argument("--rate", type=float, default=20.0, help="served: Poisson arrival rate per second")
    ap.add_argument("--policy", choices=sorted(POLICIES), default="shortest",
                    help="served: admission and run order (least remaining work first, or arrival order)")
    ap.add_argument("--active-cells", type=int, default=0,
                    help="served: cells advanced per round (0 = enough to fill the device, -1 = no cap)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-served", action="store_true",
                    help="default run: skip the served stream's sub-record")
    ap.add_argument("--no-pcie", action="store_true",
                    help="single-grid configs: skip timing the RHS callback on host buffers")
    ap.add_argument("--no-solo", action="store_true",
                    help="batched and served: skip the one-at-a-time comparison")
    ap.add_argument("--no-fma", action="store_true",
                    help="single-grid configs: skip timing the FMA-contracted build (HYPE_OPT_FMA) beside the exact one")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU (independent instances would map one per GPU; "
                                   "a single grid would split into blocks with one halo exchange per RK "
                                   "stage, not built here), so --gpus must be 1"}))
        sys.exit(2)
    if args.config == "served":
        out = run_served(args)
    elif args.config == "batched":
        out = run_batched(args)
    elif args.config == "all":
        out = run_batched(args)
        # the fixed grids at their own step counts, each with its roofline
        out["configs"] = {}
        for name in ("kapila", "euler", "bn"):
            rec = run_single(name, args, fma=(name == "kapila"), pcie=(name == "kapila"))
            out["configs"][rec["config"]["workload"]] = {k: rec[k] for k in SUB_KEYS if k in rec}
        # the served stream's first 64 requests, against one at a time
        if not args.no_served:
            rec = run_served(argparse.Namespace(**{**vars(args), "requests": 64}))
            out["configs"][rec["config"]["workload"]] = {k: rec[k] for k in SERVED_KEYS if k in rec}
    else:
        out = run_single(args.config, args, steps=args.steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
