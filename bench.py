#!/usr/bin/env python3
"""Throughput of the CompactLES per-step path on one MI355X.

Workload: the Taylor-Green vortex of the reference quick start (README.md:33-44)
at 256^3 -- periodic [0,2pi)^3, one species gamma = 1.4, rho = 1, p = 71.43,
u = (sin x cos y, -cos x sin y, 0), Lele C6 derivatives, C8 filter every step,
Cook artificial properties on, cfl 0.5 -- advanced by full run! steps
(boundary conditions, max_rate with its host-side dt decision, five LSRK(5,4)
stages, filter_state) with the state resident in HBM.  The metric is
Mpt-steps/s = 256^3 * steps / wall seconds.

Prints one JSON line.  Also reports, from a separate profiling pass, the HIP
event timing of the dominant kernel -- the z-sweep C6 flux divergence fused
with the LSRK(5,4) register update, 48 algorithmic bytes per point (flux and
partial dQ read, du and Q read and written) -- against the 8 TB/s HBM peak,
and the oracle (C restatement of the reference CPU path) timed on one step of
the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "compactles.jl_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

HBM_PEAK_GBPS = 8000.0
ALG_BYTES_PER_POINT_STEP = 8088.0   # reference-dataflow byte model, TGV art on, FP64
DOMINANT = ("fluxes",)
DOMINANT_KERNEL = {"fp64": "cles::k_fluxes<double>", "fp32": "cles::k_fluxes<float>"}


def tgv_problem(ch, n, precision):
    import numpy as np

    def ic(X, Y, Z):
        return ch.Prim(u=(np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y), 0.0 * X), p=71.43, rho=1.0)

    prob = ch.Problem(domain=((0.0, 2 * np.pi),) * 3, bcs=((ch.PeriodicBC(), ch.PeriodicBC()),) * 3, ic=ic,
                      name="taylor-green")
    num = ch.Numerics(n_global=(n, n, n), precision=precision, cfl=0.5)
    return prob, num


def read_traffic(n, precision, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy), or None."""
    path = os.path.join(ROOT, "profiles", f"tgv{n}_{precision}_hbm_counters.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch\[" + re.escape(kernel) + r"\]:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(n, steps=1):
    """The oracle on the host cores: one run! step of the same workload."""
    import numpy as np

    import cases
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    L = (2 * np.pi,) * 3
    X, Y, Z = cases.grid((n, n, n), L)
    q = cases.conserved(cases.tgv_prim(X, Y, Z))
    del X, Y, Z
    s = cases.oracle_solver((n, n, n), L, q, art=True, cfl=0.5)
    del q
    t0 = time.perf_counter()
    s.run(1e9, steps)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt / 1e6, "unit": "Mpt-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement of the reference CPU path, OpenMP, FP64) TGV {n}^3, "
                      f"{steps} full run! step(s), {dt:.2f} s"}


def counters_run(ch, ctx, n, steps):
    """Workload of the rocprofv3 --pmc passes: `steps` full run! steps of the TGV
    case, then a 1 GiB streaming copy three times (the counter calibration)."""
    import ctypes as C
    prob, num = tgv_problem(ch, n, "fp64")
    solver, _ = ch.setup(prob, num, ctx)
    solver.run(1e9, steps)
    ctx.synchronize()
    solver.close()
    nbytes = 1 << 30
    a = ch.DeviceArray(ctx, (nbytes // 8,))
    b = ch.DeviceArray(ctx, (nbytes // 8,))
    a.zero()
    for _ in range(3):
        ch._lib.check(ctx.lib.cles_memcpy_d2d(ctx.handle, b.ptr, a.ptr, C.c_size_t(nbytes)))
    ctx.synchronize()
    print(json.dumps({"kernel": DOMINANT_KERNEL["fp64"], "steps": steps, "copy_bytes_read": nbytes,
                      "copy_bytes_written": nbytes, "copy_launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--counters", type=int, default=0, metavar="STEPS",
                    help="run only STEPS run! steps and a 1 GiB streaming copy 3 times (the workload of the "
                         "rocprofv3 --pmc byte-counter passes), print a short JSON line and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design in DESIGN.md, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import compactles_hip as ch
    ctx = ch.Context(0)
    n = args.n
    if args.counters:
        counters_run(ch, ctx, n, args.counters)
        return
    prob, num = tgv_problem(ch, n, args.precision)
    solver, _ = ch.setup(prob, num, ctx)
    solver.run(1e9, args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    solver.run(1e9, args.warmup + args.steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    npts = n ** 3
    value = npts * args.steps / wall / 1e6
    ms_per_step = wall / args.steps * 1e3

    # live per-phase HIP-event timing on the solver stream (separate pass)
    solver.profile(True)
    solver.run(1e9, args.warmup + args.steps + args.profile_steps)
    phases = solver.profile_read()
    solver.profile(False)
    ms = sum(phases[p][0] for p in DOMINANT if p in phases)
    launches = sum(phases[p][1] for p in DOMINANT if p in phases)
    bytes_ = sum(phases[p][2] for p in DOMINANT if p in phases)
    avg_ms = ms / max(launches, 1)
    alg_per_launch = bytes_ / max(launches, 1)
    achieved = alg_per_launch / (avg_ms * 1e-3) / 1e9 if launches else 0.0
    total_ms = sum(v[0] for v in phases.values())
    step_alg = ALG_BYTES_PER_POINT_STEP * npts
    out = {
        "metric": f"Mpt-steps/s TGV {n}^3 {args.precision.upper()}",
        "value": round(value, 3),
        "unit": "Mpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-mixed (f64 state, RK update, reductions)",
        "data": "synthetic (Taylor-Green initial state, no checkpoint)",
        "config": {"workload": f"Taylor-Green vortex {n}^3 periodic, full run! step (max_rate + 5 LSRK(5,4) "
                               "stages + filter_state), C6 derivative, C8 filter, Cook artificial "
                               "properties on, cfl 0.5", "n": n, "precision": args.precision},
        "roofline": {"bound": "hbm", "kernel": DOMINANT_KERNEL[args.precision], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": read_traffic(n, args.precision, DOMINANT_KERNEL[args.precision]),
                     "avg_launch_us": round(avg_ms * 1e3, 2),
                     "alg_bytes_per_launch": alg_per_launch, "launches_profiled": launches},
        "step_roofline": {"alg_bytes_per_step": step_alg,
                          "achieved_GBps": round(step_alg / (ms_per_step * 1e-3) / 1e9, 1),
                          "frac": round(step_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)},
        "phases_ms_per_step": {k: round(v[0] / args.profile_steps, 3) for k, v in
                               sorted(phases.items(), key=lambda kv: -kv[1][0])},
        "phases_GBps": {k: round(v[2] / (v[0] * 1e-3) / 1e9, 1) for k, v in phases.items() if v[0] > 0},
        "profiled_kernel_ms_per_step": round(total_ms / args.profile_steps, 3),
    }
    solver.close()
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
