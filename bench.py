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
sys.path[:0] = [os.path.join(ROOT, "compactles.jl_amd")]

HBM_PEAK_GBPS = 8000.0
# the kernel with the largest share of the step's kernel time in each form
# (rocprofv3 --stats), with the solver phase that launches it
DOMINANT = {"fp64": "fluxes", "fp32": "line_div_lsrk_z"}
DOMINANT_KERNEL = {"fp64": "cles::k_fluxes<double, true, true>", "fp32": "cles::k_line_yz<float, 1, 2, 0, 3, true, 32, false>"}


def tgv_problem(ch, n, precision):
    from compactles_hip import workloads
    return workloads.tgv_problem(ch, n, precision)


def read_traffic(n, precision, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy), or None."""
    path = os.path.join(ROOT, "profiles", f"tgv{n}_{precision}_hbm_counters.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"(?m)^traffic_bytes_per_launch\[" + re.escape(kernel) + r"\]:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, steps=3, warmup=1):
    """The oracle (the C restatement of the reference CPU path, the checker of
    the tests, loaded only here) on the host cores: `warmup` untimed and
    `steps` timed run! steps of the same workload."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    from compactles_hip import workloads as W
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    L = (2 * np.pi,) * 3
    X, Y, Z = W.grid((n, n, n), L)
    q = W.conserved(W.tgv_prim(X, Y, Z))
    del X, Y, Z
    s = O.Solver(O.make_desc((n, n, n), L, art=True, cfl=0.5))
    for c, qc in enumerate(q):
        s.Q[s.halo[0]:s.halo[0] + n, s.halo[1]:s.halo[1] + n, s.halo[2]:s.halo[2] + n, c] = qc
    del q
    s.run(1e9, warmup)
    t0 = time.perf_counter()
    s.run(1e9, steps)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt / 1e6, "unit": "Mpt-steps/s", "cores": threads, "kind": "port",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"oracle (C restatement of the reference CPU path, OpenMP, {threads} threads, FP64) TGV "
                      f"{n}^3, {steps} full run! steps after {warmup} untimed, {dt:.2f} s"}


def counters_run(ch, ctx, n, steps, precision):
    """Workload of the rocprofv3 --pmc passes: `steps` full run! steps of the TGV
    case, then a 1 GiB streaming copy three times (the counter calibration)."""
    import ctypes as C
    prob, num = tgv_problem(ch, n, precision)
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
    print(json.dumps({"kernel": DOMINANT_KERNEL[precision], "steps": steps, "copy_bytes_read": nbytes,
                      "copy_bytes_written": nbytes, "copy_launches": 3}), flush=True)


def step_passes(S, D, nscbc_ops=0):
    """Per-step passes of the reference-dataflow byte model, split into
    (FP64-stored, working-precision) passes.  Per RHS: primitives 2S+12
    (C reads of Q), velocity gradients 6D plus the reference's zero-fill of
    the 3(3-D) collapsed ones, artificial properties 30+4D+[S>1]S(5+2D), scalar
    gradients 2D(1+S), fluxes 35+8S (one read of Q), divergence C(3D-1), LSRK
    update 5C (du and Q read and written), 2 per NSCBC line op; stage 1 skips
    the primitives; max_rate 4S+21 (C reads of Q); filter_state 2CD on Q.  In
    FP64 every pass is 8 bytes; in the FP32-mixed form the working passes are
    4 bytes."""
    C = S + 4
    p1 = (C, 8 + S)
    rhs_rest = (1 + 4 * C, 6 * D + 3 * (3 - D) + 30 + 4 * D + (S * (5 + 2 * D) if S > 1 else 0)
                + 2 * D * (1 + S) + 34 + 8 * S + C * (3 * D - 1) + C + 2 * nscbc_ops)
    rhs = (p1[0] + rhs_rest[0], p1[1] + rhs_rest[1])
    rate = (C, 4 * S + 21 - C)
    filt = (2 * C * D, 0)
    f64 = 4 * rhs[0] + rhs_rest[0] + rate[0] + filt[0]
    wp = 4 * rhs[1] + rhs_rest[1] + rate[1] + filt[1]
    return f64, wp


def alg_bytes_per_point_step(S, D, precision, nscbc_ops=0):
    f64, wp = step_passes(S, D, nscbc_ops)
    return 8.0 * f64 + (8.0 if precision == "fp64" else 4.0) * wp


def measure(ch, ctx, n, precision, steps, warmup, profile_steps, run_mode="host_dt"):
    prob, num = tgv_problem(ch, n, precision)
    solver, _ = ch.setup(prob, num, ctx)
    solver.set_run_mode(run_mode)
    solver.run(1e9, warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    solver.run(1e9, warmup + steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    npts = n ** 3
    value = npts * steps / wall / 1e6
    ms_per_step = wall / steps * 1e3

    # live per-phase HIP-event timing on the solver stream (separate pass)
    solver.profile(True)
    solver.run(1e9, warmup + steps + profile_steps)
    phases = solver.profile_read()
    solver.profile(False)
    ke, ens = solver.diagnostics()
    ms, launches, bytes_ = phases.get(DOMINANT[precision], (0.0, 0, 0.0))
    avg_ms = ms / max(launches, 1)
    alg_per_launch = bytes_ / max(launches, 1)
    achieved = alg_per_launch / (avg_ms * 1e-3) / 1e9 if launches else 0.0
    total_ms = sum(v[0] for v in phases.values())
    step_alg = alg_bytes_per_point_step(1, 3, precision) * npts
    solver.close()
    return {
        "value": round(value, 3), "ms_per_step": round(ms_per_step, 4), "ke": ke, "enstrophy": ens,
        "steps_total": warmup + steps + profile_steps,
        "roofline": {"bound": "hbm", "kernel": DOMINANT_KERNEL[precision], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": read_traffic(n, precision, DOMINANT_KERNEL[precision]),
                     "avg_launch_us": round(avg_ms * 1e3, 2),
                     "alg_bytes_per_launch": alg_per_launch, "launches_profiled": launches},
        "step_roofline": {"alg_bytes_per_point_step": alg_bytes_per_point_step(1, 3, precision),
                          "alg_bytes_per_step": step_alg,
                          "achieved_GBps": round(step_alg / (ms_per_step * 1e-3) / 1e9, 1),
                          "frac": round(step_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)},
        "phases_ms_per_step": {k: round(v[0] / profile_steps, 3) for k, v in
                               sorted(phases.items(), key=lambda kv: -kv[1][0])},
        "phases_GBps": {k: round(v[2] / (v[0] * 1e-3) / 1e9, 1) for k, v in phases.items() if v[0] > 0},
        "profiled_kernel_ms_per_step": round(total_ms / profile_steps, 3),
    }


def measure_tubes(ch, ctx, count, steps, warmup, precision, plain_sample=32, run_mode="step_graph"):
    """The batched form: `count` independent one-dimensional two-gas shock
    tubes (N_i ~ U{128..2048}, p_L ~ U[2, 20], rho_L ~ U[0.1, 10] from one
    seeded generator; He | CO2, Dirichlet ends, art on, cfl 0.4, filter every
    step) advanced together, each with its own dt.  In step-graph mode the
    warm-up covers at least one 8-iteration replay, so the graph is captured
    before the timed steps."""
    from compactles_hip import workloads as W
    specs = W.tube_specs(count, seed=20261202)
    npts = sum(sp["n"] for sp in specs)
    t0 = time.perf_counter()
    b = W.tube_batch(ctx, specs, precision=precision)
    setup_s = time.perf_counter() - t0
    b.set_run_mode(run_mode)
    b.run(max_steps=max(warmup, 8) if run_mode == "step_graph" else warmup)
    t0 = time.perf_counter()
    b.run(max_steps=steps)
    wall = time.perf_counter() - t0
    _, st, status = b.state()
    b.close()
    ms = wall / steps * 1e3
    alg = alg_bytes_per_point_step(2, 1, precision) * npts
    # the plain path on the same instances: a seeded sample of the tubes, each
    # through its own single-domain solver, one after another
    sample = specs[:plain_sample]
    t_plain, pts = 0.0, 0
    for sp in sample:
        s1 = W.tube_solver(ctx, sp, precision)
        s1.run(1e9, warmup)
        ctx.synchronize()
        t0 = time.perf_counter()
        s1.run(1e9, warmup + steps)
        ctx.synchronize()
        t_plain += time.perf_counter() - t0
        pts += sp["n"]
        s1.close()
    plain = pts * steps / t_plain / 1e6
    return {"value": round(npts * steps / wall / 1e6, 3), "unit": "Mpt-steps/s", "ms_per_step": round(ms, 4),
            "instance_steps_per_s": round(count * steps / wall, 1), "instances": count, "points": npts,
            "plain_path_same_instances": {"value": round(plain, 3), "unit": "Mpt-steps/s",
                                          "instances_sampled": len(sample),
                                          "batched_over_plain": round(npts * steps / wall / 1e6 / plain, 1)},
            "all_ok": bool((status == 0).all()), "setup_s": round(setup_s, 2), "run_mode": run_mode,
            "step_roofline": {"alg_bytes_per_point_step": alg_bytes_per_point_step(2, 1, precision),
                              "achieved_GBps": round(alg / (ms * 1e-3) / 1e9, 1),
                              "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)},
            "workload": f"{count} one-dimensional He|CO2 shock tubes, N_i ~ U{{128..2048}} (seed 20261202), "
                        "Dirichlet ends, art on, cfl 0.4, filter every step, per-instance dt"}


def measure_shock_2d(ch, ctx, precision, steps, warmup, n=(1024, 512, 1), tune=True):
    """The two-dimensional form: He|CO2 Mach-1.5 shock / perturbed interface on
    1024 x 512 (two species, NSCBC inflow and outflow in x, periodic y, art on,
    cfl 0.4, filter every step).  With tune, the line tiles are tuned for the
    grid first (outside the timed region)."""
    from compactles_hip import workloads as W
    prob, num, Q = W.shock_problem(ch, n, precision)
    if tune:
        ch.tune.tune(ctx, ch.SolverSpec(prob, num), reps=5)
    tiles = ch.tune.lookup(ch.SolverSpec(prob, num))
    s = ch.Solver(ctx, prob, num)
    s.upload(Q)
    s.run(1e9, warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    s.run(1e9, warmup + steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    s.close()
    npts = n[0] * n[1] * n[2]
    ms = wall / steps * 1e3
    # byte model with S = 2, D = 2 and the NSCBC inflow (p, rho) and outflow (p normal and transverse) line ops
    model = alg_bytes_per_point_step(2, 2, precision, nscbc_ops=4)
    return {"value": round(npts * steps / wall / 1e6, 3), "unit": "Mpt-steps/s", "ms_per_step": round(ms, 4),
            "step_roofline": {"alg_bytes_per_point_step": model,
                              "achieved_GBps": round(model * npts / (ms * 1e-3) / 1e9, 1),
                              "frac": round(model * npts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)},
            "line_tiles": {"chunks": list(tiles[0]), "tuned": list(tiles[1])},
            "workload": f"He|CO2 shock / interface {n[0]}x{n[1]}, NSCBC in / out, art on, cfl 0.4"}


def measure_operators(ch, ctx, n=256, iters=20, seed=20261202):
    """The operator form: one compact-operator application (apply_along!) over
    an n^3 periodic field of seeded Fourier modes (|k| <= 32, amplitude
    |k|^(-5/3), unit RMS) for the C6 and C10 first derivatives and the C8
    filter along x, y and z, in FP64 and in FP32 on the same input (cast once).
    Each is timed with HIP events over `iters` back-to-back launches; the
    algorithmic bytes are one read and one write of the field (16 B/pt FP64,
    8 B/pt FP32)."""
    import numpy as np

    from compactles_hip import workloads as W
    f64 = W.fourier_field((n, n, n), kmax=32, seed=seed)
    dec = ch.Decomp((n, n, n), (True, True, True), n_halo=0)
    h = 2 * np.pi / n
    out = {}
    for prec, dt in (("fp64", np.float64), ("fp32", np.float32)):
        f = ch.DeviceArray.from_numpy(ctx, f64, dt)
        o = ch.DeviceArray(ctx, (n, n, n), dt)
        esz = np.dtype(dt).itemsize
        for name, sch in (("C6", ch.lele_d1_6()), ("C10", ch.lele_d1_10()), ("C8_filter", ch.compact_filter())):
            row = {}
            for dim, ax in ((1, "x"), (2, "y"), (3, "z")):
                plan = ch.plan_direction(ctx, dec, sch, dim, h, dtype=dt)
                ch.apply_along_timed(o, plan, f, dec, 3)
                ms = ch.apply_along_timed(o, plan, f, dec, iters)
                gbps = 2 * esz * n ** 3 / (ms * 1e-3) / 1e9
                row[ax] = {"us": round(ms * 1e3, 2), "GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4)}
                del plan
            out.setdefault(name, {})[prec] = row
        f.free()
        o.free()
    return {"per_operator": out, "alg_bytes_per_point": {"fp64": 16, "fp32": 8},
            "workload": f"{n}^3 periodic field of seeded Fourier modes (|k| <= 32, |k|^-5/3, seed {seed}), "
                        "one apply_along! per operator and direction, timed over "
                        f"{iters} launches with HIP events"}


def measure_serving(ch, ctx=None, count=32, rate_per_s=1.5, seed=20261202, budget=64 << 30, runners=4):
    """The serving form: `count` requests with Poisson arrivals (rate
    `rate_per_s`, seeded) mixing Taylor-Green cubes of 32^3..128^3 to t = 0.05
    and two-gas shock / interface grids of 256^2..1024x512 to t = 0.03, admitted
    by device footprint under an HBM budget and run concurrently.  With a
    context, the line tiles of every distinct shape are tuned first."""
    import numpy as np

    from compactles_hip import workloads as W
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_per_s, count))
    kinds = []
    for _ in range(count):
        if rng.random() < 0.5:
            kinds.append(("tgv", int(rng.choice([32, 48, 64, 96, 128]))))
        else:
            kinds.append(("shock", [(256, 256, 1), (512, 256, 1), (1024, 512, 1)][int(rng.integers(0, 3))]))
    reqs = []
    for kind, n in kinds:
        if kind == "tgv":
            prob, num = tgv_problem(ch, n, "fp64")
            spec = ch.SolverSpec(prob, num)
            reqs.append((prob, num, ch.initialize_(spec, spec.allocate_state(), prob.ic), 0.05))
        else:
            prob, num, Q = W.shock_problem(ch, n)
            reqs.append((prob, num, Q, 0.03))
    tuned = 0
    if ctx is not None:
        seen = set()
        for prob, num, _, _ in reqs:
            spec = ch.SolverSpec(prob, num)
            key = (spec.n, num.precision)
            if key not in seen:
                seen.add(key)
                ch.tune.tune(ctx, spec, reps=5)
        tuned = len(seen)
    sv = ch.Server(device=0, hbm_budget=budget, runners=runners)
    t0 = time.perf_counter()
    ids = []
    for (prob, num, Q, tf), at in zip(reqs, arrivals):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(sv.submit(prob, num, tfinal=tf, Q=Q))
    ok = all(sv.wait(rid) == ch.server.DONE for rid in ids)
    times = [sv.timing(rid) for rid in ids]
    sv.close()
    lat = np.array([d - s for s, a, d, fp in times])
    span = max(d for s, a, d, fp in times) - min(s for s, a, d, fp in times)
    # throughput at saturation (every request submitted at once) against the
    # plain path solving the same requests one after another
    sv = ch.Server(device=0, hbm_budget=budget, runners=runners)
    t0 = time.perf_counter()
    ids2 = [sv.submit(prob, num, tfinal=tf, Q=Q) for prob, num, Q, tf in reqs]
    ok = all(sv.wait(rid) == ch.server.DONE for rid in ids2) and ok
    sat = time.perf_counter() - t0
    sv.close()
    seq = 0.0
    if ctx is not None:
        for prob, num, Q, tf in reqs:
            t1 = time.perf_counter()
            s1 = ch.Solver(ctx, prob, num)
            s1.upload(Q)
            s1.run(tf)
            ctx.synchronize()
            s1.close()
            seq += time.perf_counter() - t1
    # launch sharing: same-shape 2-D requests as one Stack against one at a time
    stack = {}
    if ctx is not None:
        for n in ((1024, 512, 1), (256, 256, 1)):
            prob, num, Q = W.shock_problem(ch, n)
            row = {}
            for K in (1, 5):
                st = ch.Stack(ctx, prob, num, K)
                for i in range(K):
                    st.upload(i, Q)
                st.run(1e9, 5)
                ctx.synchronize()
                t1 = time.perf_counter()
                st.run(1e9, 25)
                ctx.synchronize()
                row[f"K{K}_Mpt_steps_s"] = round(n[0] * n[1] * K * 20 / (time.perf_counter() - t1) / 1e6, 1)
                st.close()
            stack[f"{n[0]}x{n[1]}"] = row
    return {"solves_per_s": round(count / span, 2), "p50_s": round(float(np.percentile(lat, 50)), 4),
            "p99_s": round(float(np.percentile(lat, 99)), 4), "requests": count, "arrival_rate_per_s": rate_per_s,
            "all_done": ok, "hbm_budget_GB": budget / 2 ** 30, "runners": runners, "tuned_shapes": tuned,
            "saturated_solves_per_s": round(count / sat, 2),
            "plain_path_sequential_solves_per_s": round(count / seq, 2) if seq > 0 else None,
            "stack_limit": 8, "stack_throughput": stack,
            "workload": "Poisson arrivals (seed 20261202): Taylor-Green 32^3..128^3 to t = 0.05 and He|CO2 "
                        "shock / interface 256^2..1024x512 (NSCBC in / out) to t = 0.03, FP64"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the FP32-mixed and batched forms that the default FP64 run also reports")
    ap.add_argument("--tubes", type=int, default=4096, help="instances of the batched form")
    ap.add_argument("--run-mode", choices=["step_graph", "device_dt", "host_dt"], default="host_dt",
                    help="how run! issues its steps (identical results)")
    ap.add_argument("--only", choices=["tubes", "serving", "shock2d", "operators"], default=None,
                    help="measure only the batched-tubes (FP64) or serving form and print its JSON (profiling runs)")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=2, help="steps of the per-phase event pass (>= 1)")
    ap.add_argument("--counters", type=int, default=0, metavar="STEPS",
                    help="run only STEPS run! steps and a 1 GiB streaming copy 3 times (the workload of the "
                         "rocprofv3 --pmc byte-counter passes), print a short JSON line and exit")
    args = ap.parse_args()
    if args.profile_steps < 1:
        ap.error("--profile-steps must be at least 1 (the roofline figures come from that pass)")
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding over GPUs is designed, not built or measured)", file=sys.stderr)
        sys.exit(2)

    import compactles_hip as ch
    ctx = ch.Context(0)
    n = args.n
    if args.counters:
        counters_run(ch, ctx, n, args.counters, args.precision)
        return
    if args.only == "tubes":
        print(json.dumps(measure_tubes(ch, ctx, args.tubes, args.steps, args.warmup, args.precision)), flush=True)
        return
    if args.only == "shock2d":
        print(json.dumps(measure_shock_2d(ch, ctx, args.precision, args.steps, args.warmup)), flush=True)
        return
    if args.only == "serving":
        print(json.dumps(measure_serving(ch, ctx)), flush=True)
        return
    if args.only == "operators":
        print(json.dumps(measure_operators(ch, ctx, n)), flush=True)
        return
    primary = measure(ch, ctx, n, args.precision, args.steps, args.warmup, args.profile_steps, args.run_mode)
    out = {
        "metric": f"Mpt-steps/s TGV {n}^3 {args.precision.upper()}",
        "value": primary["value"],
        "unit": "Mpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": primary["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-mixed (f64 state, RK update, reductions)",
        "data": "synthetic (Taylor-Green initial state, no checkpoint)",
        "config": {"workload": f"Taylor-Green vortex {n}^3 periodic, full run! step (max_rate + 5 LSRK(5,4) "
                               "stages + filter_state), C6 derivative, C8 filter, Cook artificial "
                               "properties on, cfl 0.5", "n": n, "precision": args.precision,
                   "run_mode": args.run_mode},
    }
    for k in ("roofline", "step_roofline", "phases_ms_per_step", "phases_GBps", "profiled_kernel_ms_per_step"):
        out[k] = primary[k]
    if not args.no_forms and args.precision == "fp64":
        # the FP32-mixed form of the same workload, and its kinetic-energy
        # difference from the FP64 run after the same number of steps
        mixed = measure(ch, ctx, n, "fp32", args.steps, args.warmup, args.profile_steps, args.run_mode)
        out["forms"] = {"fp32_mixed": {
            "value": mixed["value"], "unit": "Mpt-steps/s", "ms_per_step": mixed["ms_per_step"],
            "step_roofline": mixed["step_roofline"],
            "roofline": {k: mixed["roofline"][k] for k in ("kernel", "achieved", "frac", "avg_launch_us")},
            "ke_relerr_vs_fp64_same_steps": abs(mixed["ke"] - primary["ke"]) / primary["ke"],
            "enstrophy_relerr_vs_fp64_same_steps": abs(mixed["enstrophy"] - primary["enstrophy"]) /
            primary["enstrophy"],
            "steps_compared": mixed["steps_total"]}}
        out["forms"]["batched_tubes"] = {p: measure_tubes(ch, ctx, args.tubes, args.steps, args.warmup, p)
                                         for p in ("fp64", "fp32")}
        out["forms"]["shock_interface_2d"] = {p: measure_shock_2d(ch, ctx, p, args.steps, args.warmup)
                                              for p in ("fp64", "fp32")}
        out["forms"]["serving"] = measure_serving(ch, ctx)
        out["forms"]["operators"] = measure_operators(ch, ctx, n)
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
