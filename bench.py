#!/usr/bin/env python3
"""Benchmark of the (T) perturbative-triples hot path on one MI355X.

Workload: the plain FP64 (T) energy of a synthetic closed-shell system with
No=40 occupied and Nv=400 virtual orbitals (10,746,400 tuples a<=b<=c), all
tensors generated on the device from a seed and resident in HBM before the
timed region starts.  A step is one slice of the tuple space: the tuples whose
first index falls in a contiguous range of index blocks, with the slices
balanced so that 16 steps cover the whole space once.  The default run
(16 timed steps) therefore computes the full E(T) of the system; with more
steps the E(T) reported is that of the first pass.

Prints one JSON line with tuples/s, TFLOP/s (doubles FLOP definition of the
reference, 12 No^3 (No+Nv) per tuple), the FP64-MFMA roofline of the doubles
kernel measured with HIP events on the stream it runs on, and a CPU baseline
(the oracle's port of the reference CPU path on the host cores, a bounded
sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = ("(T) tuples/s & TFLOP/s, FP64 vs mixed @|ΔE|≤1e-7 Eh; "
          "serve p50/p99")
PEAK_TFLOPS = {"fp64": 78.6, "split_fp32": 157.3, "fp32": 157.3,
               "bf16x3": 2500.0}
# MFMA work per algorithmic FLOP: bf16x3 forms six bf16 part products
MFMA_PER_FLOP = {"fp64": 1.0, "split_fp32": 1.0, "fp32": 1.0, "bf16x3": 6.0}
KERNEL_NAME = {"fp64": "doubles", "split_fp32": "doubles_mixed",
               "fp32": "doubles_fp32", "bf16x3": "doubles_bf16x3"}


def neumaier(values):
    """Compensated sum, in order."""
    s = c = 0.0
    for x in values:
        t = s + x
        c += (s - t) + x if abs(s) >= abs(x) else (x - t) + s
        s = t
    return s + c


def log(msg: str):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def n_tuples(nv: int) -> int:
    return nv * (nv + 1) * (nv + 2) // 6 - nv


def tuples_before(a: int, nv: int) -> int:
    """Tuples whose first index is < a (lexicographic order)."""
    tot = 0
    for x in range(a):
        m = nv - x
        tot += m * (m + 1) // 2 - 1
    return tot


def step_slices(nv: int, nb: int, nsteps: int):
    """Canonical [first, count) ranges aligned to index blocks of edge nb,
    balanced by tuple count, covering the whole tuple space."""
    nblk = (nv + nb - 1) // nb
    bounds = [tuples_before(min(b * nb, nv), nv) for b in range(nblk + 1)]
    total = bounds[-1]
    cuts = [0]
    for s in range(1, nsteps):
        target = total * s / nsteps
        b = min(range(len(bounds)), key=lambda q: abs(bounds[q] - target))
        cuts.append(max(b, cuts[-1]))
    cuts.append(nblk)
    out = []
    for s in range(nsteps):
        lo, hi = bounds[cuts[s]], bounds[cuts[s + 1]]
        if hi > lo:
            out.append((lo, hi - lo))
    return out


def measured_traffic(kernel: str, no: int, nv: int):
    """HBM bytes per launch of `kernel` at this shape, as measured with
    rocprofv3 FETCH_SIZE + WRITE_SIZE runs and recorded in profiles/ (None
    when no such measurement exists)."""
    import glob
    import re

    key = f"roofline_traffic_bytes_per_launch {kernel} No={no} Nv={nv}:"
    for path in sorted(glob.glob(os.path.join(HERE, "profiles", "*.md"))):
        for line in open(path):
            if line.startswith(key):
                m = re.search(r":\s*([0-9.eE+]+)", line[len(key) - 1:])
                if m:
                    return float(m.group(1))
    return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _timed_oracle(O, no, nv, t, first, count, threads):
    t0 = time.time()
    e, _ = O.run(no, nv, t, first=first, count=count, use_gemm=True, nthreads=threads)
    return time.time() - t0, e


def _sized_sample(O, no, nv, t, threads, budget_s):
    """First K tuples (a-major order, as the reference's max_iterations),
    K sized from a short probe so that the tuple loop takes about budget_s.
    The oracle's one-time setup per run (cutting the T(x) slices of the
    whole Tabij) is timed with an empty range and taken out of the rate."""
    setup, _ = _timed_oracle(O, no, nv, t, 0, 0, threads)
    log(f"CPU baseline {threads} thread(s): setup {setup:.1f} s")
    # the probe grows until its tuple loop is well above the setup's noise,
    # so the rate it gives cannot be a noise artefact
    probe, loop = 4 * threads, 0.0
    while True:
        dt, _ = _timed_oracle(O, no, nv, t, 0, probe, threads)
        loop = dt - setup
        log(f"CPU baseline {threads} thread(s): probe {probe} tuples {dt:.1f} s")
        if loop >= max(1.0, 0.1 * budget_s) or probe >= n_tuples(nv):
            break
        probe = min(n_tuples(nv), 4 * probe)
    per = loop / probe
    count = int(max(probe, min(n_tuples(nv), budget_s / per)))
    count = max(threads, count // threads * threads)
    dt, _ = _timed_oracle(O, no, nv, t, 0, count, threads)
    log(f"CPU baseline {threads} thread(s): {count} tuples {dt:.1f} s")
    return count, max(dt - setup, 1e-6), setup


def cpu_baseline(no: int, nv: int, seed: int, budget_s: float):
    """The oracle (C port of the reference np=1 CPU path: per tuple 12 host
    BLAS dgemm, permuted accumulates, singles, energy) on the host cores: C1
    (No=10, Nv=50) to completion, and the first tuples of the GPU workload
    sized to about budget_s, each on 1 core and on the process's CPU share
    (tuple shards, single-threaded BLAS per thread)."""
    from oracle import oracle as O  # the checker / CPU baseline only

    blas = O.blas_init()
    affinity = len(os.sched_getaffinity(0))
    # worker pools stay within the CPU share of one GPU (16 threads)
    cores = max(1, min(affinity, 16))
    flop_tuple = 12.0 * no ** 3 * (no + nv)
    c1 = {}
    t1 = O.synthetic(10, 50, 1)
    for label, th in (("1_core", 1), ("all_cores", cores)):
        dt, e = _timed_oracle(O, 10, 50, t1, 0, n_tuples(50), th)
        log(f"CPU baseline No=10 Nv=50 on {th} thread(s): {dt:.1f} s")
        c1[label] = {"threads": th, "seconds": dt, "tuples_per_s": n_tuples(50) / dt,
                     "energy_Eh": e}
    del t1
    t0 = time.time()
    t = O.synthetic(no, nv, seed, nthreads=cores)
    gen_s = time.time() - t0
    log(f"CPU baseline No={no} Nv={nv} host tensors generated in {gen_s:.1f} s")
    samples = {}
    for label, th in (("1_core", 1), ("all_cores", cores)):
        count, dt, setup = _sized_sample(O, no, nv, t, th, budget_s)
        rate = count / dt
        samples[label] = {
            "threads": th, "tuples": count, "seconds": dt, "setup_s": setup,
            "tuples_per_s": rate,
            "tflops": rate * flop_tuple / 1e12,
            "extrapolated_full_space_s": n_tuples(nv) / rate,
        }
    del t
    rate = samples["all_cores"]["tuples_per_s"]
    return {
        "value": rate,
        "unit": "tuples/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"No={no} Nv={nv} seed={seed}: the first "
                   f"{samples['all_cores']['tuples']} tuples in lexicographic order "
                   f"({samples['all_cores']['seconds']:.1f} s) on {cores} tuple-shard "
                   f"threads, extrapolated linearly to the {n_tuples(nv)} tuples "
                   f"(per-tuple cost is uniform; the run's one-time slice setup "
                   f"excluded); oracle port with dgemm from "
                   f"{blas or 'plain C'}, single-threaded BLAS per thread; "
                   f"{gen_s:.1f} s host tensor generation untimed"),
        "tflops": rate * flop_tuple / 1e12,
        "cpu_model": cpu_model(),
        "affinity_cpus": affinity,
        "note": ("all-core runs use min(affinity, 16) threads, the host CPU "
                 "share of one GPU"),
        "one_core": samples["1_core"],
        "all_cores": samples["all_cores"],
        "c1_full": c1,
    }


def bench_batched(ctx, n_systems: int, seed: int, shape=None):
    """Batched form: n seeded small systems, autotuned, one persistent launch
    (the seeded workload's shapes, or n systems of one (No, Nv) shape)."""
    from atrip_amd import workloads as W

    if shape:
        shapes = [(shape[0], shape[1], 1000 + s) for s in range(n_systems)]
    else:
        shapes = W.batch_shapes(n_systems, seed=seed)
    t0 = time.time()
    systems = [ctx.upload_synthetic(no, nv, s) for no, nv, s in shapes]
    up_s = time.time() - t0
    log(f"{n_systems} systems resident; tuning")
    t0 = time.time()
    ctx.batch_tune(systems, measure=True)
    tune_s = time.time() - t0
    ctx.batch_run(systems)  # warm
    ctx.reset_stats()
    ctx.set_option("timing", 1)
    t0 = time.perf_counter()
    e = ctx.batch_run(systems)
    wall = time.perf_counter() - t0
    ctx.set_option("timing", 0)
    k = ctx.kernel_stats("batched")
    tuples = sum(W.n_tuples(nv) for _, nv, _ in shapes)
    flop = sum(W.doubles_flop(no, nv) for no, nv, _ in shapes)
    for s_ in systems:
        ctx.release(s_)
    return {
        "workload": ((f"{n_systems} systems, No~U{{4..16}}, Nv~U{{20..120}}, "
                      f"shape seed {seed}, tensor seeds 1000+s; FP64") if not shape else
                     (f"{n_systems} systems of No={shape[0]} Nv={shape[1]}, "
                      f"tensor seeds 1000+s; FP64")),
        "tuples": tuples,
        "value": tuples / wall, "unit": "tuples/s",
        "tflops": flop / wall / 1e12,
        "kernel_ms": k.total_ms,
        "kernel_tflops": flop / (k.total_ms * 1e-3) / 1e12 if k.total_ms else None,
        "kernel_frac_of_fp64_peak": (flop / (k.total_ms * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"]) if k.total_ms else None,
        "wall_s": wall, "upload_s": up_s, "tune_s": tune_s,
        "energy_sum_Eh": float(sum(e)),
    }


def bench_serving(ctx, n_requests: int, rate: float, seed: int,
                  precisions=("fp64", "bf16x3", "split_fp32")):
    """The same Poisson request stream served once per precision; the mixed
    runs also report each request's relative energy difference to FP64."""
    from atrip_amd import serve

    out = {}
    ref = None
    window = ctx.get_option("window")
    ctx.set_option("window", 0)  # plain-path requests: edge per request shape
    for prec in precisions:
        log(f"serving form, {prec}")
        r = serve.run_poisson(ctx, n_requests, rate=rate, seed=seed,
                              hbm_budget_gb=200.0, iter_ms=100.0, precision=prec,
                              log=log)
        e = r.pop("energies")
        r["workload"] = (f"{n_requests} requests, Poisson {rate} req/s, No log-U[8,40], "
                         f"Nv log-U[40,320], seed {seed}, tensor seeds 5000+r; "
                         f"HBM budget 200 GB; {prec}; requests with No >= "
                         f"{ctx.get_option('serve_plain_no')} on the plain path "
                         f"(window edge per request shape), the rest in the "
                         f"persistent batched launch")
        if ref is None:
            ref = e
        else:
            rel = [abs(e[q] - ref[q]) / abs(ref[q]) for q in e if q in ref and ref[q]]
            r["max_rel_dE_vs_fp64"] = max(rel) if rel else None
        out[prec] = r
    ctx.set_option("window", window)
    return out


C3_ROWS = None  # window rows (block pairs) at No=60 Nv=600, window 8


def bench_c3(ctx, no: int, nv: int, seed: int, stride: int, window: int,
             scale: float):
    """Plain FP64 against the mixed forms at No=60 Nv=600 on a stratified
    sample: every stride-th row of windows (all windows of one block pair
    (alpha, beta)), spanning the whole first-index range.  Each form runs on
    its own resident copy (FP64 operands; split-only; FP32-only;
    bf16x3-only).  Then the same sample with every amplitude and integral
    times `scale` (full-space E(T) of molecular magnitude) in FP64, bf16x3
    and split FP32: the absolute |dE| at molecular size."""
    nblk = (nv + window - 1) // window
    rows = nblk * (nblk + 1) // 2
    flop_tuple = 12.0 * no ** 3 * (no + nv)
    out = {}
    ref = None
    tuples = None
    for prec, store in (("fp64", 0), ("split_fp32", 1), ("fp32", 2), ("bf16x3", 3)):
        log(f"No={no} Nv={nv} stratified sample, {prec}")
        ctx.set_option("operand_store", store)
        try:
            t0 = time.time()
            s_ = ctx.upload_synthetic(no, nv, seed)
            up = time.time() - t0
        finally:
            ctx.set_option("operand_store", 0)
        try:
            ctx.run(s_, prec, first=0, count=1)  # warm
            kname = KERNEL_NAME[prec]
            ctx.reset_stats()
            ctx.set_option("timing", 1)
            t0 = time.perf_counter()
            e, _, nt = ctx.run_windows(s_, stride, 0, prec)
            dt = time.perf_counter() - t0
            ctx.set_option("timing", 0)
            k = ctx.kernel_stats(kname)
            info = ctx.sys_info(s_)
        finally:
            ctx.release(s_)
        tuples = nt
        ent = {"value": nt / dt, "unit": "tuples/s",
               "tflops": nt * flop_tuple / dt / 1e12,
               "doubles_kernel_tflops": k.flop / (k.total_ms * 1e-3) / 1e12 if k.total_ms else None,
               "energy_Eh": e, "resident_gb": info["bytes"] / 1e9, "upload_s": up,
               "seconds": dt}
        if k.total_ms:
            ent["doubles_kernel_mfma_frac"] = (ent["doubles_kernel_tflops"] * MFMA_PER_FLOP[prec]
                                               / PEAK_TFLOPS[prec])
        if ref is None:
            ref = ent
        else:
            ent["speedup_vs_fp64"] = ent["value"] / ref["value"]
            ent["abs_dE_Eh"] = abs(e - ref["energy_Eh"])
            ent["rel_dE"] = abs(e - ref["energy_Eh"]) / abs(ref["energy_Eh"])
        out[prec] = ent
    out["workload"] = (f"(T) No={no} Nv={nv} seed {seed}: stratified sample, every "
                       f"{stride}-th of the {rows} rows of windows (block pairs "
                       f"(alpha,beta) of edge {window}), spanning a in [0,{nv}): "
                       f"{tuples} tuples = {tuples / n_tuples(nv):.2%} of {n_tuples(nv)}")
    out["tuples"] = tuples
    out["fraction"] = tuples / n_tuples(nv)
    scaled = {"scale": scale}
    e64 = None
    for prec, store in (("fp64", 0), ("bf16x3", 3), ("split_fp32", 1)):
        log(f"No={no} Nv={nv} stratified sample, inputs x {scale}, {prec}")
        ctx.set_option("operand_store", store)
        try:
            s_ = ctx.upload_synthetic_scaled(no, nv, seed, scale)
        finally:
            ctx.set_option("operand_store", 0)
        try:
            e, _, _ = ctx.run_windows(s_, stride, 0, prec)
        finally:
            ctx.release(s_)
        if e64 is None:
            e64 = e
            scaled["fp64_energy_Eh"] = e
            scaled["scaling_law_rel_err"] = (abs(e - ref["energy_Eh"] * scale ** 4)
                                             / abs(ref["energy_Eh"] * scale ** 4))
        else:
            scaled[prec] = {"energy_Eh": e, "abs_dE_Eh": abs(e - e64),
                            "rel_dE": abs(e - e64) / abs(e64)}
    scaled["note"] = ("amplitudes and integrals times the scale: the full-space "
                      "E(T) of this shape is then about scale^4 x the unscaled "
                      "one (0.5 Eh order); |dE| is measured on the sample")
    out["molecular_scale"] = scaled
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no", type=int, default=40)
    ap.add_argument("--nv", type=int, default=400)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precision", default="fp64",
                    choices=["fp64", "split_fp32", "fp32", "bf16x3"])
    ap.add_argument("--window", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--operand-store", type=int, default=0, choices=[0, 1, 2, 3],
                    help="operand store of the timed system: 0 FP64 (mixed "
                         "copies made on first use), 1 split FP32, 2 FP32, "
                         "3 bf16x3 only (No=60 Nv=600 mixed runs)")
    ap.add_argument("--slices", type=int, default=16,
                    help="steps that cover the tuple space once")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="length of each CPU-baseline sample of the GPU workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision comparison pass")
    ap.add_argument("--mixed", default="split_fp32,bf16x3",
                    help="comma-separated forms of the comparison passes "
                         "(split_fp32, fp32, bf16x3)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-systems", type=int, default=256)
    ap.add_argument("--batch-shape", default=None,
                    help="No,Nv: batched systems all of this shape instead of "
                         "the seeded workload's")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the No=60 Nv=600 FP64-vs-mixed stratified sample")
    ap.add_argument("--c3-stride", type=int, default=55,
                    help="every n-th row of windows in the No=60 Nv=600 sample")
    ap.add_argument("--c3-scale", type=float, default=4e-3,
                    help="amplitude/integral scale of the molecular-size check")
    ap.add_argument("--step-offset", type=int, default=0,
                    help="first slice of the timed steps (to split one pass of "
                         "the tuple space over several runs)")
    args = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (tuples "
              "would shard by FLOP-balanced index ranges, one per GPU, joined by "
              "one scalar all-reduce of E(T)). Run with --gpus 1.", file=sys.stderr)
        return 2

    from atrip_amd import hip as H

    ctx = H.Context(0)
    ctx.set_option("window", args.window)
    ctx.set_option("batch", args.batch)
    ctx.set_option("tile", args.tile)
    t0 = time.time()
    ctx.set_option("operand_store", args.operand_store)
    try:
        sys_ = ctx.upload_synthetic(args.no, args.nv, args.seed)
    finally:
        ctx.set_option("operand_store", 0)
    gen_s = time.time() - t0
    slices = step_slices(args.nv, args.window, args.slices)

    for i in range(args.warmup):
        first, count = slices[(len(slices) - 1 - i) % len(slices)]
        ctx.run(sys_, args.precision, first=first, count=count)

    def step_range(i):
        return slices[(args.step_offset + i) % len(slices)]

    ctx.reset_stats()
    ctx.set_option("timing", 1)
    done = 0
    e_sum = 0.0
    pass_e = []  # energies of the first pass over the slices
    t_start = time.perf_counter()
    for i in range(args.steps):
        first, count = step_range(i)
        e, _ = ctx.run(sys_, args.precision, first=first, count=count)
        e_sum += e
        if i < len(slices):
            pass_e.append(e)
        done += count
        log(f"step {i + 1}/{args.steps}: {count} tuples")
    t_end = time.perf_counter()
    elapsed = t_end - t_start
    ctx.set_option("timing", 0)

    flop_tuple = 12.0 * args.no ** 3 * (args.no + args.nv)
    kname = KERNEL_NAME[args.precision]
    ks = ctx.kernel_stats(kname)
    es = ctx.kernel_stats("energy")
    achieved = ks.flop / (ks.total_ms * 1e-3) / 1e12 if ks.total_ms else 0.0
    peak = PEAK_TFLOPS[args.precision]
    value = done / elapsed

    mixed = None
    if args.precision == "fp64" and args.operand_store == 0 and not args.no_mixed:
        # the same steps in each mixed form on the same resident inputs
        mixed = {}
        for form in [f for f in args.mixed.split(",") if f]:
            if form not in ("split_fp32", "fp32", "bf16x3"):
                raise SystemExit(f"unknown mixed form {form!r}")
            ctx.run(sys_, form, first=slices[-1][0], count=slices[-1][1])
            ctx.reset_stats()
            ctx.set_option("timing", 1)
            m_sum = 0.0
            t0m = time.perf_counter()
            for i in range(args.steps):
                first, count = step_range(i)
                em, _ = ctx.run(sys_, form, first=first, count=count)
                m_sum += em
                log(f"{form} step {i + 1}/{args.steps}")
            m_elapsed = time.perf_counter() - t0m
            ctx.set_option("timing", 0)
            mk = ctx.kernel_stats(KERNEL_NAME[form])
            m_value = done / m_elapsed
            k_tf = mk.flop / (mk.total_ms * 1e-3) / 1e12 if mk.total_ms else 0.0
            mixed[form] = {
                "value": m_value,
                "unit": "tuples/s",
                "speedup_vs_fp64": m_value / value,
                "tflops": m_value * flop_tuple / 1e12,
                "doubles_kernel_tflops": k_tf,
                "doubles_kernel_mfma_frac": k_tf * MFMA_PER_FLOP[form] / PEAK_TFLOPS[form],
                "mfma_peak_tflops": PEAK_TFLOPS[form],
                "energy_Eh": m_sum,
                "abs_dE_Eh": abs(m_sum - e_sum),
                "rel_dE": abs(m_sum - e_sum) / abs(e_sum) if e_sum else None,
            }

    ctx.release(sys_)
    log("batched form")
    bshape = tuple(int(x) for x in args.batch_shape.split(",")) if args.batch_shape else None
    batched = (None if args.no_batched else
               bench_batched(ctx, args.batch_systems, 0, bshape))
    log("serving form")
    serving = None if args.no_serve else bench_serving(ctx, args.serve_requests, 20.0, 0)
    c3 = None
    if not args.no_c3:
        c3 = bench_c3(ctx, 60, 600, 1, args.c3_stride, args.window, args.c3_scale)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline")
        cpu = cpu_baseline(args.no, args.nv, args.seed, args.cpu_seconds)

    # E(T) of the first pass over the slices: the whole tuple space when the
    # pass started at slice 0 and completed, else the covered slices only
    first_pass = None
    if pass_e:
        covered = sorted({(args.step_offset + i) % len(slices) for i in range(len(pass_e))})
        first_pass = {"slices": [covered[0], covered[-1]] if covered else None,
                      "of": len(slices), "energy_Eh": neumaier(pass_e),
                      "whole_space": len(covered) == len(slices)}
    traffic_kernel = "doubles_f64" if args.precision == "fp64" else "doubles_split"
    traffic = measured_traffic(traffic_kernel, args.no, args.nv)
    traffic_src = ("recorded rocprofv3 FETCH_SIZE + WRITE_SIZE per launch of the "
                   f"{traffic_kernel} kernel at No={args.no} Nv={args.nv}, read from "
                   "profiles/ (not counted in this run)")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "split_fp32": "f32x2(split)+f64",
                  "fp32": "f32+f64", "bf16x3": "bf16x3+f32acc+f64"}[args.precision],
        "data": "synthetic (seeded counter-based generator, generated on device)",
        "config": {
            "workload": (f"plain (T) No={args.no} Nv={args.nv}, all tuples "
                         f"a<=b<=c ({n_tuples(args.nv)}), {args.precision}"),
            "No": args.no, "Nv": args.nv, "seed": args.seed,
            "precision": args.precision, "window": args.window,
            "batch": args.batch, "tile": args.tile,
            "parallelism": "single GPU",
        },
        "tflops": value * flop_tuple / 1e12,
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "traffic_source": (traffic_src if traffic is not None else None),
            "kernel": kname,
            "launches": ks.launches,
            "avg_launch_ms": ks.avg_ms,
        },
        "energy_kernel_ms": es.total_ms,
        "tuples_timed": done,
        "energy_Eh": (first_pass["energy_Eh"]
                      if first_pass and first_pass["whole_space"] else None),
        "energy_first_pass": first_pass,
        "setup_s": gen_s,
        "mixed_precision": mixed,
        "c3_stratified": c3,
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
