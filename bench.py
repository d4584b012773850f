#!/usr/bin/env python3
"""Benchmark of the (T) perturbative-triples hot path on one MI355X.

Workload: the plain FP64 (T) energy of a synthetic closed-shell system with
No=40 occupied and Nv=400 virtual orbitals (10,746,400 tuples a<=b<=c), all
tensors generated on the device from a seed and resident in HBM before the
timed region starts.  A step is one slice of the tuple space: the tuples whose
first index falls in a contiguous range of index blocks, with the slices
balanced so that 16 steps cover the whole space once.  The default run
(16 timed steps) therefore computes the full E(T) of the system.

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


def cpu_baseline(no: int, nv: int, seed: int, budget_s: float, cores: int):
    """Oracle (C port of the reference np=1 CPU path, host BLAS dgemm) on the
    first tuples of the same workload, all host cores as tuple shards."""
    from oracle import oracle as O  # the checker / CPU baseline only

    blas = O.blas_init()
    t0 = time.time()
    t = O.synthetic(no, nv, seed, nthreads=cores)
    gen_s = time.time() - t0
    flop_tuple = 12.0 * no ** 3 * (no + nv)
    # calibrate on a small sample, then size the measured sample to the budget
    probe = max(cores, 4)
    t0 = time.time()
    O.run(no, nv, t, first=0, count=probe, use_gemm=True, nthreads=cores)
    dt = max(time.time() - t0, 1e-3)
    count = int(max(probe, min(n_tuples(nv), probe * budget_s / dt)))
    count = (count // cores) * cores or cores
    t0 = time.time()
    O.run(no, nv, t, first=0, count=count, use_gemm=True, nthreads=cores)
    dt = time.time() - t0
    rate = count / dt
    return {
        "value": rate,
        "unit": "tuples/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"first {count} tuples (lexicographic order) of No={no} "
                   f"Nv={nv} seed={seed}; oracle port with dgemm from "
                   f"{blas or 'plain C'}; {cores} tuple-shard threads, "
                   f"single-threaded BLAS each; {dt:.1f} s measured, "
                   f"{gen_s:.1f} s host generation untimed"),
        "tflops": rate * flop_tuple / 1e12,
    }


def bench_batched(ctx, n_systems: int, seed: int):
    """Batched form: n seeded small systems, autotuned, one persistent launch."""
    from atrip_amd import workloads as W

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
        "workload": (f"{n_systems} systems, No~U{{4..16}}, Nv~U{{20..120}}, "
                     f"shape seed {seed}, tensor seeds 1000+s; FP64"),
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
                  precisions=("fp64", "split_fp32")):
    """The same Poisson request stream served once per precision; the mixed
    runs also report each request's relative energy difference to FP64."""
    from atrip_amd import serve

    out = {}
    ref = None
    for prec in precisions:
        log(f"serving form, {prec}")
        r = serve.run_poisson(ctx, n_requests, rate=rate, seed=seed,
                              hbm_budget_gb=200.0, iter_ms=100.0, precision=prec,
                              log=log)
        e = r.pop("energies")
        r["workload"] = (f"{n_requests} requests, Poisson {rate} req/s, No log-U[8,40], "
                         f"Nv log-U[40,320], seed {seed}, tensor seeds 5000+r; "
                         f"HBM budget 200 GB; {prec}")
        if ref is None:
            ref = e
        else:
            rel = [abs(e[q] - ref[q]) / abs(ref[q]) for q in e if q in ref and ref[q]]
            r["max_rel_dE_vs_fp64"] = max(rel) if rel else None
        out[prec] = r
    return out


def bench_c3(ctx, no: int, nv: int, seed: int, a0: int, a1: int, window: int):
    """Plain FP64 vs the mixed forms at No=60 Nv=600 on the tuples whose first
    index is in [a0, a1) (whole windows), each form on its own resident copy
    (FP64 operands; split-only; FP32-only; bf16x3-only)."""
    first = tuples_before(a0, nv)
    count = tuples_before(a1, nv) - first
    flop_tuple = 12.0 * no ** 3 * (no + nv)
    out = {"workload": (f"(T) No={no} Nv={nv} seed {seed}, tuples with first index "
                        f"in [{a0},{a1}) ({count})"), "tuples": count}
    ref = None
    for prec, store in (("fp64", 0), ("split_fp32", 1), ("fp32", 2), ("bf16x3", 3)):
        log(f"No={no} Nv={nv} sample, {prec}")
        ctx.set_option("operand_store", store)
        try:
            t0 = time.time()
            s_ = ctx.upload_synthetic(no, nv, seed)
            up = time.time() - t0
        finally:
            ctx.set_option("operand_store", 0)
        try:
            ctx.run(s_, prec, first=first, count=min(count, 1))  # warm
            kname = KERNEL_NAME[prec]
            ctx.reset_stats()
            ctx.set_option("timing", 1)
            t0 = time.perf_counter()
            e, _ = ctx.run(s_, prec, first=first, count=count)
            dt = time.perf_counter() - t0
            ctx.set_option("timing", 0)
            k = ctx.kernel_stats(kname)
            info = ctx.sys_info(s_)
        finally:
            ctx.release(s_)
        ent = {"value": count / dt, "unit": "tuples/s",
               "tflops": count * flop_tuple / dt / 1e12,
               "doubles_kernel_tflops": k.flop / (k.total_ms * 1e-3) / 1e12 if k.total_ms else None,
               "energy_Eh": e, "resident_gb": info["bytes"] / 1e9, "upload_s": up}
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
    ap.add_argument("--slices", type=int, default=16,
                    help="steps that cover the tuple space once")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision comparison pass")
    ap.add_argument("--mixed", default="split_fp32,bf16x3",
                    help="comma-separated forms of the comparison passes "
                         "(split_fp32, fp32, bf16x3)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-systems", type=int, default=256)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the No=60 Nv=600 FP64-vs-mixed sample")
    ap.add_argument("--c3-a0", type=int, default=400)
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
    sys_ = ctx.upload_synthetic(args.no, args.nv, args.seed)
    gen_s = time.time() - t0
    slices = step_slices(args.nv, args.window, args.slices)

    for i in range(args.warmup):
        first, count = slices[(len(slices) - 1 - i) % len(slices)]
        ctx.run(sys_, args.precision, first=first, count=count)

    ctx.reset_stats()
    ctx.set_option("timing", 1)
    done = 0
    e_sum = 0.0
    t_start = time.perf_counter()
    for i in range(args.steps):
        first, count = slices[i % len(slices)]
        e, _ = ctx.run(sys_, args.precision, first=first, count=count)
        e_sum += e
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
    if args.precision == "fp64" and not args.no_mixed:
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
                first, count = slices[i % len(slices)]
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
    batched = None if args.no_batched else bench_batched(ctx, args.batch_systems, 0)
    log("serving form")
    serving = None if args.no_serve else bench_serving(ctx, args.serve_requests, 20.0, 0)
    c3 = None
    if not args.no_c3:
        c3 = bench_c3(ctx, 60, 600, 1, args.c3_a0, args.c3_a0 + args.window,
                      args.window)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline")
        cores = min(len(os.sched_getaffinity(0)), 16)
        cpu = cpu_baseline(args.no, args.nv, args.seed, args.cpu_seconds, cores)

    full = args.steps % len(slices) == 0 and args.steps > 0
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
        "dtype": "f64" if args.precision == "fp64" else "f32x2(split)+f64",
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
            "traffic": measured_traffic(
                "doubles_f64" if args.precision == "fp64" else "doubles_split",
                args.no, args.nv),
            "kernel": kname,
            "launches": ks.launches,
            "avg_launch_ms": ks.avg_ms,
        },
        "energy_kernel_ms": es.total_ms,
        "tuples_timed": done,
        "energy_Eh": e_sum if full else None,
        "setup_s": gen_s,
        "mixed_precision": mixed,
        "mixed_c3": c3,
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
