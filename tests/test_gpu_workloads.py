"""The benchmark workloads themselves, checked on the GPU at their stated
sizes:

* No=60, Nv=600 (36,179,600 tuples, 149 GB of FP64 operands resident): FP64
  per-tuple energies at the start, middle and end of the tuple space against
  the slice-wise oracle (slices drawn from the generator per tuple, so no
  125 GB host tensors), and every FP64-free operand store (split FP32,
  FP32, bf16x3) within the mixed tolerance of FP64 tuple by tuple over
  20,000 tuples; the amplitude scaling law at that shape.
* The 256 seeded systems of the batched workload (No ~ U{4..16},
  Nv ~ U{20..120}), tuned and run in one persistent launch: each E(T)
  within 1e-12 of the plain path, eight spread systems against the oracle.
* The first requests of the served Poisson stream (No log-U[8,40], Nv
  log-U[40,320], 20 req/s) under a 200 GB budget: each result bitwise the
  request served alone, within 1e-12 of the plain path.
"""
import numpy as np
import pytest

from atrip_amd import serve, workloads

pytestmark = pytest.mark.gpu

REL_TUPLE = 1e-12        # FP64, per tuple, relative to the largest |E_abc|
REL_E = 1e-10            # FP64, total
MIXED_REL_TUPLE = 1e-5   # mixed forms, per tuple, relative to the largest |E_abc|
MIXED_REL_E = 1e-7       # mixed forms, total, relative
CPU_THREADS = 16


def _close(a, b, rel):
    return abs(a - b) <= rel * abs(b) + 1e-18


def test_c3_no60_nv600(hip_ctx, oracle):
    no, nv, seed = 60, 600, 1
    n = oracle.n_tuples(nv)
    assert n == 36_179_600
    ranges = [(0, 48), (n // 2, 48), (n - 48, 48)]
    first, count = 12_000_000, 20_000
    print("C3: FP64 upload", flush=True)
    s = hip_ctx.upload_synthetic(no, nv, seed)
    try:
        sampled = [hip_ctx.tuple_energies(s, f, c) for f, c in ranges]
        print("C3: FP64 range", first, count, flush=True)
        base = hip_ctx.tuple_energies(s, first, count)
        e_base, _ = hip_ctx.run(s, first=first, count=count)
    finally:
        hip_ctx.release(s)
    assert _close(e_base, -np.sum(base), 1e-12)
    scale_base = np.max(np.abs(base))
    for store, precision in ((1, "split_fp32"), (2, "fp32"), (3, "bf16x3")):
        print("C3: operand_store", store, precision, flush=True)
        hip_ctx.set_option("operand_store", store)
        try:
            s = hip_ctx.upload_synthetic(no, nv, seed)
        finally:
            hip_ctx.set_option("operand_store", 0)
        try:
            pm = hip_ctx.tuple_energies(s, first, count, precision=precision)
            em, _ = hip_ctx.run(s, precision, first=first, count=count)
        finally:
            hip_ctx.release(s)
        assert np.max(np.abs(pm - base)) <= MIXED_REL_TUPLE * scale_base, precision
        assert _close(em, e_base, MIXED_REL_E), (precision, em, e_base)
    print("C3: slice-wise oracle", flush=True)
    oracle.blas_init()
    for (f, c), pt in zip(ranges, sampled):
        ref, _ = oracle.run_synthetic(no, nv, seed, f, c, nthreads=CPU_THREADS)
        assert np.max(np.abs(pt - ref)) <= REL_TUPLE * np.max(np.abs(ref)), (f, c)


def test_c3_scaled_to_molecular_magnitude(hip_ctx):
    """Amplitudes and integrals times s = 4e-3 put the full-space E(T) of
    No=60, Nv=600 near 0.5 Eh.  The energies scale by s^4 (bilinear doubles
    and singles terms in an energy quadratic in them), and bf16x3 stays
    within 1e-7 Eh absolute of FP64."""
    no, nv, seed, sc = 60, 600, 1, 4e-3
    first, count = 12_000_000, 20_000
    out = {}
    for store, precision, scale in ((0, "fp64", 1.0), (0, "fp64", sc), (3, "bf16x3", sc)):
        print("C3 scaled:", precision, scale, flush=True)
        hip_ctx.set_option("operand_store", store)
        try:
            s = hip_ctx.upload_synthetic_scaled(no, nv, seed, scale)
        finally:
            hip_ctx.set_option("operand_store", 0)
        try:
            out[(precision, scale)] = hip_ctx.run(s, precision, first=first, count=count)[0]
        finally:
            hip_ctx.release(s)
    e1, e64, etri = out[("fp64", 1.0)], out[("fp64", sc)], out[("bf16x3", sc)]
    assert _close(e64, e1 * sc ** 4, 1e-12), (e64, e1 * sc ** 4)
    assert abs(etri - e64) <= 1e-7, (etri, e64)


def test_c4_all_256_systems_tuned(hip_ctx, oracle):
    shapes = workloads.batch_shapes(256)
    print("C4: upload 256 systems", flush=True)
    systems = [hip_ctx.upload_synthetic(no, nv, seed) for no, nv, seed in shapes]
    try:
        print("C4: tune", flush=True)
        hip_ctx.batch_tune(systems, measure=True)
        print("C4: batched launch", flush=True)
        e = hip_ctx.batch_run(systems)
        print("C4: plain path per system", flush=True)
        for s, ei, sh in zip(systems, e, shapes):
            ep, _ = hip_ctx.run(s)
            assert _close(ei, ep, 1e-12), (sh, ei, ep)
    finally:
        for s in systems:
            hip_ctx.release(s)
    print("C4: oracle on spread systems", flush=True)
    oracle.blas_init()
    for q in (0, 37, 73, 110, 146, 183, 219, 255):
        no, nv, seed = shapes[q]
        t = oracle.synthetic(no, nv, seed, nthreads=CPU_THREADS)
        e_ref, _ = oracle.run(no, nv, t, use_gemm=True, nthreads=CPU_THREADS)
        assert _close(e[q], e_ref, REL_E), (shapes[q], e[q], e_ref)


@pytest.mark.parametrize("precision", ["fp64", "split_fp32"])
def test_c5_served_stream_head(hip_ctx, precision):
    """The first seven requests of the served workload, arriving by their
    Poisson times under a 200 GB budget: each result is bitwise the
    request served alone, and within 1e-12 (FP64) or the mixed tolerance
    of the plain path."""
    nreq = 7
    reqs = workloads.serve_requests(nreq, rate=20.0, seed=0)
    print("C5:", precision, reqs, flush=True)
    r = serve.run_poisson(hip_ctx, nreq, rate=20.0, seed=0, hbm_budget_gb=200.0,
                          iter_ms=100.0, precision=precision)
    assert r["failed"] == 0 and len(r["energies"]) == nreq
    for q, (_, no, nv, seed) in enumerate(reqs):
        with serve.Server(hip_ctx, hbm_budget_gb=200.0, precision=precision) as srv:
            rid = srv.submit_synthetic(no, nv, seed)
            alone = srv.drain([rid])[rid]
        assert alone["state"] == serve.DONE
        assert r["energies"][q] == alone["energy"], (q, no, nv)
        s = hip_ctx.upload_synthetic(no, nv, seed)
        try:
            ep, _ = hip_ctx.run(s, precision)
        finally:
            hip_ctx.release(s)
        assert _close(alone["energy"], ep, 1e-12 if precision == "fp64" else MIXED_REL_E)
