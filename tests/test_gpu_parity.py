"""Parity of the HIP path against the CPU oracle, through the C-ABI.

Bar: per-tuple energies within 1e-12 of the largest per-tuple magnitude and
total E(T) within 1e-10 relative (the plain FP64 tolerance), on the same
seeded inputs; bitwise identity where the arithmetic is the same (device vs
host generation, window/batch partitioning).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

REL_E = 1e-10       # total E(T), relative
REL_TUPLE = 1e-12   # per tuple, relative to max |E_abc|
FLOOR = 1e-18       # absolute, Eh


def _check(ctx, oracle, no, nv, seed, window=8, batch=4, with_j=False):
    t = oracle.synthetic(no, nv, seed, with_j=with_j)
    e_ref, ect_ref, pt_ref, ptc_ref = oracle.run(no, nv, t, per_tuple=True)
    ctx.set_option("window", window)
    ctx.set_option("batch", batch)
    s = ctx.upload(no, nv, t)
    try:
        e, ect = ctx.run(s)
        n = oracle.n_tuples(nv)
        pt, ptc = ctx.tuple_energies(s, 0, n, with_ct=True)
    finally:
        ctx.release(s)
    # No = 1 has E = 0 in exact arithmetic: FLOOR absorbs pure roundoff
    scale = np.max(np.abs(pt_ref))
    assert np.max(np.abs(pt - pt_ref)) <= REL_TUPLE * scale + FLOOR
    assert np.max(np.abs(ptc - ptc_ref)) <= REL_TUPLE * np.max(np.abs(ptc_ref)) + FLOOR
    assert abs(e - e_ref) <= REL_E * abs(e_ref) + FLOOR
    assert abs(ect - ect_ref) <= REL_E * abs(ect_ref) + FLOOR
    return e, pt


@pytest.mark.parametrize("no,nv,window,batch", [
    (1, 3, 2, 1), (2, 2, 8, 4), (3, 4, 2, 3), (4, 7, 3, 2), (5, 9, 8, 1),
    (7, 13, 4, 5), (16, 17, 8, 2), (17, 10, 3, 4), (9, 33, 16, 1), (20, 24, 8, 4),
])
def test_small_systems(hip_ctx, oracle, no, nv, window, batch):
    _check(hip_ctx, oracle, no, nv, seed=100 + no * nv, window=window, batch=batch)


def test_c1_full_golden(hip_ctx, oracle):
    """No=10, Nv=50, all 22,050 tuples: the CPU-runnable configuration."""
    gold = json.load(open(os.path.join(GOLDEN, "c1_seed1.json")))
    e, pt = _check(hip_ctx, oracle, 10, 50, seed=1)
    assert abs(e - gold["energy"]) <= REL_E * abs(gold["energy"])
    for idx, val in gold["tuple_energies"].items():
        assert abs(pt[int(idx)] - val) <= REL_TUPLE * gold["max_abs_tuple_energy"]


def test_cT_pass(hip_ctx, oracle):
    _check(hip_ctx, oracle, 5, 11, seed=21, window=4, with_j=True)


def test_ijkabc(hip_ctx, oracle):
    no, nv = 6, 12
    t = oracle.synthetic(no, nv, 31)
    e_ref, _ = oracle.run(no, nv, t, ijkabc=True)
    s = hip_ctx.upload(no, nv, t)
    try:
        e, ect = hip_ctx.run(s, ijkabc=True)
    finally:
        hip_ctx.release(s)
    assert abs(e - e_ref) <= REL_E * abs(e_ref) and ect == e


@pytest.mark.parametrize("first,count", [(0, 1), (0, 100), (37, 411), (700, 10 ** 9),
                                         (1000, 0)])
def test_ranges(hip_ctx, oracle, first, count):
    no, nv = 6, 20
    t = oracle.synthetic(no, nv, 41)
    n = oracle.n_tuples(nv)
    cnt = max(0, min(count, n - first))
    s = hip_ctx.upload(no, nv, t)
    try:
        hip_ctx.set_option("window", 4)
        e, _ = hip_ctx.run(s, first=first, count=count)
        pt = hip_ctx.tuple_energies(s, first, cnt) if cnt else np.zeros(0)
    finally:
        hip_ctx.release(s)
    if cnt == 0:
        assert e == 0.0
        return
    e_ref, _, pt_ref, _ = oracle.run(no, nv, t, first=first, count=cnt, per_tuple=True)
    assert abs(e - e_ref) <= REL_E * abs(e_ref)
    assert np.max(np.abs(pt - pt_ref)) <= REL_TUPLE * np.max(np.abs(pt_ref))


def test_device_generation_is_bit_identical(hip_ctx, oracle):
    no, nv, seed = 11, 23, 5
    t = oracle.synthetic(no, nv, seed, with_j=True)
    s1 = hip_ctx.upload(no, nv, t)
    s2 = hip_ctx.upload_synthetic(no, nv, seed, with_j=True)
    try:
        n = oracle.n_tuples(nv)
        a, ac = hip_ctx.tuple_energies(s1, 0, n, with_ct=True)
        b, bc = hip_ctx.tuple_energies(s2, 0, n, with_ct=True)
    finally:
        hip_ctx.release(s1)
        hip_ctx.release(s2)
    assert np.array_equal(a, b) and np.array_equal(ac, bc)


def test_partition_independence(hip_ctx, oracle):
    """Per-tuple energies do not depend on which tuples share a launch."""
    no, nv = 12, 40
    s = hip_ctx.upload_synthetic(no, nv, 77)
    try:
        n = oracle.n_tuples(nv)
        outs = []
        for window, batch, tile in [(8, 4, 0), (4, 1, 0), (16, 2, 1), (5, 7, 2), (8, 2, 5), (0, 4, -1),
                                    (3, 5, 3), (8, 1, -1)]:
            hip_ctx.set_option("window", window)
            hip_ctx.set_option("batch", batch)
            hip_ctx.set_option("tile", tile)
            outs.append(hip_ctx.tuple_energies(s, 0, n))
        hip_ctx.set_option("tile", -1)
        hip_ctx.set_option("window", 0)
        hip_ctx.set_option("batch", 4)
        # a single tuple alone gives the same bits as in the full run
        alone = hip_ctx.tuple_energies(s, 1234, 1)
    finally:
        hip_ctx.release(s)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    assert alone[0] == outs[0][1234]


def test_mid_size_sample(hip_ctx, oracle):
    """No=24, Nv=120 (Vabci 330 MB): sampled tuple ranges against the oracle."""
    no, nv, seed = 24, 120, 3
    t = oracle.synthetic(no, nv, seed, nthreads=8)
    oracle.blas_init()
    s = hip_ctx.upload_synthetic(no, nv, seed)
    try:
        for first, count in [(0, 300), (150000, 200), (oracle.n_tuples(nv) - 250, 250)]:
            pt = hip_ctx.tuple_energies(s, first, count)
            _, _, pt_ref, _ = oracle.run(no, nv, t, first=first, count=count,
                                         use_gemm=True, nthreads=8, per_tuple=True)
            assert np.max(np.abs(pt - pt_ref)) <= REL_TUPLE * np.max(np.abs(pt_ref))
    finally:
        hip_ctx.release(s)


def test_benchmark_size_sample_and_properties(hip_ctx, oracle):
    """No=40, Nv=400, seed 1: the size and inputs of the FP64 benchmark
    (Vabci 20.5 GB).  Sampled tuple ranges at the start, middle and end of
    the tuple space against the oracle; over a larger range, the
    size-independent properties: the per-tuple energies do not depend on the
    window edge, batch or tile (bitwise), and every mixed form stays within
    the mixed tolerance of FP64 tuple by tuple."""
    no, nv, seed = 40, 400, 1
    n = oracle.n_tuples(nv)
    ranges = [(0, 96), (n // 2, 96), (n - 96, 96)]
    first, count = 5_000_000, 20_000
    print("benchmark size: upload", flush=True)
    s = hip_ctx.upload_synthetic(no, nv, seed)
    try:
        sampled = [hip_ctx.tuple_energies(s, f, c) for f, c in ranges]
        base = hip_ctx.tuple_energies(s, first, count)
        for window, batch, tile in [(4, 1, 0), (8, 2, 3)]:
            print("benchmark size: window/batch/tile", window, batch, tile, flush=True)
            hip_ctx.set_option("window", window)
            hip_ctx.set_option("batch", batch)
            hip_ctx.set_option("tile", tile)
            assert np.array_equal(hip_ctx.tuple_energies(s, first, count), base)
        hip_ctx.set_option("tile", -1)
        hip_ctx.set_option("window", 0)
        hip_ctx.set_option("batch", 4)
        for precision in MIXED_FORMS:
            print("benchmark size:", precision, flush=True)
            pm = hip_ctx.tuple_energies(s, first, count, precision=precision)
            assert np.max(np.abs(pm - base)) <= MIXED_REL_TUPLE * np.max(np.abs(base))
    finally:
        hip_ctx.release(s)
    print("benchmark size: oracle sample", flush=True)
    t = oracle.synthetic(no, nv, seed, nthreads=16)
    oracle.blas_init()
    for (f, c), pt in zip(ranges, sampled):
        _, _, pt_ref, _ = oracle.run(no, nv, t, first=f, count=c, use_gemm=True,
                                     nthreads=16, per_tuple=True)
        assert np.max(np.abs(pt - pt_ref)) <= REL_TUPLE * np.max(np.abs(pt_ref)), (f, c)
    del t


def test_empty_tuple_space(hip_ctx):
    """Nv = 1 holds no a <= b <= c other than a = b = c: E(T) = E(cT) = 0
    in every form, and an empty tuple range returns no energies."""
    s = hip_ctx.upload_synthetic(3, 1, 9, with_j=True)
    try:
        for precision in ["fp64"] + MIXED_FORMS:
            assert hip_ctx.run(s, precision) == (0.0, 0.0), precision
        assert hip_ctx.tuple_energies(s, 0, 0).size == 0
    finally:
        hip_ctx.release(s)


def test_errors_are_reported(hip_ctx):
    from atrip_amd import hip

    with pytest.raises(hip.HipError, match="unknown system"):
        hip_ctx.run(987654)
    with pytest.raises(hip.HipError, match="window"):
        hip_ctx.set_option("window", -1)
    with pytest.raises(hip.HipError, match="unknown option"):
        hip_ctx.set_option("no-such-knob", 1)


# ---- mixed precision (split FP32 doubles, FP64 energy) ----------------------

MIXED_REL_TUPLE = 1e-5   # per tuple, relative to max |E_abc| (FP32 accumulation)
MIXED_REL_E = 1e-7       # total E(T), relative


MIXED_FORMS = ["split_fp32", "bf16x3"]
_ORACLE_RUNS = {}


def _oracle_per_tuple(oracle, no, nv, seed):
    """Synthetic tensors and the oracle's per-tuple run, once per shape."""
    key = (no, nv, seed)
    if key not in _ORACLE_RUNS:
        t = oracle.synthetic(no, nv, seed)
        _ORACLE_RUNS[key] = (t, oracle.run(no, nv, t, per_tuple=True))
    return _ORACLE_RUNS[key]


@pytest.mark.parametrize("precision", MIXED_FORMS)
def test_mixed_precision_against_fp64(hip_ctx, oracle, precision):
    for no, nv, seed in [(4, 7, 1), (10, 50, 1), (17, 23, 2), (32, 20, 3)]:
        print(precision, no, nv, seed, flush=True)
        t, (e_ref, _, pt_ref, _) = _oracle_per_tuple(oracle, no, nv, seed)
        s = hip_ctx.upload(no, nv, t)
        try:
            e, ect = hip_ctx.run(s, precision)
            n = oracle.n_tuples(nv)
            pt = hip_ctx.tuple_energies(s, 0, n, precision=precision)
        finally:
            hip_ctx.release(s)
        assert abs(e - e_ref) <= MIXED_REL_E * abs(e_ref)
        assert ect == e
        assert np.max(np.abs(pt - pt_ref)) <= MIXED_REL_TUPLE * np.max(np.abs(pt_ref))


@pytest.mark.parametrize("precision", MIXED_FORMS)
def test_mixed_precision_cT_and_partition(hip_ctx, oracle, precision):
    no, nv = 9, 21
    t = oracle.synthetic(no, nv, 5, with_j=True)
    e_ref, ect_ref = oracle.run(no, nv, t)
    s = hip_ctx.upload(no, nv, t)
    try:
        e, ect = hip_ctx.run(s, precision)
        n = oracle.n_tuples(nv)
        a = hip_ctx.tuple_energies(s, 0, n, precision=precision)
        hip_ctx.set_option("window", 4)
        hip_ctx.set_option("batch", 1)
        b = hip_ctx.tuple_energies(s, 0, n, precision=precision)
        hip_ctx.set_option("window", 0)
        hip_ctx.set_option("batch", 4)
    finally:
        hip_ctx.release(s)
    assert abs(e - e_ref) <= MIXED_REL_E * abs(e_ref)
    assert abs(ect - ect_ref) <= MIXED_REL_E * abs(ect_ref)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("no,nv,seed", [(1, 5, 31), (2, 3, 32), (3, 2, 33), (33, 12, 34),
                                        (65, 4, 35)])
def test_mixed_edge_shapes(hip_ctx, oracle, no, nv, seed):
    """Tiny and odd shapes through every mixed form: No = 1..3 (one padded
    fragment), No = 33 (No^2 = 1089 rows, padded slices), No = 65 (> 64)."""
    t, (e_ref, _, pt_ref, _) = _oracle_per_tuple(oracle, no, nv, seed)
    s = hip_ctx.upload(no, nv, t)
    try:
        for precision in ("split_fp32", "fp32", "bf16x3"):
            print(precision, no, nv, flush=True)
            e, _ = hip_ctx.run(s, precision)
            assert abs(e - e_ref) <= MIXED_REL_E * abs(e_ref) + 1e-15, (precision, e, e_ref)
            pt = hip_ctx.tuple_energies(s, 0, oracle.n_tuples(nv), precision=precision)
            assert np.max(np.abs(pt - pt_ref)) <= MIXED_REL_TUPLE * np.max(np.abs(pt_ref)) + 1e-18
    finally:
        hip_ctx.release(s)


def test_mixed_absolute_error_at_molecular_scale(hip_ctx, oracle):
    """The mixed forms' error is relative to |E(T)|: with the synthetic
    amplitudes and integrals scaled down so that |E(T)| is of molecular size
    (0.01-0.1 Eh), every mixed form is within 1e-7 Eh of FP64 (and FP64 of
    the oracle).  The unscaled synthetic inputs give |E(T)| ~ 1e7-1e8 Eh,
    where no FP64 or mixed result can be 1e-7 Eh absolute."""
    no, nv, seed = 12, 36, 21
    t = oracle.synthetic(no, nv, seed)
    # T2 and V scaled by 0.1: E(T) = -0.092 Eh here
    for name in ("Tabij", "Vabij", "Vijka", "Vabci"):
        t[name] = t[name] * 0.1
    e_ref, _ = oracle.run(no, nv, t)
    print("molecular-scale E(T) oracle", e_ref, flush=True)
    assert 1e-3 < abs(e_ref) < 1.0
    s = hip_ctx.upload(no, nv, t)
    try:
        e64, _ = hip_ctx.run(s)
        assert abs(e64 - e_ref) <= 1e-10 * abs(e_ref)
        for precision in ("split_fp32", "fp32", "bf16x3"):
            e, _ = hip_ctx.run(s, precision)
            print(precision, "abs dE", abs(e - e64), flush=True)
            assert abs(e - e64) <= 1e-7, (precision, e, e64)
    finally:
        hip_ctx.release(s)


@pytest.mark.parametrize("precision", MIXED_FORMS)
def test_mixed_tiles_and_windows(hip_ctx, oracle, precision):
    """Every tile variant and window shape gives the same per-tuple bits
    (each W element accumulates the same chunks in the same order), within
    the mixed tolerance of the oracle."""
    no, nv, seed = 12, 40, 78
    t, (e_ref, _, pt_ref, _) = _oracle_per_tuple(oracle, no, nv, seed)
    s = hip_ctx.upload(no, nv, t)
    n = oracle.n_tuples(nv)
    outs = []
    try:
        for window, batch, tile in [(8, 4, 0), (4, 1, 1), (16, 2, 2), (5, 7, 3),
                                    (8, 2, 4), (3, 5, 5), (8, 1, 6), (6, 3, 7),
                                    (8, 4, -1)]:
            print(precision, "window", window, "batch", batch, "tile", tile, flush=True)
            hip_ctx.set_option("window", window)
            hip_ctx.set_option("batch", batch)
            hip_ctx.set_option("tile", tile)
            outs.append(hip_ctx.tuple_energies(s, 0, n, precision=precision))
    finally:
        hip_ctx.set_option("tile", -1)
        hip_ctx.set_option("window", 0)
        hip_ctx.set_option("batch", 4)
        hip_ctx.release(s)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    assert np.max(np.abs(outs[0] - pt_ref)) <= MIXED_REL_TUPLE * np.max(np.abs(pt_ref))


def test_operand_store_modes(hip_ctx, oracle):
    """Systems kept without FP64 operands (split, FP32 or bf16x3 only) give
    the same bits as the copies an FP64-resident system makes; runs needing
    operands the system does not hold are refused."""
    from atrip_amd import hip

    no, nv, seed = 7, 19, 11
    t = oracle.synthetic(no, nv, seed, with_j=True)
    ref = hip_ctx.upload_synthetic(no, nv, seed, with_j=True)
    try:
        want = {p: hip_ctx.run(ref, p) for p in ("split_fp32", "fp32", "bf16x3")}
        for store, precs in ((1, ("split_fp32", "fp32")), (2, ("fp32",)),
                             (3, ("bf16x3",))):
            hip_ctx.set_option("operand_store", store)
            try:
                sys_dev = hip_ctx.upload_synthetic(no, nv, seed, with_j=True)
                sys_host = hip_ctx.upload(no, nv, t)
            finally:
                hip_ctx.set_option("operand_store", 0)
            try:
                for p in precs:
                    assert hip_ctx.run(sys_dev, p) == want[p], (store, p)
                    assert hip_ctx.run(sys_host, p) == want[p], (store, p)
                with pytest.raises(hip.HipError, match="FP64 operands"):
                    hip_ctx.run(sys_dev, "fp64")
                if store >= 2:
                    with pytest.raises(hip.HipError, match="without split operands"):
                        hip_ctx.run(sys_dev, "split_fp32")
                if store in (1, 2):
                    with pytest.raises(hip.HipError, match="without bf16x3 operands"):
                        hip_ctx.run(sys_dev, "bf16x3")
                if store == 3:
                    with pytest.raises(hip.HipError, match="bf16x3 operands only"):
                        hip_ctx.run(sys_dev, "fp32")
                assert hip_ctx.sys_info(sys_dev)["bytes"] < hip_ctx.sys_info(ref)["bytes"]
            finally:
                hip_ctx.release(sys_dev)
                hip_ctx.release(sys_host)
    finally:
        hip_ctx.release(ref)


@pytest.mark.parametrize("no,nv,seed,with_j", [
    (1, 3, 61, False), (3, 5, 62, True), (4, 9, 63, False), (9, 14, 64, True),
    (17, 10, 65, False),
])
def test_complex_against_oracle(hip_ctx, oracle, no, nv, seed, with_j):
    """Complex (T): device-generated and host-uploaded complex systems
    against the complex oracle (E, E(cT), per tuple)."""
    print("complex", no, nv, seed, with_j, flush=True)
    t = oracle.synthetic_z(no, nv, seed, with_j=with_j)
    e_ref, ect_ref, pt_ref, ptc_ref = oracle.run_z(no, nv, t, per_tuple=True)
    s_host = hip_ctx.upload_complex(no, nv, t)
    s_dev = hip_ctx.upload_synthetic_complex(no, nv, seed, with_j=with_j)
    try:
        e, ect = hip_ctx.run(s_host)
        n = oracle.n_tuples(nv)
        pt, ptc = hip_ctx.tuple_energies(s_host, 0, n, with_ct=True)
        assert hip_ctx.run(s_dev) == (e, ect)  # same inputs bit for bit
        e_ijk, _ = hip_ctx.run(s_dev, ijkabc=True)
        with pytest.raises(Exception, match="FP64 only"):
            hip_ctx.run(s_dev, "split_fp32")
    finally:
        hip_ctx.release(s_host)
        hip_ctx.release(s_dev)
    scale = np.max(np.abs(pt_ref))
    assert np.max(np.abs(pt - pt_ref)) <= REL_TUPLE * scale + FLOOR
    assert np.max(np.abs(ptc - ptc_ref)) <= REL_TUPLE * np.max(np.abs(ptc_ref)) + FLOOR
    assert abs(e - e_ref) <= REL_E * abs(e_ref) + FLOOR
    assert abs(ect - ect_ref) <= REL_E * abs(ect_ref) + FLOOR
    e_ijk_ref, _ = oracle.run_z(no, nv, t, ijkabc=True)
    assert abs(e_ijk - e_ijk_ref) <= REL_E * abs(e_ijk_ref) + FLOOR


def test_complex_with_zero_imaginary_is_real(hip_ctx, oracle):
    no, nv = 6, 11
    t = oracle.synthetic(no, nv, 70, with_j=True)
    s_r = hip_ctx.upload(no, nv, t)
    s_z = hip_ctx.upload_complex(no, nv, {k: v.astype(np.complex128) for k, v in t.items()})
    try:
        er, ectr = hip_ctx.run(s_r)
        ez, ectz = hip_ctx.run(s_z)
    finally:
        hip_ctx.release(s_r)
        hip_ctx.release(s_z)
    assert abs(ez - er) <= 1e-13 * abs(er) and abs(ectz - ectr) <= 1e-13 * abs(ectr)



def test_fused_and_cube_epilogues_agree(hip_ctx):
    """The fused epilogue (W gathered per block triple from the class
    buffers, energies summed per triple) and the cube epilogue (W written
    as cubes first, one workgroup per tuple) form the same W bits and
    differ only in the order the points' terms are added: per-tuple
    energies within 1e-14 of the largest, E(T) and E(cT) within 1e-14
    relative, for FP64, bf16x3 and split FP32 with (cT)."""
    hip_ctx.set_option("window", 4)
    hip_ctx.set_option("batch", 3)
    try:
        s_ = hip_ctx.upload_synthetic(11, 37, 21, with_j=True)
        n = 37 * 38 * 39 // 6 - 37
        for prec in ("fp64", "bf16x3", "split_fp32"):
            print("epilogues", prec, flush=True)
            out = {}
            for epi in (0, 1):
                hip_ctx.set_option("epilogue", epi)
                out[epi] = (hip_ctx.run(s_, prec),
                            hip_ctx.tuple_energies(s_, 0, n, precision=prec, with_ct=True))
            (e0, c0), (t0, u0) = out[0]
            (e1, c1), (t1, u1) = out[1]
            assert abs(e0 - e1) <= 1e-14 * abs(e1) and abs(c0 - c1) <= 1e-14 * abs(c1)
            assert np.max(np.abs(t0 - t1)) <= 1e-14 * np.max(np.abs(t1))
            assert np.max(np.abs(u0 - u1)) <= 1e-14 * np.max(np.abs(u1))
        hip_ctx.release(s_)
    finally:
        hip_ctx.set_option("epilogue", 0)
        hip_ctx.set_option("window", 0)
        hip_ctx.set_option("batch", 4)
