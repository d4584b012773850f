"""Host ingestion on the GPU: the chunked device staging of host uploads, the
slab-wise streaming upload through a reader, complex orbital energies with
imaginary parts, and a context that keeps working after an upload ran out of
device memory.

Bar: a system uploaded from the host in many chunks (or through a reader in
many slabs) is bit for bit the system generated on the device from the same
seed, so its per-tuple energies are bitwise equal in every precision its
operand store allows; complex systems match the complex oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STORE_PRECISIONS = {0: ("fp64", "split_fp32", "fp32", "bf16x3"),
                    1: ("split_fp32", "fp32"), 2: ("fp32",), 3: ("bf16x3",)}


def _energies(ctx, s, nv, oracle, precisions):
    n = oracle.n_tuples(nv)
    return {p: ctx.tuple_energies(s, 0, n, precision=p, with_ct=True) for p in precisions}


def _same(a, b):
    for p in a:
        assert np.array_equal(a[p][0], b[p][0]), p
        assert np.array_equal(a[p][1], b[p][1]), p


SMALL = ("ei", "ea", "Tai")  # read whole, as the reference's read_all


class _Reader:
    """Reads slabs out of whole host arrays and records the largest read of
    a large tensor."""

    def __init__(self, t):
        self.t = t
        self.max_count = 0
        self.tensors = set()

    def __call__(self, name, first, count):
        if name not in SMALL:
            self.max_count = max(self.max_count, count)
        self.tensors.add(name)
        return self.t[name][first:first + count]


@pytest.mark.parametrize("store", [0, 1, 2, 3])
def test_chunked_host_upload_is_the_device_system(hip_ctx, oracle, store):
    """Every large tensor staged in many chunks (one slab per chunk, then
    several slabs per chunk with a short last chunk): per-tuple energies
    bitwise equal to the device-generated system, in each precision the
    operand store keeps, (cT) included."""
    no, nv, seed = 7, 19, 11
    t = oracle.synthetic(no, nv, seed, with_j=True)
    precs = STORE_PRECISIONS[store]
    hip_ctx.set_option("operand_store", store)
    try:
        ref = hip_ctx.upload_synthetic(no, nv, seed, with_j=True)
        want = _energies(hip_ctx, ref, nv, oracle, precs)
        hip_ctx.release(ref)
        # 1 KiB: one slab per chunk; 48 KiB: Tabij 2 slabs (of 2527 values)
        # per chunk with a short last chunk, Vijka 17 slabs then 2
        for stage_kb in (1, 48):
            print("store", store, "stage_kb", stage_kb, flush=True)
            hip_ctx.set_option("stage_kb", stage_kb)
            s = hip_ctx.upload(no, nv, t)
            try:
                _same(_energies(hip_ctx, s, nv, oracle, precs), want)
            finally:
                hip_ctx.release(s)
            rd = _Reader(t)
            s = hip_ctx.upload_stream(no, nv, rd, with_j=True)
            try:
                _same(_energies(hip_ctx, s, nv, oracle, precs), want)
            finally:
                hip_ctx.release(s)
            # pinned buffer of stage_kb bytes in two halves
            assert rd.max_count <= stage_kb * 1024 // 16
            assert rd.tensors == set(t)
    finally:
        hip_ctx.set_option("stage_kb", 4 << 20)
        hip_ctx.set_option("operand_store", 0)


def test_chunked_complex_upload(hip_ctx, oracle):
    """Complex systems staged in many chunks, from interleaved host arrays
    and through a reader, equal the device-generated complex system bit for
    bit, and match the complex oracle."""
    no, nv, seed = 5, 11, 64
    t = oracle.synthetic_z(no, nv, seed, with_j=True)
    e_ref, ect_ref = oracle.run_z(no, nv, t)
    ref = hip_ctx.upload_synthetic_complex(no, nv, seed, with_j=True)
    n = oracle.n_tuples(nv)
    try:
        want = hip_ctx.tuple_energies(ref, 0, n, with_ct=True)
        e_dev = hip_ctx.run(ref)
    finally:
        hip_ctx.release(ref)
    assert abs(e_dev[0] - e_ref) <= 1e-10 * abs(e_ref)
    assert abs(e_dev[1] - ect_ref) <= 1e-10 * abs(ect_ref)
    try:
        hip_ctx.set_option("stage_kb", 1)
        s = hip_ctx.upload_complex(no, nv, t)
        try:
            got = hip_ctx.tuple_energies(s, 0, n, with_ct=True)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        finally:
            hip_ctx.release(s)
        rd = _Reader(t)
        s = hip_ctx.upload_stream(no, nv, rd, with_j=True, complex_=True)
        try:
            got = hip_ctx.tuple_energies(s, 0, n, with_ct=True)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
            assert hip_ctx.run(s) == e_dev
        finally:
            hip_ctx.release(s)
        assert rd.max_count <= 1024 // 32  # complex values: 16 bytes each
    finally:
        hip_ctx.set_option("stage_kb", 4 << 20)


def test_complex_orbital_energies_with_imaginary_parts(hip_ctx, oracle):
    """Host-uploaded complex system whose ea and ei carry imaginary parts:
    the denominator uses Re(ea_a + ea_b + ea_c) and the complex ei, as the
    reference (src/atrip/Atrip.cxx:643-644), per tuple and in total."""
    no, nv = 6, 10
    t = oracle.synthetic_z(no, nv, 91, with_j=True)
    rng = np.random.default_rng(91)
    t["ea"] = t["ea"] + 1j * rng.uniform(0.05, 0.3, nv)
    t["ei"] = t["ei"] + 1j * rng.uniform(-0.2, 0.2, no)
    e_ref, ect_ref, pt_ref, ptc_ref = oracle.run_z(no, nv, t, per_tuple=True)
    s = hip_ctx.upload_complex(no, nv, t)
    try:
        e, ect = hip_ctx.run(s)
        pt, ptc = hip_ctx.tuple_energies(s, 0, oracle.n_tuples(nv), with_ct=True)
    finally:
        hip_ctx.release(s)
    assert np.max(np.abs(pt - pt_ref)) <= 1e-12 * np.max(np.abs(pt_ref))
    assert np.max(np.abs(ptc - ptc_ref)) <= 1e-12 * np.max(np.abs(ptc_ref))
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)
    assert abs(ect - ect_ref) <= 1e-10 * abs(ect_ref)


def test_reader_failure_is_an_error(hip_ctx, oracle):
    """A reader that fails aborts the upload with an error; the context
    stays usable."""
    from atrip_amd import hip

    no, nv = 4, 9
    t = oracle.synthetic(no, nv, 3)

    def bad(name, first, count):
        if name == "Vabci":
            raise IOError("disk gone")
        return t[name][first:first + count]

    with pytest.raises(hip.HipError, match="reader failed"):
        hip_ctx.upload_stream(no, nv, bad)
    s = hip_ctx.upload_stream(no, nv, _Reader(t))
    try:
        e, _ = hip_ctx.run(s)
    finally:
        hip_ctx.release(s)
    e_ref, _ = oracle.run(no, nv, t)
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)


def test_context_works_after_out_of_memory(hip_ctx, oracle):
    """An upload larger than the device fails with out of memory; on the
    same context the plain and batched runs of a small system then succeed
    with the energies they had before (no stale error reported by a
    launch)."""
    from atrip_amd import hip

    no, nv, seed = 6, 14, 17
    s = hip_ctx.upload_synthetic(no, nv, seed)
    try:
        before = hip_ctx.run(s)
        before_b = hip_ctx.batch_run([s])
        # No=100, Nv=1000: the V(xy) slabs alone are 0.8 TB
        with pytest.raises(hip.HipError, match="out of memory"):
            hip_ctx.upload_synthetic(100, 1000, 1)
        assert hip_ctx.run(s) == before
        assert np.array_equal(hip_ctx.batch_run([s]), before_b)
        with pytest.raises(hip.HipError, match="out of memory"):
            hip_ctx.upload_synthetic(100, 1000, 1)
        assert np.array_equal(hip_ctx.batch_run([s]), before_b)
        assert hip_ctx.run(s) == before
    finally:
        hip_ctx.release(s)
