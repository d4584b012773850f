"""The allocation-free entry point atrip_hip_run_noalloc against
atrip_hip_run, through the C-ABI.

Bar: with the same window edge and windows per launch the two entry points
compute the same launches, so E(T) and E(cT) are bitwise equal; two calls in
a row (the second on a smaller system, over the first call's leftovers in
the shared region) both match.  Sizes beyond the region and mixed forms
without resident operand copies are refused before any launch.
"""
import numpy as np
import pytest

from atrip_amd import hip

pytestmark = pytest.mark.gpu


@pytest.fixture()
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


def _both(c, s, precision="fp64", **kw):
    return c.run(s, precision, **kw), c.run_noalloc(s, precision, **kw)


def test_two_calls_in_a_row_second_smaller(ctx):
    ctx.set_option("window", 4)
    ctx.set_option("batch", 4)
    big = ctx.upload_synthetic(12, 40, 3, with_j=True)
    small = ctx.upload_synthetic(6, 20, 4)
    for s, label in ((big, "No=12 Nv=40 (cT)"), (small, "No=6 Nv=20")):
        print("noalloc:", label, flush=True)
        (e, ect), (e2, ect2) = _both(ctx, s)
        assert e2 == e and ect2 == ect, (label, e, e2, ect, ect2)
        assert e != 0.0
    # and back to the larger one, over the smaller one's leftovers
    (e, ect), (e2, ect2) = _both(ctx, big)
    assert e2 == e and ect2 == ect
    ctx.release(big)
    ctx.release(small)


def test_ranges_ijkabc_and_complex(ctx):
    ctx.set_option("window", 4)
    ctx.set_option("batch", 2)
    s = ctx.upload_synthetic(9, 33, 5)
    n = 33 * 34 * 35 // 6 - 33
    for first, count in ((0, 1), (37, 411), (n - 100, None), (n, 5)):
        print("noalloc: range", first, count, flush=True)
        a, b = _both(ctx, s, first=first, count=count)
        assert a == b, (first, count, a, b)
    a, b = _both(ctx, s, ijkabc=True)
    assert a == b
    ctx.release(s)
    z = ctx.upload_synthetic_complex(5, 14, 6, with_j=True)
    print("noalloc: complex", flush=True)
    a, b = _both(ctx, z)
    assert a == b
    ctx.release(z)


def test_mixed_forms_need_resident_copies(ctx):
    ctx.set_option("window", 4)
    s = ctx.upload_synthetic(8, 30, 7)
    for prec in ("bf16x3", "split_fp32"):
        with pytest.raises(hip.HipError, match="not resident"):
            ctx.run_noalloc(s, prec)
        print("noalloc:", prec, flush=True)
        a, b = _both(ctx, s, prec)  # the plain run makes the copies first
        assert a == b, (prec, a, b)
    ctx.release(s)
    ctx.set_option("operand_store", 3)
    try:
        t = ctx.upload_synthetic(8, 30, 7)
    finally:
        ctx.set_option("operand_store", 0)
    a, b = _both(ctx, t, "bf16x3")
    assert a == b
    ctx.release(t)


def test_window_beyond_the_region_is_refused(ctx):
    # one window of edge 8 at No=40 needs 512 slots x 3 x 40^3 doubles of
    # class buffers alone (786 MB), beyond the 256 MiB region
    ctx.set_option("window", 8)
    ctx.set_option("batch", 4)
    s = ctx.upload_synthetic(40, 12, 8)
    with pytest.raises(hip.HipError, match="window edge 5 fits"):
        ctx.run_noalloc(s)
    # edge 5 fits one window per launch; with one window per launch in the
    # plain run too, both compute the same launches
    ctx.set_option("window", 5)
    ctx.set_option("batch", 1)
    print("noalloc: No=40 window 5", flush=True)
    a, b = _both(ctx, s)
    assert a == b
    # with four windows per launch in the plain run the sums group
    # differently: equal to rounding
    ctx.set_option("batch", 4)
    (e, _), (e2, _) = _both(ctx, s)
    assert abs(e2 - e) <= 1e-12 * abs(e)
    ctx.release(s)
