"""Batched form: the full tuple spaces of many resident systems in one
persistent launch, through the C-ABI.

Bar: each system's E(T) within 1e-10 relative of the oracle (the plain FP64
tolerance) and within 1e-12 relative of the plain path on the same system;
bitwise identity where the arithmetic is the same (a system alone or in a
mix, any dispatch order or grid, any tile variant at a fixed window edge:
the K accumulation order of every W element does not depend on the tile).
"""
import numpy as np
import pytest

from atrip_amd import hip, workloads

pytestmark = pytest.mark.gpu

REL_ORACLE = 1e-10
REL_PLAIN = 1e-12
N_TILES = 10  # wave-tile variants of the batched kernel (batch_tile 0..9)

SHAPES = [(4, 20, 1000), (5, 23, 1001), (7, 31, 1002), (16, 40, 1003),
          (9, 57, 1004), (12, 33, 1005), (1, 5, 1006), (2, 3, 1007),
          (3, 2, 1008), (33, 12, 1009)]


@pytest.fixture(scope="module")
def batch(hip_ctx):
    systems = [hip_ctx.upload_synthetic(no, nv, seed) for no, nv, seed in SHAPES]
    yield systems
    for s in systems:
        hip_ctx.release(s)


def _reset(ctx):
    ctx.set_option("batch_nb", 0)
    ctx.set_option("batch_tile", -1)
    ctx.set_option("batch_order", 0)
    ctx.set_option("batch_grid", 0)


def _close(a, b, rel):
    return abs(a - b) <= rel * abs(b) + 1e-18


def test_batched_matches_oracle(hip_ctx, oracle, batch):
    print("batched vs oracle", SHAPES, flush=True)
    e = hip_ctx.batch_run(batch)
    for (no, nv, seed), ei in zip(SHAPES, e):
        t = oracle.synthetic(no, nv, seed)
        e_ref, _ = oracle.run(no, nv, t)
        assert _close(ei, e_ref, REL_ORACLE), (no, nv, ei, e_ref)


def test_batched_matches_plain(hip_ctx, batch):
    print("batched vs plain", flush=True)
    e = hip_ctx.batch_run(batch)
    for s, ei, shape in zip(batch, e, SHAPES):
        ep, _ = hip_ctx.run(s)
        assert _close(ei, ep, REL_PLAIN), (shape, ei, ep)


def test_batched_alone_equals_mix(hip_ctx, batch):
    print("batched alone vs mix", flush=True)
    mix = hip_ctx.batch_run(batch)
    for q, s in enumerate(batch):
        alone = hip_ctx.batch_run([s])
        assert alone[0] == mix[q], SHAPES[q]
    # repeated and reordered lists
    rev = hip_ctx.batch_run(batch[::-1] + batch[:3])
    assert np.array_equal(rev[:len(batch)], mix[::-1])
    assert np.array_equal(rev[len(batch):], mix[:3])


def test_batched_order_and_grid_independent(hip_ctx, batch):
    print("batched order/grid", flush=True)
    ref = hip_ctx.batch_run(batch)
    try:
        hip_ctx.set_option("batch_order", 1)
        assert np.array_equal(hip_ctx.batch_run(batch), ref)
        hip_ctx.set_option("batch_grid", 3)
        assert np.array_equal(hip_ctx.batch_run(batch), ref)
    finally:
        _reset(hip_ctx)


def test_batched_tiles_and_windows(hip_ctx, batch):
    print("batched tiles x windows", flush=True)
    plain = np.array([hip_ctx.run(s)[0] for s in batch])
    try:
        for nb in (1, 2, 4, 8):
            hip_ctx.set_option("batch_nb", nb)
            first = None
            for tile in range(N_TILES):
                print(f"  nb={nb} tile={tile}", flush=True)
                hip_ctx.set_option("batch_tile", tile)
                e = hip_ctx.batch_run(batch)
                for q in range(len(batch)):
                    assert _close(e[q], plain[q], REL_PLAIN), (nb, tile, SHAPES[q])
                if first is None:
                    first = e
                else:
                    assert np.array_equal(e, first), (nb, tile)
    finally:
        _reset(hip_ctx)


def test_batched_tuner(hip_ctx, batch):
    print("batched tuner", flush=True)
    ref = hip_ctx.batch_run(batch)
    hip_ctx.batch_tune(batch, measure=True)
    for s in batch:
        p = hip_ctx.batch_plan(s)
        assert p["measured"], p
        assert p["nb"] in (1, 2, 4, 8) and p["fm"] in (1, 2, 3, 4) and p["fn"] in (1, 2, 3, 4)
        assert p["sec_window"] > 0
    e = hip_ctx.batch_run(batch)
    for q in range(len(batch)):
        assert _close(e[q], ref[q], REL_PLAIN)


def test_batched_seeded_sample(hip_ctx):
    """The first 24 systems of the seeded 256-system workload against the
    plain path."""
    shapes = workloads.batch_shapes(24)
    print("batched seeded sample", shapes, flush=True)
    systems = [hip_ctx.upload_synthetic(no, nv, seed) for no, nv, seed in shapes]
    try:
        e = hip_ctx.batch_run(systems)
        for s, ei, sh in zip(systems, e, shapes):
            ep, _ = hip_ctx.run(s)
            assert _close(ei, ep, REL_PLAIN), (sh, ei, ep)
    finally:
        for s in systems:
            hip_ctx.release(s)


def test_batched_edge_sizes(hip_ctx):
    """The largest No the batched form takes (64) and an empty tuple space
    (Nv = 1: no a <= b <= c other than a = b = c) in one launch with an
    ordinary system: each equals the plain path, the empty one exactly 0."""
    shapes = [(64, 4, 1010), (3, 1, 1011), (6, 9, 1012)]
    print("batched edge sizes", shapes, flush=True)
    systems = [hip_ctx.upload_synthetic(no, nv, seed) for no, nv, seed in shapes]
    try:
        e = hip_ctx.batch_run(systems)
        for s, ei, sh in zip(systems, e, shapes):
            ep, _ = hip_ctx.run(s)
            assert _close(ei, ep, REL_PLAIN), (sh, ei, ep)
        assert e[1] == 0.0 and hip_ctx.run(systems[1])[0] == 0.0
    finally:
        for s in systems:
            hip_ctx.release(s)


def test_batched_errors(hip_ctx):
    assert len(hip_ctx.batch_run([])) == 0
    with pytest.raises(hip.HipError, match="unknown system"):
        hip_ctx.batch_run([987654])
    big = hip_ctx.upload_synthetic(65, 3, 1)
    try:
        with pytest.raises(hip.HipError, match="No <= 64"):
            hip_ctx.batch_run([big])
    finally:
        hip_ctx.release(big)
    with pytest.raises(hip.HipError, match="batch_nb"):
        hip_ctx.set_option("batch_nb", 3)


# The plain mixed kernels accumulate on 32x32 MFMAs, the batched form on
# 16x16 ones: the same products summed in other FP32 groupings, so the two
# agree to the FP32-accumulation level, and each is within the mixed
# tolerance of FP64.
REL_MIXED_PLAIN = 1e-7
REL_MIXED_FP64 = 1e-7


@pytest.mark.parametrize("precision", ["split_fp32", "fp32"])
def test_batched_mixed_forms(hip_ctx, batch, precision):
    """The mixed forms in the batched kernel: E(T) agrees with the plain
    path's mixed run and with FP64 to the mixed tolerances; alone equals
    mix."""
    print("batched", precision, flush=True)
    e = hip_ctx.batch_run(batch, precision)
    for s, ei, shape in zip(batch, e, SHAPES):
        ep, _ = hip_ctx.run(s, precision)
        e64, _ = hip_ctx.run(s)
        assert _close(ei, ep, REL_MIXED_PLAIN), (shape, ei, ep)
        assert _close(ei, e64, REL_MIXED_FP64), (shape, ei, e64)
    assert hip_ctx.batch_run(batch[3:4], precision)[0] == e[3]
    hip_ctx.batch_tune(batch[:4], measure=True, precision=precision)
    assert hip_ctx.batch_plan(batch[0], precision)["measured"]
    e2 = hip_ctx.batch_run(batch, precision)
    for q in range(len(batch)):
        assert _close(e2[q], e[q], REL_PLAIN)


BF16X3_OFF = ("the batched kernel's bf16x3 wave tile went silent for 180 s on "
              "its first GPU run without a cause found in its code; it is "
              "taken out and batch_run refuses bf16x3")


def test_batched_refuses_bf16x3(hip_ctx, batch):
    with pytest.raises(hip.HipError, match="not bf16x3"):
        hip_ctx.batch_run(batch[:2], "bf16x3")


@pytest.mark.skip(reason=BF16X3_OFF)
def test_batched_bf16x3_mixed_form(hip_ctx, batch):
    """bf16x3 in the batched kernel: E(T) within the mixed tolerance of
    FP64 and of the plain path's bf16x3 run; alone equals mix."""
    e = hip_ctx.batch_run(batch, "bf16x3")
    for s, ei, shape in zip(batch, e, SHAPES):
        ep, _ = hip_ctx.run(s, "bf16x3")
        e64, _ = hip_ctx.run(s)
        assert _close(ei, ep, REL_MIXED_PLAIN), (shape, ei, ep)
        assert _close(ei, e64, REL_MIXED_FP64), (shape, ei, e64)
    assert hip_ctx.batch_run(batch[3:4], "bf16x3")[0] == e[3]


@pytest.mark.skip(reason=BF16X3_OFF)
def test_batched_bf16x3_tiles_and_windows(hip_ctx, batch):
    """bf16x3 in the batched kernel for every window edge and wave tile
    (half-empty last K steps for No or Nv not a multiple of 32, A rows past
    the padded slice): each system within the mixed tolerance of FP64; the
    K accumulation order of a W element does not depend on the tile, so at
    a fixed window edge every tile gives the same bits."""
    print("batched bf16x3 tiles x windows", flush=True)
    plain = np.array([hip_ctx.run(s)[0] for s in batch])
    try:
        for nb in (1, 2, 4):
            hip_ctx.set_option("batch_nb", nb)
            first = None
            for tile in range(N_TILES):
                print(f"  nb={nb} tile={tile}", flush=True)
                hip_ctx.set_option("batch_tile", tile)
                e = hip_ctx.batch_run(batch, "bf16x3")
                for q in range(len(batch)):
                    assert _close(e[q], plain[q], REL_MIXED_FP64), (nb, tile, SHAPES[q])
                if first is None:
                    first = e
                else:
                    assert np.array_equal(e, first), (nb, tile)
    finally:
        _reset(hip_ctx)
