"""Host logic of the batched and served forms that needs no GPU: the seeded
workload shapes and the batched cost model."""
import math

import pytest

from atrip_amd import workloads


def test_splitmix64_known_values():
    # SplitMix64 outputs for states 0 and 1 (the published generator)
    assert workloads.mix64(0) == 0xE220A8397B1DCDAF
    assert workloads.mix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_batch_shapes_ranges_and_reproducible():
    sh = workloads.batch_shapes(256, seed=0)
    assert sh == workloads.batch_shapes(256, seed=0)
    assert sh != workloads.batch_shapes(256, seed=1)
    assert len(sh) == 256
    assert all(4 <= no <= 16 and 20 <= nv <= 120 for no, nv, _ in sh)
    assert [s for _, _, s in sh] == list(range(1000, 1256))
    # every value of the ranges shows up (uniform draws)
    assert {no for no, _, _ in sh} == set(range(4, 17))
    mean_nv = sum(nv for _, nv, _ in sh) / 256
    assert 60 < mean_nv < 80


def test_serve_requests_poisson_and_log_uniform():
    req = workloads.serve_requests(2000, rate=20.0, seed=0)
    t = [r[0] for r in req]
    assert all(b > a for a, b in zip(t, t[1:]))
    rate = len(t) / t[-1]
    assert 18.0 < rate < 22.0
    assert all(8 <= r[1] <= 40 and 40 <= r[2] <= 320 for r in req)
    # log-uniform: the geometric midpoint splits the draws about evenly
    below = sum(r[2] < math.sqrt(40 * 320) for r in req) / len(req)
    assert 0.45 < below < 0.55


def test_flop_and_tuple_counts():
    assert workloads.n_tuples(50) == 22050
    assert workloads.n_tuples(400) == 10746400
    assert workloads.doubles_flop(10, 50) == pytest.approx(1.588e10, rel=1e-3)


def test_batch_model_empty_tuple_space():
    """Nv = 1 holds no tuple: the model still returns a valid window edge."""
    from atrip_amd import hip

    nb, fm, fn = hip.batch_model(3, 1)
    assert nb >= 1 and fm >= 1 and fn >= 1


def test_batch_model_choices():
    from atrip_amd import hip

    for no, nv, _ in workloads.batch_shapes(64):
        nb, fm, fn = hip.batch_model(no, nv)
        assert nb in (1, 2, 4, 8) and fm in (1, 2, 3, 4) and fn in (1, 2, 3, 4)
        # the chosen window never needs more than the scratch bound
        assert nb == 1 or nb ** 3 * no ** 3 * 8 * 256 <= 8192 * 2 ** 20
    # a tight scratch bound forces small windows
    nb, _, _ = hip.batch_model(16, 120, scratch_mb=16, grid=256)
    assert nb <= 2
    with pytest.raises(hip.HipError):
        hip.batch_model(0, 10)
