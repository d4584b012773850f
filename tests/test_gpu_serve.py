"""Serving form: continuous batching over the persistent batched kernel.

Bar: every served request's E(T) is bitwise the same alone and in any mix
(any arrival pattern, admission order, policy or HBM budget), within 1e-12
relative of the plain path on the same system and within 1e-10 relative of
the oracle for host-tensor requests; the resident bytes never exceed the
budget.
"""
import pytest

from atrip_amd import serve

pytestmark = pytest.mark.gpu

REL_PLAIN = 1e-12
REL_ORACLE = 1e-10

MIX = [(8, 40, 5000), (12, 33, 5001), (9, 57, 5002), (16, 45, 5003),
       (10, 28, 5004), (20, 41, 5005), (8, 66, 5006)]


def _close(a, b, rel):
    return abs(a - b) <= rel * abs(b) + 1e-18


@pytest.fixture(scope="module")
def alone(hip_ctx):
    """E(T) of every MIX request served alone, and by the plain path."""
    out = {}
    for no, nv, seed in MIX:
        with serve.Server(hip_ctx, hbm_budget_gb=50) as srv:
            rid = srv.submit_synthetic(no, nv, seed)
            r = srv.drain([rid])[rid]
        assert r["state"] == serve.DONE
        s = hip_ctx.upload_synthetic(no, nv, seed)
        try:
            plain, _ = hip_ctx.run(s)
        finally:
            hip_ctx.release(s)
        out[seed] = (r["energy"], plain)
    return out


def test_alone_matches_plain(alone):
    for seed, (e, plain) in alone.items():
        assert _close(e, plain, REL_PLAIN), (seed, e, plain)


@pytest.mark.parametrize("policy,iter_ms", [(0, 20.0), (1, 5.0), (0, 1000.0)])
def test_same_alone_and_in_the_mix(hip_ctx, alone, policy, iter_ms):
    """Requests arrive while others are on the device: three first, the rest
    one per step; each result equals the request served alone, bitwise."""
    print("serve mix policy", policy, "iter_ms", iter_ms, flush=True)
    with serve.Server(hip_ctx, hbm_budget_gb=50, iter_ms=iter_ms, policy=policy) as srv:
        ids = [srv.submit_synthetic(*MIX[q]) for q in range(3)]
        for q in range(3, len(MIX)):
            srv.step()
            ids.append(srv.submit_synthetic(*MIX[q]))
        res = srv.drain(ids)
        st = srv.stats()
    assert st["launches"] >= 2
    for rid, (no, nv, seed) in zip(ids, MIX):
        r = res[rid]
        assert r["state"] == serve.DONE
        assert r["energy"] == alone[seed][0], (no, nv, seed)
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]


def test_empty_request_completes(hip_ctx):
    """A request with no tuple (Nv = 1) retires at admission with E(T) = 0,
    next to an ordinary request whose result it does not change."""
    with serve.Server(hip_ctx, hbm_budget_gb=50.0, iter_ms=10.0) as srv:
        empty = srv.submit_synthetic(3, 1, 5100)
        other = srv.submit_synthetic(5, 9, 5101)
        for _ in range(1000):
            srv.step()
            if all(srv.result(r)["state"] == serve.DONE for r in (empty, other)):
                break
        assert srv.result(empty)["state"] == serve.DONE
        assert srv.result(empty)["energy"] == 0.0
        assert srv.result(other)["state"] == serve.DONE
        e_other = srv.result(other)["energy"]
    s = hip_ctx.upload_synthetic(5, 9, 5101)
    try:
        assert _close(e_other, hip_ctx.run(s)[0], 1e-12)
    finally:
        hip_ctx.release(s)


def test_budget_admission(hip_ctx, alone):
    """A budget that holds about two of the systems: requests wait for
    memory, the resident bytes stay under the budget, results unchanged."""
    from atrip_amd import hip

    sizes = []
    for no, nv, seed in MIX:
        s = hip_ctx.upload_synthetic(no, nv, seed)
        sizes.append(hip_ctx.sys_info(s)["bytes"])
        hip_ctx.release(s)
    budget = sorted(sizes)[-1] + sorted(sizes)[-2]
    print("serve budget bytes", budget, flush=True)
    with serve.Server(hip_ctx, hbm_budget_gb=budget / 1e9, iter_ms=10.0) as srv:
        ids = [srv.submit_synthetic(*m) for m in MIX]
        res = {}
        while len(res) < len(ids):
            srv.step()
            st = srv.stats()
            assert st["resident_bytes"] <= budget
            for rid in ids:
                r = srv.result(rid)
                if r["state"] == serve.DONE:
                    res[rid] = r
        assert srv.stats()["peak_bytes"] <= budget
    for rid, (no, nv, seed) in zip(ids, MIX):
        assert res[rid]["energy"] == alone[seed][0]
    # a request larger than the whole budget fails with its reason
    with serve.Server(hip_ctx, hbm_budget_gb=1e-6) as srv:
        rid = srv.submit_synthetic(8, 40, 1)
        srv.step()
        with pytest.raises(hip.HipError, match="larger than the HBM budget"):
            srv.result(rid)


def test_host_tensor_requests(hip_ctx, oracle):
    print("serve host tensors", flush=True)
    shapes = [(6, 14, 41), (9, 20, 42)]
    with serve.Server(hip_ctx, hbm_budget_gb=10) as srv:
        ids = []
        refs = []
        for no, nv, seed in shapes:
            t = oracle.synthetic(no, nv, seed)
            refs.append(oracle.run(no, nv, t)[0])
            ids.append(srv.submit(no, nv, t))
        res = srv.drain(ids)
    for rid, ref in zip(ids, refs):
        assert _close(res[rid]["energy"], ref, REL_ORACLE)


def test_poisson_driver_reports_percentiles(hip_ctx):
    print("serve poisson", flush=True)
    r = serve.run_poisson(hip_ctx, 12, rate=40.0, seed=3, hbm_budget_gb=50,
                          iter_ms=20.0, no=(8, 12), nv=(20, 48))
    assert r["requests"] == 12 and len(r["energies"]) == 12
    assert 0 < r["p50_ms"] <= r["p99_ms"] <= r["max_ms"]
    assert r["tuples"] > 0 and r["tuples_per_s"] > 0


def test_mixed_precision_serving(hip_ctx):
    """A split-FP32 server: each request keeps only its split operands, and
    its energy is the same alone and in the mix and agrees with the plain
    path's split-FP32 run to the FP32-accumulation level."""
    print("serve split_fp32", flush=True)
    mix = MIX[:4]
    alone = {}
    for m in mix:
        with serve.Server(hip_ctx, hbm_budget_gb=50, precision="split_fp32") as srv:
            rid = srv.submit_synthetic(*m)
            alone[m] = srv.drain([rid])[rid]["energy"]
    with serve.Server(hip_ctx, hbm_budget_gb=50, iter_ms=5.0,
                      precision="split_fp32") as srv:
        ids = [srv.submit_synthetic(*mix[0])]
        for m in mix[1:]:
            srv.step()
            ids.append(srv.submit_synthetic(*m))
        res = srv.drain(ids)
    for rid, m in zip(ids, mix):
        assert res[rid]["energy"] == alone[m]
        s = hip_ctx.upload_synthetic(*m)
        try:
            ep, _ = hip_ctx.run(s, "split_fp32")
        finally:
            hip_ctx.release(s)
        # the plain path's split kernel accumulates on 32x32 MFMAs, the
        # batched kernel behind the server on 16x16 ones
        assert _close(alone[m], ep, 1e-7)


@pytest.mark.parametrize("plain_no", [1, 12, 1000])
def test_plain_path_routing(hip_ctx, alone, plain_no):
    """Requests with No >= serve_plain_no run their windows through the plain
    class-GEMM launches, the rest through the persistent batched launch;
    whatever the split, every result is bitwise the same alone and in the
    mix, and within 1e-12 of the plain path's atrip_hip_run."""
    print("serve plain routing, serve_plain_no", plain_no, flush=True)
    hip_ctx.set_option("serve_plain_no", plain_no)
    try:
        solo = {}
        for m in MIX:
            with serve.Server(hip_ctx, hbm_budget_gb=50) as srv:
                rid = srv.submit_synthetic(*m)
                solo[m] = srv.drain([rid])[rid]["energy"]
        with serve.Server(hip_ctx, hbm_budget_gb=50, iter_ms=5.0) as srv:
            ids = [srv.submit_synthetic(*MIX[q]) for q in range(2)]
            for q in range(2, len(MIX)):
                srv.step()
                ids.append(srv.submit_synthetic(*MIX[q]))
            res = srv.drain(ids)
    finally:
        hip_ctx.set_option("serve_plain_no", 16)
    for rid, m in zip(ids, MIX):
        assert res[rid]["state"] == serve.DONE
        assert res[rid]["energy"] == solo[m], m
        assert _close(solo[m], alone[m[2]][1], REL_PLAIN), m


def test_serving_keeps_window_edge_of_admission(hip_ctx):
    """A plain-path request keeps the window edge it was admitted with: the
    context's window option changed in mid-stream leaves every result
    bitwise the same as served alone at the edge of admission."""
    print("serve window edge pinned", flush=True)
    mix = MIX[:4]
    hip_ctx.set_option("serve_plain_no", 1)
    try:
        solo = {}
        for m in mix:
            with serve.Server(hip_ctx, hbm_budget_gb=50) as srv:
                rid = srv.submit_synthetic(*m)
                solo[m] = srv.drain([rid])[rid]["energy"]
        with serve.Server(hip_ctx, hbm_budget_gb=50, iter_ms=2.0) as srv:
            ids = [srv.submit_synthetic(*m) for m in mix]
            srv.step()  # admits all four at the automatic edge
            hip_ctx.set_option("window", 4)
            res = srv.drain(ids)
    finally:
        hip_ctx.set_option("window", 0)
        hip_ctx.set_option("serve_plain_no", 16)
    for rid, m in zip(ids, mix):
        assert res[rid]["state"] == serve.DONE
        assert res[rid]["energy"] == solo[m], m


def test_bf16x3_serving(hip_ctx):
    """A bf16x3 server (every request on the plain path, bf16x3 operands
    only): each result bitwise the same alone and in the mix, equal to the
    plain path's bf16x3 run to rounding, and within the mixed tolerance of
    FP64."""
    print("serve bf16x3", flush=True)
    mix = MIX[:5]
    solo = {}
    for m in mix:
        with serve.Server(hip_ctx, hbm_budget_gb=50, precision="bf16x3") as srv:
            rid = srv.submit_synthetic(*m)
            solo[m] = srv.drain([rid])[rid]["energy"]
    with serve.Server(hip_ctx, hbm_budget_gb=50, iter_ms=2.0,
                      precision="bf16x3") as srv:
        ids = [srv.submit_synthetic(*mix[0])]
        for m in mix[1:]:
            srv.step()
            ids.append(srv.submit_synthetic(*m))
        res = srv.drain(ids)
    for rid, m in zip(ids, mix):
        assert res[rid]["energy"] == solo[m], m
        s = hip_ctx.upload_synthetic(*m)
        try:
            e3, _ = hip_ctx.run(s, "bf16x3")
            e64, _ = hip_ctx.run(s)
        finally:
            hip_ctx.release(s)
        assert _close(solo[m], e3, 1e-12), m
        assert _close(solo[m], e64, 1e-7), m
