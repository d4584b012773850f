"""Serving form: the continuous-batching request loop of libatrip_hip
(atrip_hip_serve_*), and a Poisson-arrival driver that reports latency
percentiles and throughput.

The loop itself (admission by HBM budget, window selection, the persistent
launch, retirement) is native; this module only binds it and replays an
arrival process against its clock.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import hip, workloads

QUEUED, ACTIVE, DONE, FAILED = 0, 1, 2, 3


class Server:
    """A serving loop on a hip.Context (which must outlive it)."""

    def __init__(self, ctx: hip.Context, hbm_budget_gb: float = 200.0,
                 max_active: int = 64, iter_ms: float = 100.0, policy: int = 0,
                 precision: str = "fp64"):
        self.ctx = ctx
        self._L = ctx._L
        self._h = C.c_void_p()
        ctx._check(self._L.atrip_hip_serve_create(
            ctx._h, int(hbm_budget_gb * 1e9), max_active, iter_ms,
            hip.PRECISIONS[precision], C.byref(self._h)), "serve_create")
        if policy:
            ctx._check(self._L.atrip_hip_serve_set_policy(self._h, policy),
                       "serve_set_policy")
        self._keep = {}

    def close(self):
        if self._h:
            self._L.atrip_hip_serve_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def clock_ms(self) -> float:
        return self._L.atrip_hip_serve_clock_ms(self._h)

    def submit_synthetic(self, no: int, nv: int, seed: int, arrival_ms: float = -1.0) -> int:
        rid = C.c_int64()
        self.ctx._check(self._L.atrip_hip_serve_submit_synthetic(
            self._h, no, nv, seed, arrival_ms, C.byref(rid)), "serve_submit")
        return rid.value

    def submit(self, no: int, nv: int, tensors: dict, arrival_ms: float = -1.0) -> int:
        d = hip.SysDesc()
        d.No, d.Nv = no, nv
        keep = []
        for name in hip.TENSORS:
            arr = tensors.get(name)
            if arr is None or name in ("Jijka", "Jabci"):
                setattr(d, name, C.POINTER(C.c_double)())
                continue
            arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
            keep.append(arr)
            setattr(d, name, hip._p(arr))
        rid = C.c_int64()
        self.ctx._check(self._L.atrip_hip_serve_submit(
            self._h, C.byref(d), arrival_ms, C.byref(rid)), "serve_submit")
        self._keep[rid.value] = keep  # read at admission
        return rid.value

    def step(self) -> int:
        n = C.c_int()
        self.ctx._check(self._L.atrip_hip_serve_step(self._h, C.byref(n)), "serve_step")
        return n.value

    def result(self, rid: int, raise_failed: bool = True) -> dict:
        """State, energy and times of a request; a failed request raises
        HipError with its reason (or is returned with state FAILED and the
        reason under "error" when raise_failed is False)."""
        st, e = C.c_int(), C.c_double()
        ts, ta, td = C.c_double(), C.c_double(), C.c_double()
        rc = self._L.atrip_hip_serve_result(
            self._h, rid, C.byref(st), C.byref(e), C.byref(ts), C.byref(ta),
            C.byref(td))
        out = dict(state=st.value, energy=e.value, t_submit=ts.value,
                   t_admit=ta.value, t_done=td.value)
        if rc != 0:
            if st.value != FAILED or raise_failed:
                self.ctx._check(rc, f"serve_result({rid})")
            out["error"] = self._L.atrip_hip_last_error(self.ctx._h).decode()
        if st.value in (DONE, FAILED):
            self._keep.pop(rid, None)
        return out

    def stats(self) -> dict:
        q, a = C.c_int(), C.c_int()
        res, peak, tup, nl = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        busy = C.c_double()
        self.ctx._check(self._L.atrip_hip_serve_stats(
            self._h, C.byref(q), C.byref(a), C.byref(res), C.byref(peak),
            C.byref(tup), C.byref(nl), C.byref(busy)), "serve_stats")
        return dict(queued=q.value, active=a.value, resident_bytes=res.value,
                    peak_bytes=peak.value, tuples_done=tup.value,
                    launches=nl.value, busy_ms=busy.value)

    def drain(self, ids):
        """Step until every request in ids has ended; returns their results."""
        pending = set(ids)
        out = {}
        while pending:
            self.step()
            for rid in list(pending):
                r = self.result(rid, raise_failed=False)
                if r["state"] in (DONE, FAILED):
                    out[rid] = r
                    pending.discard(rid)
        return out


def run_poisson(ctx: hip.Context, n_requests: int, rate: float = 20.0,
                seed: int = 0, hbm_budget_gb: float = 200.0,
                max_active: int = 64, iter_ms: float = 100.0, policy: int = 0,
                no=(8, 40), nv=(40, 320), precision: str = "fp64",
                log=None) -> dict:
    """Replay a Poisson arrival window of n_requests against the server's
    clock (requests are submitted when their arrival time has passed, stamped
    with that arrival time) and report latency percentiles and throughput."""
    reqs = workloads.serve_requests(n_requests, rate=rate, seed=seed, no=no, nv=nv)
    with Server(ctx, hbm_budget_gb, max_active, iter_ms, policy, precision) as srv:
        t0 = srv.clock_ms()
        ids = []
        nxt = 0
        done = {}
        while len(done) < len(reqs):
            now = srv.clock_ms()
            while nxt < len(reqs) and t0 + 1e3 * reqs[nxt][0] <= now:
                a, rno, rnv, rseed = reqs[nxt]
                ids.append(srv.submit_synthetic(rno, rnv, rseed, t0 + 1e3 * a))
                nxt += 1
            if srv.stats()["queued"] + srv.stats()["active"] == 0:
                if nxt < len(reqs):  # idle until the next arrival
                    time.sleep(max(0.0, (t0 + 1e3 * reqs[nxt][0] - now) * 1e-3))
                continue
            srv.step()
            for q, rid in enumerate(ids):
                if rid in done:
                    continue
                r = srv.result(rid, raise_failed=False)
                if r["state"] in (DONE, FAILED):
                    done[rid] = r
                    if log:
                        log(f"request {q} No={reqs[q][1]} Nv={reqs[q][2]} "
                            f"latency {r['t_done'] - r['t_submit']:.0f} ms")
        st = srv.stats()
        t_end = max(r["t_done"] for r in done.values())
    ok = [r for r in done.values() if r["state"] == DONE]
    if not ok:
        raise hip.HipError("no served request finished")
    lat = np.array([r["t_done"] - r["t_submit"] for r in ok])
    queue = np.array([r["t_admit"] - r["t_submit"] for r in ok])
    span_s = (t_end - t0) * 1e-3
    flop = sum(workloads.doubles_flop(r[1], r[2]) for r in reqs)
    return dict(
        precision=precision, requests=len(reqs), rate_req_s=rate, p50_ms=float(np.percentile(lat, 50)),
        p99_ms=float(np.percentile(lat, 99)), mean_ms=float(lat.mean()),
        max_ms=float(lat.max()), p50_queue_ms=float(np.percentile(queue, 50)),
        span_s=span_s, tuples=int(st["tuples_done"]),
        tuples_per_s=st["tuples_done"] / span_s,
        tflops=flop / span_s / 1e12, busy_fraction=st["busy_ms"] * 1e-3 / span_s,
        launches=st["launches"], peak_resident_gb=st["peak_bytes"] / 1e9,
        failed=len(done) - len(ok),
        energies={ids.index(k): v["energy"] for k, v in done.items()
                  if v["state"] == DONE},
    )
