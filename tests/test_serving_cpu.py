"""Host logic of the serving loop (no GPU): admission order, per-system frame
order, overlap across systems, HBM budget with eviction, latency summary."""
import threading
import time

import numpy as np
import pytest

from dmff_amd.admp.serving import ADMPServer, Request, Result, SystemSpec, estimate_bytes, latency_summary


def _spec(key, n=300, rc=8.0, q=None):
    return SystemSpec(key, n, None, None, None, q, None, None, None, None, None, rc)


class _Recorder:
    def __init__(self, dt=0.01):
        self.dt = dt
        self.lock = threading.Lock()
        self.active = {}
        self.max_conc = 0
        self.log = []
        self.batches = []

    def __call__(self, reqs, lives, t_start):
        keys = [r.spec.key for r in reqs]
        assert len(set(keys)) == len(keys), "two frames of one system in one batch"
        assert len({r.spec.class_key() for r in reqs}) == 1, "a batch mixes classes"
        with self.lock:
            for key in keys:
                assert self.active.get(key, 0) == 0, "two frames of one system ran at once"
                self.active[key] = 1
            self.max_conc = max(self.max_conc, sum(self.active.values()))
            self.batches.append([r.rid for r in reqs])
            for r in reqs:
                self.log.append((r.spec.key, r.rid))
        time.sleep(self.dt)
        with self.lock:
            for key in keys:
                self.active[key] = 0
        return [Result(r.rid, r.spec.key, 0.0, None, None, 0, 0.0, False, r.arrival, t_start, time.perf_counter(),
                       batch=len(reqs)) for r in reqs]


def _box(L=40.0):
    return np.eye(3) * L


def test_frames_of_a_system_keep_order_and_systems_overlap():
    rec = _Recorder()
    srv = ADMPServer(device="cpu", slots=4, hbm_budget=1 << 40, runner=rec)
    specs = [_spec(f"s{k}") for k in range(4)]
    futs = []
    rid = 0
    for rep in range(3):
        for sp in specs:
            futs.append(srv.submit(Request(rid, sp, None, _box())))
            rid += 1
    for f in futs:
        f.result(timeout=30)
    srv.close()
    assert rec.max_conc > 1
    for sp in specs:
        order = [r for k, r in rec.log if k == sp.key]
        assert order == sorted(order)
    assert len(srv.results) == len(futs)


def test_waiting_frames_of_one_class_run_as_one_batch():
    """With one slot busy, frames of four same-class systems and of a different
    class wait; the next admission takes the same-class frames as one batch
    (one frame per system, in arrival order) and leaves the other class and the
    second frame of a batched system for later."""
    rec = _Recorder(dt=0.05)
    srv = ADMPServer(device="cpu", slots=1, hbm_budget=1 << 40, max_batch=8, runner=rec)
    a = [_spec(f"a{k}") for k in range(4)]
    other = _spec("b0", q=np.ones(3))
    futs = [srv.submit(Request(0, _spec("first", n=100), None, _box()))]
    rid = 1
    for sp in (a[0], other, a[1], a[2], a[0], a[3]):
        futs.append(srv.submit(Request(rid, sp, None, _box())))
        rid += 1
    for f in futs:
        f.result(timeout=30)
    srv.close()
    assert rec.batches[0] == [0]
    assert rec.batches[1] == [1, 3, 4, 6]          # a0, a1, a2, a3; a0's second frame waits
    assert sorted(map(tuple, rec.batches[2:])) == [(2,), (5,)]
    assert [r for k, r in rec.log if k == "a0"] == [1, 5]


def test_budget_is_respected_and_idle_systems_are_evicted():
    need = estimate_bytes(300, (26, 26, 26), 8.0)
    rec = _Recorder(dt=0.005)
    srv = ADMPServer(device="cpu", slots=2, hbm_budget=int(2.5 * need), max_batch=1, runner=rec)
    peak = []
    orig = srv._admit_locked

    def watch():
        orig()
        peak.append(srv.used)

    srv._admit_locked = watch
    # six systems of six classes: six contexts, at most two of which fit
    futs = [srv.submit(Request(k, _spec(f"s{k}", q=np.full(3, float(k))), None, _box())) for k in range(6)]
    for f in futs:
        f.result(timeout=30)
    srv.close()
    assert max(peak) <= srv.budget
    assert len(srv.ctxs) <= 2 and len(rec.batches) == 6


def test_latency_summary_percentiles():
    class R:
        def __init__(self, v):
            self.latency_ms = v

    s = latency_summary([R(float(v)) for v in range(1, 101)])
    assert s["n"] == 100
    assert s["p50_ms"] == pytest.approx(50.5)
    assert s["p99_ms"] == pytest.approx(99.01)
