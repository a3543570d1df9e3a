"""Host logic of the serving loop (no GPU): admission order, per-system frame
order, overlap across systems, HBM budget with eviction, latency summary."""
import threading
import time

import numpy as np
import pytest

from dmff_amd.admp.serving import ADMPServer, Request, Result, SystemSpec, estimate_bytes, latency_summary


def _spec(key, n=300, rc=8.0):
    return SystemSpec(key, n, None, None, None, None, None, None, None, None, None, rc)


class _Recorder:
    def __init__(self, dt=0.01):
        self.dt = dt
        self.lock = threading.Lock()
        self.active = {}
        self.max_conc = 0
        self.log = []

    def __call__(self, req, lv, t_start):
        key = req.spec.key
        with self.lock:
            assert self.active.get(key, 0) == 0, "two frames of one system ran at once"
            self.active[key] = 1
            self.max_conc = max(self.max_conc, sum(self.active.values()))
            self.log.append((key, req.rid))
        time.sleep(self.dt)
        with self.lock:
            self.active[key] = 0
        return Result(req.rid, key, 0.0, None, None, 0, 0.0, False, req.arrival, t_start, time.perf_counter())


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


def test_budget_is_respected_and_idle_systems_are_evicted():
    need = estimate_bytes(300, (26, 26, 26), 8.0)
    rec = _Recorder(dt=0.005)
    srv = ADMPServer(device="cpu", slots=2, hbm_budget=int(2.5 * need), runner=rec)
    peak = []
    orig = srv._admit_locked

    def watch():
        orig()
        peak.append(srv.used)

    srv._admit_locked = watch
    futs = [srv.submit(Request(k, _spec(f"s{k}"), None, _box())) for k in range(6)]
    for f in futs:
        f.result(timeout=30)
    srv.close()
    assert max(peak) <= srv.budget
    assert len(srv.systems) <= 2


def test_latency_summary_percentiles():
    class R:
        def __init__(self, v):
            self.latency_ms = v

    s = latency_summary([R(float(v)) for v in range(1, 101)])
    assert s["n"] == 100
    assert s["p50_ms"] == pytest.approx(50.5)
    assert s["p99_ms"] == pytest.approx(99.01)
