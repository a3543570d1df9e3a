"""Continuous-batching request loop for ADMP PME evaluations on one GPU.

A request is one configuration (positions, box) of a periodic system with its
force-field data.  The server keeps every live system on the device -- its
native context (PME grids, FFT tables, neighbour lists, solver scratch) and the
last induced dipoles as the warm start of its next frame -- and runs requests
concurrently in ``slots`` worker threads, each on its own stream.  The library
releases the interpreter lock inside every native call and each context works
on streams of its own, so the work of different requests overlaps on the GPU.

Admission (continuous batching): requests wait in arrival order; whenever a
slot is free the scheduler takes the first waiting request whose system is idle
and whose HBM footprint fits the budget (evicting idle systems, least recently
used first, when that makes room).  A request never overtakes an earlier
request of the same system, so the warm-start chain of a system -- and with it
every result -- is the same whether it runs alone or in a mix (all reductions
in the library have a fixed order).

Per-request latency (arrival -> result ready) and queueing time are recorded;
``latency_summary`` gives p50/p99.
"""
import collections
import concurrent.futures
import math
import threading
import time

import numpy as np
import torch

from .engine import Engine
from .pme import setup_ewald_parameters

F64 = torch.float64


class SystemSpec:
    """Static description of a periodic system (Angstrom, e, kJ/mol)."""

    def __init__(self, key, n_atoms, axis_types, axis_idx, exclusions, Q_local, pol, thole,
                 mScales, pScales, dScales, rc, lmax=2, lpol=True, ethresh=5e-4):
        self.key = key
        self.n_atoms = int(n_atoms)
        self.axis_types, self.axis_idx, self.exclusions = axis_types, axis_idx, exclusions
        self.Q_local, self.pol, self.thole = Q_local, pol, thole
        self.scales = (mScales, pScales, dScales)
        self.rc, self.lmax, self.lpol, self.ethresh = float(rc), int(lmax), bool(lpol), float(ethresh)


class Request:
    def __init__(self, rid, spec, pos, box, arrival=None):
        self.rid = rid
        self.spec = spec
        self.pos = pos            # (N, 3) Angstrom
        self.box = box            # (3, 3) Angstrom, lattice vectors in rows
        self.arrival = time.perf_counter() if arrival is None else arrival
        self.future = concurrent.futures.Future()


class Result:
    def __init__(self, rid, key, energy, dpos, dbox, n_iter, residual, warm, t_arrival, t_start, t_end):
        self.rid, self.key = rid, key
        self.energy, self.dpos, self.dbox = energy, dpos, dbox
        self.n_iter, self.residual, self.warm = n_iter, residual, warm
        self.queued_ms = (t_start - t_arrival) * 1e3
        self.latency_ms = (t_end - t_arrival) * 1e3


class _Live:
    def __init__(self, spec, engine, nbytes, kappa, K):
        self.spec, self.engine, self.nbytes = spec, engine, nbytes
        self.kappa, self.K = kappa, K
        self.U = None
        self.busy = False
        self.last_used = time.perf_counter()
        self.dev = {}


def estimate_bytes(n_atoms, K, rc, precision="mixed", density=0.1003):
    """Upper estimate of a context's device footprint before it is created."""
    w = 4 if precision == "mixed" else 8
    cap = int(1.5 * density * 4.18879 * rc ** 3) + 64
    G = K[0] * K[1] * K[2]
    Gh = K[0] * K[1] * (K[2] // 2 + 1)
    per_site = 2400 + 60 * 8 + cap * (4 + 2 * w + 4)
    grids = G * (8 + 4 * 4) + Gh * 2 * 8 + 2 * G * 8
    return int(1.25 * (n_atoms * per_site + grids)) + (4 << 20)


class ADMPServer:
    def __init__(self, device="cuda:0", precision="mixed", slots=4, hbm_budget=None, tol=1e-6, maxiter=500,
                 runner=None):
        """runner(req, live, t_start) -> Result replaces the GPU evaluation (host-logic tests)."""
        self.device = torch.device(device)
        self.runner = runner
        self.precision, self.tol, self.maxiter = precision, float(tol), int(maxiter)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.5 * free)
        self.budget = int(hbm_budget)
        self.used = 0
        self.systems = {}
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.closed = False
        self.results = []
        self.n_slots = int(slots)
        self.free_slots = self.n_slots
        self.work = collections.deque()
        self.threads = [threading.Thread(target=self._worker, args=(k,), daemon=True) for k in range(self.n_slots)]
        for t in self.threads:
            t.start()

    # ------------------------------------------------------------------ API
    def submit(self, req):
        with self.cv:
            if self.closed:
                raise RuntimeError("server closed")
            self.pending.append(req)
            self._admit_locked()
        return req.future

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for t in self.threads:
            t.join()

    def drain(self, timeout=None):
        t0 = time.perf_counter()
        with self.cv:
            while self.pending or self.free_slots < self.n_slots or self.work:
                left = None if timeout is None else timeout - (time.perf_counter() - t0)
                if left is not None and left <= 0:
                    return False
                self.cv.wait(left)
        return True

    # ------------------------------------------------------------ scheduling
    def _footprint(self, spec, box):
        kappa, K1, K2, K3 = setup_ewald_parameters(spec.rc, spec.ethresh, np.asarray(box))
        return estimate_bytes(spec.n_atoms, (K1, K2, K3), spec.rc, self.precision), kappa, (K1, K2, K3)

    def _evict_for_locked(self, need):
        idle = sorted((lv for lv in self.systems.values() if not lv.busy), key=lambda lv: lv.last_used)
        for lv in idle:
            if self.used + need <= self.budget:
                break
            self.used -= lv.nbytes
            del self.systems[lv.spec.key]
            lv.engine = None
        return self.used + need <= self.budget

    def _admit_locked(self):
        blocked = set()
        i = 0
        while self.free_slots > 0 and i < len(self.pending):
            req = self.pending[i]
            key = req.spec.key
            if key in blocked:
                i += 1
                continue
            lv = self.systems.get(key)
            if lv is not None:
                if lv.busy:
                    blocked.add(key)   # keep the frame order of a system
                    i += 1
                    continue
                lv.busy = True
            else:
                need, kappa, K = self._footprint(req.spec, req.box)
                if self.used + need > self.budget and not self._evict_for_locked(need):
                    blocked.add(key)
                    i += 1
                    continue
                lv = _Live(req.spec, None, need, kappa, K)
                lv.busy = True
                self.systems[key] = lv
                self.used += need
            del self.pending[i]
            self.free_slots -= 1
            self.work.append((req, lv))
            self.cv.notify_all()

    # --------------------------------------------------------------- workers
    def _worker(self, k):
        stream = torch.cuda.Stream(device=self.device) if self.runner is None else None
        while True:
            with self.cv:
                while not self.work and not self.closed:
                    self.cv.wait()
                if not self.work and self.closed:
                    return
                req, lv = self.work.popleft()
            t_start = time.perf_counter()
            try:
                if self.runner is not None:
                    res = self.runner(req, lv, t_start)
                else:
                    with torch.cuda.device(self.device), torch.cuda.stream(stream):
                        res = self._run(req, lv, t_start)
                        stream.synchronize()
                req.future.set_result(res)
            except Exception as e:  # noqa: BLE001 - reported through the future
                req.future.set_exception(e)
            with self.cv:
                lv.busy = False
                lv.last_used = time.perf_counter()
                self.free_slots += 1
                if req.future.done() and not req.future.exception():
                    self.results.append(req.future.result())
                self._admit_locked()
                self.cv.notify_all()

    def _run(self, req, lv, t_start):
        spec = req.spec
        if lv.engine is None:
            eng = Engine(spec.n_atoms, spec.lmax, spec.lpol, True, lv.kappa, lv.K, spec.rc,
                         precision=self.precision, device=self.device)
            eng.set_topology(spec.axis_types, spec.axis_idx, spec.exclusions)
            lv.engine = eng
            nb = eng.memory_bytes()
            with self.cv:
                self.used += nb - lv.nbytes
                lv.nbytes = nb
            t = lambda x: torch.as_tensor(np.asarray(x), dtype=F64, device=self.device)
            lv.dev = {"Q": t(spec.Q_local), "pol": t(spec.pol), "thole": t(spec.thole)}
        eng = lv.engine
        pos = torch.as_tensor(np.asarray(req.pos), dtype=F64).to(self.device, non_blocking=False)
        box = torch.as_tensor(np.asarray(req.box), dtype=F64).to(self.device, non_blocking=False)
        eng.build_neighbors(pos, box)
        warm = lv.U is not None
        E, g, U, info = eng.compute(pos, box, lv.dev["Q"], lv.dev["pol"], lv.dev["thole"], *spec.scales,
                                    U_init=lv.U, solver="pcg", tol=self.tol, maxiter=self.maxiter, want_grad=True)
        lv.U = U
        t_end = time.perf_counter()
        return Result(req.rid, spec.key, float(E[0]), g["pos"][0], g["box"][0], info["n_iter"][0],
                      info["residual"][0], warm, req.arrival, t_start, t_end)


def latency_summary(results):
    lat = np.array([r.latency_ms for r in results])
    if len(lat) == 0:
        return {}
    return {"n": int(len(lat)), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}


# ---------------------------------------------------------------------------
# Serving workload: periodic water boxes at liquid density, sizes log-uniform.
# ---------------------------------------------------------------------------
WATER_DENSITY_NM3 = 33.43


def water_spec(key, n_waters, seed, rc=8.0):
    """A polarisable water box of n_waters at liquid density with cutoff rc.
    Boxes shorter than 2 rc keep rc: the pair set is then the reference's
    single minimum image, each pair once at its rounded-fractional image
    (dmff/admp/spatial.py:13-32), and kappa follows rc as in pme.py:460-466."""
    from .water import make_water_system
    L_nm = (n_waters / WATER_DENSITY_NM3) ** (1.0 / 3.0)
    n_side = int(math.ceil(n_waters ** (1.0 / 3.0) - 1e-9))
    S = make_water_system(n_waters, L_nm, seed, polarizable=True, n_side=n_side)
    box = S.box_nm * 10.0
    spec = SystemSpec(key, S.n_atoms, S.axis_types, S.axis_idx, S.bonded, S.Q_local, S.pol, S.thole,
                      S.mScales, S.pScales, S.dScales, rc)
    return spec, S.positions_nm * 10.0, box


def make_stream(n_requests, qps, seed, n_min=64, n_max=8192, follow_frac=0.3, sigma_A=0.02):
    """Poisson arrivals; new systems with log-uniform size, follow-on frames of a
    live system displaced by N(0, sigma) per atom."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    systems = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / qps)
        if systems and rng.uniform() < follow_frac:
            spec, pos0, box = systems[rng.integers(len(systems))]
            pos = pos0 + rng.normal(0.0, sigma_A, pos0.shape)
        else:
            nw = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
            spec, pos, box = water_spec(f"sys{len(systems)}", nw, int(rng.integers(1 << 30)))
            systems.append((spec, pos, box))
        out.append((t, spec, pos, box))
    return out
