"""Continuous-batching request loop for ADMP PME evaluations on one GPU.

A request is one configuration (positions, box) of a periodic system with its
force-field data.  The server keeps every live system's induced dipoles on the
device as the warm start of its next frame, and keeps native contexts (PME
grids, FFT tables, neighbour lists, solver scratch, captured solver graphs)
alive between requests.  Requests run concurrently in ``slots`` worker threads,
each on its own stream; the library releases the interpreter lock inside every
native call, so the work of different requests overlaps on the GPU.

Batching: systems with the same force-field data, size and PME grid (one
"class": same atom count, topology, parameters, cutoff, kappa and K) share
native contexts, and up to ``max_batch`` waiting requests of one class --
one frame per system -- run as the frames of ONE multi-frame launch sequence
(every kernel of the evaluation, the solver's captured graphs included, covers
all frames at once).  Requests of different classes never share a launch: a
different K or atom count would have to be padded to a common grid, which
changes the PME discretisation and so the result.

Admission: requests wait in arrival order; whenever a slot is free the
scheduler takes the first waiting request whose system is idle, gathers the
later waiting requests of its class (systems idle, one frame each, none
overtaking an earlier request of its own system), and admits the batch when a
context for that class and frame count exists and is idle or its HBM footprint
fits the budget (evicting idle contexts, least recently used first).  Frames of
one system keep their order, so the warm-start chain of a system is the same
alone and in a mix; within a batch each frame's solve is independent, so
results agree with a lone run to the solver tolerance (to rounding when the
batch holds one frame).

Per-request latency (arrival -> result ready), queueing time, batch size and
context-creation time are recorded; ``latency_summary`` gives p50/p99.
"""
import collections
import concurrent.futures
import hashlib
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
        self._class = None

    def class_key(self):
        """Digest of everything a shared context depends on besides the box."""
        if self._class is None:
            h = hashlib.sha1(repr((self.n_atoms, self.rc, self.lmax, self.lpol, self.ethresh)).encode())
            for a in (self.axis_types, self.axis_idx, self.Q_local, self.pol, self.thole, *self.scales):
                if a is not None:
                    h.update(np.ascontiguousarray(np.asarray(a, dtype=np.float64)).tobytes())
            if self.exclusions:
                h.update(repr(sorted(self.exclusions.items())).encode())
            self._class = h.hexdigest()
        return self._class


class Request:
    def __init__(self, rid, spec, pos, box, arrival=None):
        self.rid = rid
        self.spec = spec
        self.pos = pos            # (N, 3) Angstrom
        self.box = box            # (3, 3) Angstrom, lattice vectors in rows
        self.arrival = time.perf_counter() if arrival is None else arrival
        self.future = concurrent.futures.Future()
        self._bkey = None


class Result:
    def __init__(self, rid, key, energy, dpos, dbox, n_iter, residual, warm, t_arrival, t_start, t_end,
                 batch=1, ctx_ms=0.0):
        self.rid, self.key = rid, key
        self.energy, self.dpos, self.dbox = energy, dpos, dbox
        self.n_iter, self.residual, self.warm = n_iter, residual, warm
        self.queued_ms = (t_start - t_arrival) * 1e3
        self.latency_ms = (t_end - t_arrival) * 1e3
        self.batch = int(batch)          # frames in the launch that served this request
        self.ctx_ms = float(ctx_ms)      # native context creation inside this request's latency


class _Live:
    """Per-system state: the warm start of the next frame."""

    def __init__(self, spec):
        self.spec = spec
        self.U = None
        self.busy = False
        self.last_used = time.perf_counter()


class _Ctx:
    """A native context for one class and frame count."""

    def __init__(self, key, nbytes, kappa, K):
        self.key, self.nbytes, self.kappa, self.K = key, nbytes, kappa, K
        self.engine = None
        self.dev = {}
        self.busy = False
        self.last_used = time.perf_counter()


def estimate_bytes(n_atoms, K, rc, precision="mixed", density=0.1003):
    """Upper estimate of a one-frame context's device footprint before it is created."""
    w = 4 if precision == "mixed" else 8
    cap = int(1.5 * density * 4.18879 * rc ** 3) + 64
    G = K[0] * K[1] * K[2]
    Gh = K[0] * K[1] * (K[2] // 2 + 1)
    per_site = 2400 + 60 * 8 + cap * (4 + 2 * w + 4)
    grids = G * (8 + 4 * 4) + Gh * 2 * 8 + 2 * G * 8
    return int(1.25 * (n_atoms * per_site + grids)) + (4 << 20)


class ADMPServer:
    def __init__(self, device="cuda:0", precision="mixed", slots=4, hbm_budget=None, tol=1e-6, maxiter=500,
                 max_batch=8, runner=None, max_systems=4096):
        """runner(reqs, lives, t_start) -> [Result] replaces the GPU evaluation (host-logic tests)."""
        self.device = torch.device(device)
        self.runner = runner
        self.precision, self.tol, self.maxiter = precision, float(tol), int(maxiter)
        self.max_batch, self.max_systems = max(1, int(max_batch)), int(max_systems)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.5 * free)
        self.budget = int(hbm_budget)
        self.used = 0
        self.systems = collections.OrderedDict()   # key -> _Live, least recently used first
        self.ctxs = {}                              # (class key, K, frames) -> _Ctx
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
    def _batch_key(self, req):
        """(class, K) of a request, with its one-frame footprint and kappa."""
        if req._bkey is None:
            spec = req.spec
            kappa, K1, K2, K3 = setup_ewald_parameters(spec.rc, spec.ethresh, np.asarray(req.box))
            K = (K1, K2, K3)
            req._bkey = ((spec.class_key(), K), estimate_bytes(spec.n_atoms, K, spec.rc, self.precision), kappa)
        return req._bkey

    def _evict_for_locked(self, need):
        idle = sorted((c for c in self.ctxs.values() if not c.busy), key=lambda c: c.last_used)
        for c in idle:
            if self.used + need <= self.budget:
                break
            self.used -= c.nbytes
            del self.ctxs[c.key]
            c.engine = None
        return self.used + need <= self.budget

    def _live_locked(self, spec):
        lv = self.systems.get(spec.key)
        if lv is None:
            lv = _Live(spec)
            self.systems[spec.key] = lv
            while len(self.systems) > self.max_systems:
                old = next(iter(self.systems.values()))
                if old.busy:
                    break
                del self.systems[old.spec.key]
        self.systems.move_to_end(spec.key)
        return lv

    def _admit_locked(self):
        blocked = set()
        i = 0
        while self.free_slots > 0 and i < len(self.pending):
            req = self.pending[i]
            key = req.spec.key
            lv = self.systems.get(key)
            if key in blocked or (lv is not None and lv.busy):
                blocked.add(key)   # keep the frame order of a system
                i += 1
                continue
            bkey, need1, kappa = self._batch_key(req)
            batch, members, skip = [i], {key}, set(blocked)
            j = i + 1
            while len(batch) < self.max_batch and j < len(self.pending):
                r2 = self.pending[j]
                k2 = r2.spec.key
                lv2 = self.systems.get(k2)
                if k2 not in members and k2 not in skip and not (lv2 is not None and lv2.busy) \
                        and self._batch_key(r2)[0] == bkey:
                    batch.append(j)
                    members.add(k2)
                else:
                    skip.add(k2)
                j += 1
            ck = (bkey[0], bkey[1], len(batch))
            ctx = self.ctxs.get(ck)
            if ctx is not None and ctx.busy:
                blocked.add(key)
                i += 1
                continue
            if ctx is None:
                need = need1 * len(batch)
                if self.used + need > self.budget and not self._evict_for_locked(need):
                    blocked.add(key)
                    i += 1
                    continue
                ctx = _Ctx(ck, need, kappa, bkey[1])
                self.ctxs[ck] = ctx
                self.used += need
            ctx.busy = True
            reqs = [self.pending[k] for k in batch]
            for k in reversed(batch):
                del self.pending[k]
            lives = [self._live_locked(r.spec) for r in reqs]
            for x in lives:
                x.busy = True
            self.free_slots -= 1
            self.work.append((reqs, lives, ctx))
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
                reqs, lives, ctx = self.work.popleft()
            t_start = time.perf_counter()
            try:
                if self.runner is not None:
                    res = self.runner(reqs, lives, t_start)
                else:
                    with torch.cuda.device(self.device), torch.cuda.stream(stream):
                        res = self._run(reqs, lives, ctx, t_start)
                        stream.synchronize()
                for r, x in zip(reqs, res):
                    r.future.set_result(x)
            except Exception as e:  # noqa: BLE001 - reported through the futures
                for r in reqs:
                    if not r.future.done():
                        r.future.set_exception(e)
            with self.cv:
                now = time.perf_counter()
                ctx.busy = False
                ctx.last_used = now
                for x in lives:
                    x.busy = False
                    x.last_used = now
                self.free_slots += 1
                for r in reqs:
                    if r.future.done() and not r.future.exception():
                        self.results.append(r.future.result())
                self._admit_locked()
                self.cv.notify_all()

    def _run(self, reqs, lives, ctx, t_start):
        spec = reqs[0].spec
        F, N = len(reqs), spec.n_atoms
        t_ctx = 0.0
        if ctx.engine is None:
            t0 = time.perf_counter()
            eng = Engine(N, spec.lmax, spec.lpol, True, ctx.kappa, ctx.K, spec.rc, n_frames=F,
                         precision=self.precision, device=self.device)
            eng.set_topology(spec.axis_types, spec.axis_idx, spec.exclusions)
            # the context works on the worker's stream: with a stream per context,
            # hundreds of contexts share the GPU's few hardware queues and wait on
            # each other's unrelated work
            eng.set_stream_mode(True)
            t = lambda x: torch.as_tensor(np.asarray(x), dtype=F64, device=self.device)
            ctx.dev = {"Q": t(spec.Q_local), "pol": t(spec.pol), "thole": t(spec.thole)}
            ctx.engine = eng
            nb = eng.memory_bytes()
            with self.cv:
                self.used += nb - ctx.nbytes
                ctx.nbytes = nb
            t_ctx = time.perf_counter() - t0
        eng = ctx.engine
        pos = torch.as_tensor(np.stack([np.asarray(r.pos) for r in reqs]), dtype=F64).to(self.device)
        box = torch.as_tensor(np.stack([np.asarray(r.box) for r in reqs]), dtype=F64).to(self.device)
        eng.build_neighbors(pos, box)
        warm = [x.U is not None for x in lives]
        U0 = None
        if any(warm):
            # a cold frame starts from zero dipoles, which is the library's cold start
            U0 = torch.stack([x.U if x.U is not None else torch.zeros((N, 3), dtype=F64, device=self.device)
                              for x in lives])
        E, g, U, info = eng.compute(pos, box, ctx.dev["Q"], ctx.dev["pol"], ctx.dev["thole"], *spec.scales,
                                    U_init=U0, solver="pcg", tol=self.tol, maxiter=self.maxiter, want_grad=True)
        for f, x in enumerate(lives):
            x.U = U[f].clone()
        t_end = time.perf_counter()
        Eh = E.cpu().tolist()
        return [Result(r.rid, r.spec.key, Eh[f], g["pos"][f], g["box"][f], info["n_iter"][f], info["residual"][f],
                       warm[f], r.arrival, t_start, t_end, batch=F, ctx_ms=t_ctx * 1e3)
                for f, r in enumerate(reqs)]


def latency_summary(results):
    lat = np.array([r.latency_ms for r in results])
    if len(lat) == 0:
        return {}
    out = {"n": int(len(lat)), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}
    ctx = np.array([getattr(r, "ctx_ms", 0.0) for r in results])
    if ctx.any():
        reuse = lat[ctx == 0.0]
        out.update({"ctx_created": int((ctx > 0).sum()), "ctx_ms_mean": float(ctx[ctx > 0].mean()),
                    "ctx_ms_p50": float(np.percentile(ctx[ctx > 0], 50))})
        if len(reuse):
            out.update({"reused_ctx_p50_ms": float(np.percentile(reuse, 50)),
                        "reused_ctx_p99_ms": float(np.percentile(reuse, 99))})
    b = np.array([getattr(r, "batch", 1) for r in results])
    out["mean_batch"] = float(b.mean())
    return out


# ---------------------------------------------------------------------------
# Serving workload: periodic water boxes at liquid density, sizes log-uniform.
# ---------------------------------------------------------------------------
WATER_DENSITY_NM3 = 33.43


def water_spec(key, n_waters, seed, rc=8.0):
    """A polarisable water box of n_waters at liquid density with cutoff rc,
    carrying configuration 1's multipoles (every box has the same force field;
    the seed draws the geometry).  Boxes shorter than 2 rc keep rc: the pair set
    is then the reference's single minimum image, each pair once at its
    rounded-fractional image (dmff/admp/spatial.py:13-32), and kappa follows rc
    as in pme.py:460-466."""
    from .water import config1_multipoles, make_water_system
    L_nm = (n_waters / WATER_DENSITY_NM3) ** (1.0 / 3.0)
    n_side = int(math.ceil(n_waters ** (1.0 / 3.0) - 1e-9))
    S = make_water_system(n_waters, L_nm, seed, polarizable=True, n_side=n_side,
                          Q_local_types=config1_multipoles())
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


def make_fitting_stream(n_requests, qps, seed, n_systems=8, n_waters=512, sigma_A=0.02):
    """Poisson arrivals of frames of n_systems same-size boxes (one force field,
    different geometries): the trajectory-fitting pattern, in which waiting
    frames of different systems share one class and run as one batch."""
    rng = np.random.default_rng(seed)
    systems = [water_spec(f"fit{k}", n_waters, int(rng.integers(1 << 30))) for k in range(n_systems)]
    t = 0.0
    out = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / qps)
        spec, pos0, box = systems[int(rng.integers(n_systems))]
        out.append((t, spec, pos0 + rng.normal(0.0, sigma_A, pos0.shape), box))
    return out
