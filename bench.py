#!/usr/bin/env python
"""ADMP multipolar PME benchmark on one MI355X.

Workload (default): 4096 rigid polarisable waters, 12,288 atoms, L = 4.967 nm,
rc 0.8 nm, ethresh 5e-4 (kappa 0.3285 1/A, 50^3 grid), order-6 PME, lmax 2,
Thole a = 0.39, induced dipoles solved to max|dE/dU| < 1e-6 from a cold start.

One evaluation ("step") = GPU neighbour build + induced-dipole solve + energy
+ dE/dpos + dE/dbox + dE/d{Q_local, pol, thole} (per atom, summed per type),
with inputs resident in HBM.  Timed: K steps between two device synchronisations.
`value` is the mixed-precision throughput (evals/s); the FP64 path is timed the
same way and reported beside it.

The CPU baseline (reported only) times the FP64 oracle on the host cores on a
bounded sample: one full energy+gradient pass and one induced-dipole gradient
pass at the same 12,288-atom size, scaled to one evaluation with the GPU
solver's iteration count.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(cfg_id):
    import numpy as np
    import torch
    from dmff_amd.admp.water import config_system
    from dmff_amd.admp.pme import setup_ewald_parameters
    S, c = config_system(cfg_id)
    box = S.box_nm * 10.0
    rc = c["rc"] * 10.0
    kappa, K1, K2, K3 = setup_ewald_parameters(rc, 5e-4, box)
    dev = "cuda:0"
    t = lambda x, dt=torch.float64: torch.as_tensor(np.asarray(x), dtype=dt, device=dev)
    D = dict(S=S, rc=rc, kappa=kappa, K=(K1, K2, K3), pos=t(S.positions_nm * 10.0), box=t(box),
             Q=t(S.Q_local), pol=t(S.pol), thole=t(S.thole), map=t(S.map_atomtype, torch.int64), dev=dev)
    from dmff_amd.admp.pme import TypeReducer
    D["qtype"] = TypeReducer(S.map_atomtype, device=dev)
    D["ptype"] = TypeReducer(S.map_poltype, device=dev)
    return D


def make_engine(D, precision):
    from dmff_amd.admp.engine import Engine
    S = D["S"]
    e = Engine(S.n_atoms, 2, S.polarizable, True, D["kappa"], D["K"], D["rc"], precision=precision, device=D["dev"])
    e.set_topology(S.axis_types, S.axis_idx, S.bonded)
    return e


def evaluate(eng, D, tol=1e-6, stats=True):
    import torch
    S = D["S"]
    eng.build_neighbors(D["pos"], D["box"])
    E, g, U, info = eng.compute(D["pos"], D["box"], D["Q"], D["pol"], D["thole"], S.mScales, S.pScales, S.dScales,
                                solver="pcg", tol=tol, maxiter=500, want_grad=True, stats=stats)
    dQ = D["qtype"](g["Q"][0])
    # pol and thole share the polarisation-type map: one reduction of both columns
    dpt = D["ptype"](torch.stack((g["pol"][0], g["thole"][0]), dim=1))
    dpol, dth = dpt[:, 0], dpt[:, 1]
    return E, g, U, info, dQ, dpol, dth


def time_path(eng, D, steps, warmup):
    """Seconds per evaluation over `steps` timed evaluations (no instrumentation in
    the timed region), then the per-family breakdown from as many evaluations with
    the library's event timing on (one event pair per launch, which costs time)."""
    import torch
    for _ in range(warmup):
        evaluate(eng, D, stats=False)
    torch.cuda.synchronize()
    # the timed evaluations skip the solver statistics (iteration counts, energy
    # components: a host read-back); the same evaluation with them follows
    t0 = time.perf_counter()
    for _ in range(steps):
        evaluate(eng, D, stats=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    out = evaluate(eng, D)
    eng.profile(True, True)
    for _ in range(steps):
        evaluate(eng, D)
    torch.cuda.synchronize()
    prof = eng.profile_get()
    eng.profile(False, True)
    return (t1 - t0) / steps, out, prof


def kernel_profile(D, precision, steps=2):
    """Per-kernel average launch durations with HIP events (graph capture off so each
    launch is bracketed individually)."""
    os.environ["ADMP_NO_GRAPH"] = "1"
    try:
        eng = make_engine(D, precision)
    finally:
        os.environ.pop("ADMP_NO_GRAPH", None)
    evaluate(eng, D)
    import torch
    torch.cuda.synchronize()
    eng.profile(True, True)
    for _ in range(steps):
        evaluate(eng, D)
    torch.cuda.synchronize()
    prof = eng.profile_get()
    eng.profile(False, True)
    return prof


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_pairs(pos, box, rc, bonded):
    """Reference-format pairs (i < j, nbond) on the host: a periodic k-d tree
    (scipy cKDTree, the algorithmic class of the reference's freud AABB query,
    dmff/common/nblist.py:27-55) for candidates, then the exact minimum-image
    test d < rc."""
    import numpy as np
    from scipy.spatial import cKDTree
    L = np.diag(box)
    p = np.mod(pos, L)
    cand = cKDTree(p, boxsize=L).query_pairs(rc * (1 + 1e-9), output_type="ndarray")
    d = pos[cand[:, 0]] - pos[cand[:, 1]]
    d -= L * np.floor(d / L + 0.5)
    keep = np.einsum("ij,ij->i", d, d) < rc * rc
    cand = np.sort(cand[keep], axis=1)
    n = len(pos)
    nb = np.zeros(len(cand), dtype=np.int64)
    if bonded:
        bk = np.array([i * n + j for (i, j) in bonded], dtype=np.int64)
        bv = np.array(list(bonded.values()), dtype=np.int64)
        order = np.argsort(bk)
        bk, bv = bk[order], bv[order]
        key = cand[:, 0] * n + cand[:, 1]
        pos_ = np.clip(np.searchsorted(bk, key), 0, len(bk) - 1)
        hit = bk[pos_] == key
        nb[hit] = bv[pos_[hit]]
    return np.concatenate([cand, nb[:, None]], axis=1)


def cpu_baseline(D, n_iter_cg):
    """The FP64 oracle on the host cores, on a bounded sample of one evaluation at
    the full 12,288-atom size: the pair list (periodic k-d tree), three
    induced-dipole gradient passes (the unit of work of one CG iteration; their
    mean is used), and the final energy + gradient pass.  One evaluation =
    pairs + (n_cg + 1) dE/dU passes + final pass, where n_cg is the number of
    iterations of the oracle's own (diagonally preconditioned) CG to the same
    tolerance.  Also one dE/dU pass on a single core."""
    import numpy as np
    import torch
    from oracle import admp_oracle as O
    S = D["S"]
    pos = S.positions_nm * 10.0
    box = S.box_nm * 10.0
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    t0 = time.perf_counter()
    pairs = cpu_pairs(pos, box, D["rc"], S.bonded)
    t_pairs = time.perf_counter() - t0
    f = O.ADMPPmeForce(box, S.axis_types, S.axis_idx, D["rc"], 5e-4, 2, True, True, None)
    T = lambda x: torch.tensor(np.asarray(x, dtype=np.float64))
    U = torch.zeros((S.n_atoms, 3), dtype=torch.float64)
    rest = (T(S.pol), T(S.thole), S.mScales, S.pScales, S.dScales)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        f.grad_U(T(pos), T(box), pairs, T(S.Q_local), U, *rest)
        ts.append(time.perf_counter() - t0)
    t_iter = sum(ts) / len(ts)
    p = T(pos).requires_grad_(True)
    b = T(box).requires_grad_(True)
    q = T(S.Q_local).requires_grad_(True)
    a = T(S.pol).requires_grad_(True)
    th = T(S.thole).requires_grad_(True)
    t0 = time.perf_counter()
    e = f.energy_fn(p, b, pairs, q, U, a, th, S.mScales, S.pScales, S.dScales)
    torch.autograd.grad(e, (p, b, q, a, th))
    t_final = time.perf_counter() - t0
    torch.set_num_threads(1)
    t0 = time.perf_counter()
    f.grad_U(T(pos), T(box), pairs, T(S.Q_local), U, *rest)
    t_iter1 = time.perf_counter() - t0
    torch.set_num_threads(cores)
    t_eval = t_pairs + (n_iter_cg + 1) * t_iter + t_final
    return {"value": 1.0 / t_eval, "unit": "evals/s", "cores": cores, "kind": "port",
            "cpu_model": _cpu_model(),
            "single_core_dEdU_pass_s": t_iter1,
            "sample": (f"FP64 oracle (torch CPU, {cores} threads) at 12,288 atoms, measured: pair list (periodic "
                       f"k-d tree) {t_pairs:.2f} s, dE/dU pass {t_iter:.2f} s (mean of 3; {t_iter1:.1f} s on one "
                       f"core), final E+grad pass {t_final:.2f} s; one eval = pairs + ({n_iter_cg}+1) dE/dU passes "
                       f"(the oracle's diagonally preconditioned CG to 1e-6) + final pass = {t_eval:.1f} s")}


def leg_batch(D, precision, n_frames=64, steps=2, warmup=1, sigma_nm=0.005, profile=False):
    """64 perturbed frames of the 12,288-atom system in one batched call (trajectory fitting)."""
    import numpy as np
    import torch
    from dmff_amd.admp.engine import Engine
    S = D["S"]
    rng = np.random.default_rng(1002)
    base = S.positions_nm * 10.0
    frames = np.stack([base + rng.normal(0.0, sigma_nm * 10.0, base.shape) for _ in range(n_frames)])
    pos = torch.as_tensor(frames, dtype=torch.float64, device=D["dev"])
    box = torch.as_tensor(np.stack([S.box_nm * 10.0] * n_frames), dtype=torch.float64, device=D["dev"])
    eng = Engine(S.n_atoms, 2, True, True, D["kappa"], D["K"], D["rc"], n_frames=n_frames, precision=precision,
                 device=D["dev"])
    eng.set_topology(S.axis_types, S.axis_idx, S.bonded)
    N = S.n_atoms
    holder = [eng]

    def ev(stats=True):
        eng = holder[0]
        eng.build_neighbors(pos, box)
        E, g, U, info = eng.compute(pos, box, D["Q"], D["pol"], D["thole"], S.mScales, S.pScales, S.dScales,
                                    solver="pcg", tol=1e-6, maxiter=500, want_grad=True, stats=stats)
        # per frame, per type parameter gradients
        dQ = D["qtype"](g["Q"], axis=1)
        dpol = D["ptype"](g["pol"], axis=1)
        dth = D["ptype"](g["thole"], axis=1)
        return E, g, info, dQ, dpol, dth

    for _ in range(warmup):
        ev(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ev(False)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    info = ev()[2]   # the same evaluation with the solver statistics, untimed
    holder[0] = None
    del eng
    torch.cuda.empty_cache()
    res = {"frames": n_frames, "ms_per_batch": dt * 1e3, "evals_per_s": n_frames / dt,
           "solver_iters_max": int(max(info["n_iter"][:min(n_frames, 64)])),
           "residual_max": float(max(info["residual"][:min(n_frames, 64)]))}
    if profile:
        # per-kernel launch durations of one batched evaluation (graphs off, HIP events
        # around each launch on the stream it runs on); each launch covers all frames
        os.environ["ADMP_NO_GRAPH"] = "1"
        try:
            eng = Engine(S.n_atoms, 2, True, True, D["kappa"], D["K"], D["rc"], n_frames=n_frames,
                         precision=precision, device=D["dev"])
        finally:
            os.environ.pop("ADMP_NO_GRAPH", None)
        eng.set_topology(S.axis_types, S.axis_idx, S.bonded)
        holder[0] = eng
        ev()
        torch.cuda.synchronize()
        eng.profile(True, True)
        ev()
        torch.cuda.synchronize()
        prof = eng.profile_get()
        eng.profile(False, True)
        P = int(eng.get_pairs(0).shape[0])
        holder[0] = None
        del eng
        torch.cuda.empty_cache()
        Db = dict(D)
        Db["n_pairs"] = P
        res["kernel_us_per_launch"] = {k: v[0] / v[1] * 1e3 for k, v in prof.items() if v[1]}
        res["roofline"] = {k: roofline(k, prof[k], Db, precision, frames=n_frames, prof=prof)
                           for k in ("spread", "fft", "influence", "gather", "real_matvec") if prof.get(k, (0, 0))[1]}
        for r in res["roofline"].values():
            r["traffic"], r["traffic_source"] = measured_traffic(r["kernel"].split("/")[0], batch=True)
    return res


def water_dispersion_slater(S, seed):
    """Per-atom sqrt(C6, C8, C10) (A units) with +-10% jitter; per-type Slater A (sqrt kJ/mol, the pair
    prefactor is A_i A_j as in pairwise.py:156-169) and B (1/A)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    C = np.array([[2.0e-3, 1.0e-4, 6.0e-6], [1.0e-4, 3.0e-6, 1.0e-7]])  # kJ/mol nm^p, O and H
    t = S.map_atomtype
    Ca = C[t] * rng.uniform(0.9, 1.1, size=(S.n_atoms, 3))
    Ca = Ca * np.array([1e6, 1e8, 1e10])  # nm^p -> A^p
    A_t = np.sqrt(rng.uniform(1e4, 1e5, size=2))  # pair prefactor A_i A_j in [1e4, 1e5] kJ/mol
    B_t = rng.uniform(30.0, 45.0, size=2) / 10.0  # 1/nm -> 1/A
    return np.sqrt(Ca), A_t[t], B_t[t]


def leg_large(steps=2, warmup=1, precision="mixed"):
    """32,768 waters (98,304 atoms), rc 10 A: electrostatics + polarisation + dispersion PME (pmax 10) + Slater."""
    import numpy as np
    import torch
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import setup_ewald_parameters
    from dmff_amd.admp.water import config_system
    S, c = config_system(3)
    dev = "cuda:0"
    box_np = S.box_nm * 10.0
    rc = c["rc"] * 10.0
    kappa, K1, K2, K3 = setup_ewald_parameters(rc, 5e-4, box_np)
    cl, a, b = water_dispersion_slater(S, 1003)
    T = lambda x: torch.as_tensor(np.asarray(x), dtype=torch.float64, device=dev)
    pos, box = T(S.positions_nm * 10.0), T(box_np)
    Q, pol, th, cl, a, b = T(S.Q_local), T(S.pol), T(S.thole), T(cl), T(a), T(b)
    eng = Engine(S.n_atoms, 2, True, True, kappa, (K1, K2, K3), rc, precision=precision, disp_kappa=kappa,
                 dK=(K1, K2, K3), disp_pmax=10, device=dev)
    eng.set_topology(S.axis_types, S.axis_idx, S.bonded)

    def ev():
        eng.build_neighbors(pos, box)
        E, g, U, info = eng.compute(pos, box, Q, pol, th, S.mScales, S.pScales, S.dScales, solver="pcg", tol=1e-6,
                                    maxiter=500, want_grad=True)
        Ed, gd = eng.disp_compute(pos, box, cl, S.mScales)
        Es, gs = eng.slater_compute(pos, box, a, b, S.mScales)
        return E, Ed, Es, info

    for _ in range(warmup):
        ev()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        E, Ed, Es, info = ev()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    del eng
    torch.cuda.empty_cache()
    return {"n_atoms": S.n_atoms, "grid": [K1, K2, K3], "rc_A": rc, "precision": precision,
            "ms_per_eval": dt * 1e3, "evals_per_s": 1.0 / dt, "solver_iters": int(info["n_iter"][0]),
            "E_elec": float(E[0]), "E_disp_backend": float(Ed[0]), "E_slater": float(Es[0])}


def leg_serving(stream, precision, slots=4, max_batch=8, time_scale=1.0):
    """A request stream through the continuous-batching server: arrivals at the
    stream's times x time_scale; latency per request, throughput over the span."""
    import torch
    from dmff_amd.admp.serving import ADMPServer, Request, latency_summary
    srv = ADMPServer(device="cuda:0", precision=precision, slots=slots, tol=1e-6, max_batch=max_batch)
    futs = []
    t0 = time.perf_counter()
    for rid, (t, spec, pos, box) in enumerate(stream):
        t = t * time_scale
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        futs.append(srv.submit(Request(rid, spec, pos, box, arrival=t0 + t)))
    res = [f.result(timeout=600) for f in futs]
    span = time.perf_counter() - t0
    srv.close()
    del srv
    torch.cuda.empty_cache()
    out = latency_summary(res)
    out.update({"evals_per_s": len(res) / span, "offered_qps": len(res) / (stream[-1][0] * time_scale),
                "slots": slots, "max_batch": max_batch, "precision": precision,
                "warm_frames": int(sum(r.warm for r in res))})
    return out, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-json", default=None)
    ap.add_argument("--legs", default="batch,large,serving",
                    help="extra measurements beside the headline: batch (64 frames), large (98,304 atoms with "
                         "dispersion and Slater), serving (request stream, p50/p99); 'none' for the headline only")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 256 MiB device-to-device tensor copy five times (the reference "
                         "kernel for scaling FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (requested --gpus {args.gpus}); the ADMP path here is a "
              f"single-GPU build. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    if args.calibrate_copy:
        # counter calibration: a 256 MiB fill (WRITE_SIZE reference) and a 256 MiB
        # elementwise kernel copy, dst = src * 1 (FETCH_SIZE and WRITE_SIZE reference;
        # a plain copy_ runs as a DMA blit that no kernel counter sees)
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        log(f"calibration: fill and 3 x elementwise copy of {src.numel() * 8} bytes")
        del src, dst
    D = setup(args.config)
    S = D["S"]
    log(f"config {args.config}: {S.n_atoms} atoms, kappa {D['kappa']:.6f}, K {D['K']}")
    res = {}
    for prec in ("fp64", "mixed"):
        eng = make_engine(D, prec)
        dt, out, prof = time_path(eng, D, args.steps, args.warmup)
        if "n_pairs" not in D:
            D["n_pairs"] = int(eng.get_pairs(0).shape[0])
        E, g, U, info = out[:4]
        res[prec] = dict(dt=dt, E=float(E[0]), g=g, info=info, prof=prof, dQ=out[4])
        log(f"{prec}: {dt * 1e3:.2f} ms/eval, E={float(E[0]):.8f}, iters={info['n_iter'][0]}, "
            f"res={info['residual'][0]:.2e}")
        for k, (ms, n) in sorted(prof.items(), key=lambda x: -x[1][0]):
            if n:
                log(f"   {k:16s} {ms / args.steps:9.3f} ms/eval  {n // args.steps:5d} launches/eval  "
                    f"{ms / n * 1e3:9.1f} us/launch")
    kprof = kernel_profile(D, "mixed")
    log("per-kernel (events, graphs off, mixed):")
    for k, (ms, n) in sorted(kprof.items(), key=lambda x: -x[1][0]):
        if n:
            log(f"   {k:16s} {ms / 2:9.3f} ms/eval  {n // 2:5d} launches/eval  {ms / n * 1e3:9.1f} us/launch")
    f64, mix = res["fp64"], res["mixed"]
    e_rel = abs(mix["E"] - f64["E"]) / abs(f64["E"])
    fr = float(torch.linalg.norm(mix["g"]["pos"] - f64["g"]["pos"]) / torch.linalg.norm(f64["g"]["pos"]))
    # dominant kernel of the mixed path and its roofline.  The event families
    # lump kernels ("spread": the solver's, the right-hand side's and the final
    # pass's instances; "fft": the z R2C and C2R passes), so the dominant single
    # kernel is taken from the kernel trace of the timed evaluations
    # (DOMINANT_KERNEL_SOURCE): the fused plane pass + real-space product when
    # this size runs it; the spread's roofline is reported beside it.
    prof = mix["prof"]
    cand = {k: v for k, v in kprof.items() if v[1] and k not in ("solver_vec", "solver_graph")}
    dom = ("plane_matvec", kprof["plane_matvec"]) if kprof.get("plane_matvec", (0, 0))[1] else \
        max(cand.items(), key=lambda x: x[1][0])
    roof = roofline(dom[0], dom[1], D, "mixed", prof=kprof)
    roof["traffic"], roof["traffic_source"] = measured_traffic(roof["kernel"])
    if dom[0] == "plane_matvec":
        roof["dominance"] = DOMINANT_KERNEL_SOURCE
    if dom[0] != "spread" and kprof.get("spread", (0, 0))[1]:
        sp = roofline("spread", kprof["spread"], D, "mixed", prof=kprof)
        sp["traffic"], sp["traffic_source"] = measured_traffic(sp["kernel"])
        roof["spread"] = sp
    rf = kprof.get("real_final", (0, 0))
    roof["pair_kernel_valu"] = pair_kernel_valu(rf[0] / rf[1] * 1e3 if rf[1] else None)
    out = {
        "metric": "ADMP E+F+grad evals/s @12,288 atoms (mixed precision); FP64 beside it",
        "value": 1.0 / mix["dt"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mix["dt"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair/spread/FFT with f64 accumulation, f64 solver residual",
        "data": "synthetic (4096 rigid waters, seeded recipe)",
        "config": {"workload": f"ADMP PME cfg{args.config}: {S.n_atoms} atoms, lmax 2, Thole-polarisable, "
                               f"rc 8 A, grid {D['K'][0]}^3, solve to 1e-6",
                   "n_atoms": S.n_atoms, "grid": list(D["K"]), "rc_A": D["rc"]},
        "fp64": {"value": 1.0 / f64["dt"], "ms_per_step": f64["dt"] * 1e3, "solver_iters": f64["info"]["n_iter"][0],
                 "residual": f64["info"]["residual"][0]},
        "mixed_vs_fp64": {"speedup": f64["dt"] / mix["dt"], "energy_rel": e_rel, "force_rel_rms": fr,
                          "solver_iters": mix["info"]["n_iter"][0], "residual": mix["info"]["residual"][0]},
        "roofline": roof,
        "kernels_ms_per_eval": {k: v[0] / args.steps for k, v in prof.items() if v[1]},
        "kernel_us_per_launch": {k: v[0] / v[1] * 1e3 for k, v in kprof.items() if v[1]},
    }
    legs = [x for x in args.legs.split(",") if x and x != "none"]
    if "batch" in legs:
        b = {}
        for prec in ("mixed", "fp64"):
            b[prec] = leg_batch(D, prec, profile=(prec == "mixed"))
            log(f"batch64 {prec}: {b[prec]['ms_per_batch']:.1f} ms/batch, {b[prec]['evals_per_s']:.1f} evals/s")
        out["batch64"] = b
    if "large" in legs:
        out["large"] = leg_large()
        log(f"large: {out['large']['ms_per_eval']:.1f} ms/eval ({out['large']['n_atoms']} atoms)")
    if "serving" in legs:
        from dmff_amd.admp.serving import make_fitting_stream, make_stream
        t_gen = time.perf_counter()
        stream = make_stream(500, 20.0, 1004)
        log(f"serving stream: 500 requests generated in {time.perf_counter() - t_gen:.1f} s")
        sv = {}
        res = {}
        for prec in ("mixed", "fp64"):
            sv[prec], res[prec] = leg_serving(stream, prec)
            log(f"serving {prec}: p50 {sv[prec]['p50_ms']:.1f} ms, p99 {sv[prec]['p99_ms']:.1f} ms, "
                f"{sv[prec]['evals_per_s']:.1f} evals/s, {sv[prec].get('ctx_created', 0)} contexts created "
                f"(mean {sv[prec].get('ctx_ms_mean', 0.0):.1f} ms)")
        import torch as _t
        errs = [float(_t.linalg.norm(a.dpos - f.dpos) / _t.linalg.norm(f.dpos))
                for a, f in zip(res["mixed"], res["fp64"])]
        sv["mixed_force_rel_rms_max"] = max(errs)
        sv["stream"] = ("500 requests, Poisson arrivals at 20/s, water boxes of 64-8192 molecules (log-uniform), "
                        "30% follow-on frames of a live system, 4 slots, batches of up to 8 frames of one class, "
                        "solve to 1e-6")
        # the same stream offered faster, up to saturation (mixed)
        sweep = []
        for qps in (80.0, 320.0, 1280.0):
            r, _ = leg_serving(stream, "mixed", time_scale=20.0 / qps)
            sweep.append({k: r[k] for k in ("offered_qps", "evals_per_s", "p50_ms", "p99_ms", "mean_batch")})
            log(f"serving mixed at {qps:.0f}/s offered: {r['evals_per_s']:.1f} evals/s, p50 {r['p50_ms']:.1f} ms, "
                f"p99 {r['p99_ms']:.1f} ms")
        sv["qps_sweep_mixed"] = sweep
        # trajectory-fitting pattern: frames of 8 same-size boxes (one class) batch into multi-frame launches
        fit = []
        for qps in (100.0, 400.0, 1600.0, 6400.0):
            r, _ = leg_serving(make_fitting_stream(300, qps, 5), "mixed")
            fit.append({k: r[k] for k in ("offered_qps", "evals_per_s", "p50_ms", "p99_ms", "mean_batch")})
            log(f"fitting stream at {qps:.0f}/s offered: {r['evals_per_s']:.1f} evals/s, mean batch "
                f"{r['mean_batch']:.2f}, p99 {r['p99_ms']:.1f} ms")
        r1, _ = leg_serving(make_fitting_stream(300, 6400.0, 5), "mixed", max_batch=1)
        log(f"fitting stream at 6400/s, no batching: {r1['evals_per_s']:.1f} evals/s")
        sv["fitting_sweep_mixed"] = fit
        sv["fitting_unbatched_6400qps"] = {k: r1[k] for k in ("evals_per_s", "p50_ms", "p99_ms", "mean_batch")}
        sv["fitting_stream"] = "300 frames of 8 boxes of 512 waters (one force field), sigma 0.02 A, Poisson arrivals"
        out["serving"] = sv
    if not args.no_cpu:
        # the oracle's CG is preconditioned by the diagonal: its iteration count to the
        # same tolerance is the library's with the diagonal preconditioner
        eng = make_engine(D, "fp64")
        eng.build_neighbors(D["pos"], D["box"])
        _, _, _, dinfo = eng.compute(D["pos"], D["box"], D["Q"], D["pol"], D["thole"], S.mScales, S.pScales,
                                     S.dScales, solver="pcg", tol=1e-6, maxiter=500, want_grad=False,
                                     precond="diagonal")
        del eng
        out["cpu_baseline"] = cpu_baseline(D, dinfo["n_iter"][0])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


COUNTER_SUMMARY = os.path.join("profiles", "admp_12288_hbm_counters.md")
BATCH_COUNTER_SUMMARY = os.path.join("profiles", "admp_12288x64_batch_hbm_counters.md")
VALU_SUMMARY = os.path.join("profiles", "admp_12288_pair_kernel_valu.md")
FP64_VALU_PEAK_TFLOPS, FP32_VALU_PEAK_TFLOPS = 78.6, 157.3


def _summary_rows(rel):
    """Table rows (lists of cells) of a committed profile summary; [] when it is absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), rel)
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return []
    return [[c.strip() for c in ln.strip().strip("|").split("|")] for ln in lines if ln.startswith("| `")]


def measured_traffic(kname, batch=False):
    """HBM bytes per launch of kernel `kname` at this workload, FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (the factors
    read off a 256 MiB copy), from the committed summary of two rocprofv3 --pmc runs of this script, one counter
    each (single frame: `--legs none --steps 1 --warmup 0 --calibrate-copy`; 64-frame batch: `--legs batch ...`).
    (None, None) without a row for it."""
    rel, ncol, col = (BATCH_COUNTER_SUMMARY, 9, 3) if batch else (COUNTER_SUMMARY, 7, 4)
    for cells in _summary_rows(rel):
        if len(cells) == ncol and cells[0].startswith("`" + kname + "<"):
            try:
                return int(cells[col]) * 1024, rel + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"
            except ValueError:
                return None, None
    return None, None


def pair_kernel_valu(us_mixed):
    """VALU FLOP rate of the real-space final pair kernel: lane FLOPs per launch (64 x the SQ_INSTS_VALU_FLOPS_*
    counters, committed summary of two rocprofv3 --pmc runs of `--legs none`) over the launch duration, the
    mixed path's measured live (HIP events, graphs off), the FP64 path's from the summary's kernel trace."""
    out = {}
    for cells in _summary_rows(VALU_SUMMARY):
        if len(cells) != 11:
            continue
        try:
            g64, g32, us = float(cells[6]), float(cells[7]), float(cells[8])
        except ValueError:
            continue
        if cells[0].startswith("`k_real_final<float"):
            t = us_mixed if us_mixed else us
            ach = g32 * 1e9 / (t * 1e-6) / 1e12
            out = {"kernel": "k_real_final<float, true>", "bound": "valu", "achieved": ach,
                   "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP32_VALU_PEAK_TFLOPS,
                   "frac_of_unpacked_fp32": ach / FP64_VALU_PEAK_TFLOPS, "fp32_flop_per_launch": g32 * 1e9,
                   "fp64_flop_per_launch": g64 * 1e9, "avg_launch_us": t, "source": VALU_SUMMARY, **out}
        elif cells[0].startswith("`k_real_final<double"):
            ach = g64 * 1e9 / (us * 1e-6) / 1e12
            out["fp64_path"] = {"kernel": "k_real_final<double, true>", "achieved": ach,
                                "peak": FP64_VALU_PEAK_TFLOPS, "frac": ach / FP64_VALU_PEAK_TFLOPS,
                                "fp64_flop_per_launch": g64 * 1e9, "avg_launch_us": us}
    return out or None


DOMINANT_KERNEL_SOURCE = ("k_plane_matvec<double>: 19.6 % of the kernel time of the timed evaluations, ahead of "
                          "k_spread_cw<double, 1, 48> (16.0 %), profiles/admp_12288_kernel_stats.md")


def roofline(name, tm, D, prec, frames=1, prof=None):
    """Algorithmic bytes per launch of a kernel family / its measured average launch duration.

    Per-launch algorithmic bytes per frame (w = 8 B: solver products are FP64 on both paths; G = K0 K1 K2,
    Gh = K0 K1 (K2/2 + 1); solver launches, which dominate, spread and gather dipoles only):
      spread     N (3*8 + 4 w) + G w           positions, fractional charge+dipole, grid written once
      gather     G w + N 3*8 + N 3*8           grid, positions, field out
      real_matvec 2P (4 + 2 w) + N 3 (w + 8)   ELL index + two stored coefficients per ordered pair
    FFT, fused plane path (two z launches per transform pair, one plane launch):
      fft        G w + 2 Gh w                  z R2C or C2R: real grid one way, half spectrum the other
      influence  4 Gh w + 8 Gh                 plane pass: half spectrum in and out, tabulated influence
    FFT, line path (four z/y launches per transform pair, one x launch):
      fft        (2 G w + 12 Gh w) / 4         z R2C, y, y, z C2R passes
      influence  4 Gh w + 8 Gh                 x pass: half spectrum in and out, tabulated influence
    plane_matvec (solver products with few frames: the plane pass and the real-space product in one launch)
                 4 Gh w + 8 Gh + 2P (4 + 2 w) + N 3 (w + 8)
    The path is read off the launch counts: 2 or 4 fft launches per plane (influence or plane_matvec) launch.
    """
    ms, n = tm
    t = ms / n / 1e3
    S = D["S"]
    N = S.n_atoms
    # the solver launches, which dominate every family, run FP64 products on both paths
    w = 8
    K = D["K"]
    G = K[0] * K[1] * K[2]
    Gh = K[0] * K[1] * (K[2] // 2 + 1)
    P = D.get("n_pairs", 0)
    plane = False
    n_plane = (prof.get("influence", (0, 0))[1] + prof.get("plane_matvec", (0, 0))[1]) if prof else 0
    if n_plane:
        plane = prof.get("fft", (0, 0))[1] < 3 * n_plane
    table = {
        "spread": ("k_spread_cw", N * (3 * 8 + 4 * w) + G * w),
        "fft": ("k_fz_r2c/k_fz_c2r", G * w + 2 * Gh * w) if plane else
               ("k_fz_r2c/k_fy/k_fz_c2r", (2 * G * w + 12 * Gh * w) / 4),
        "influence": ("k_fxy_plane" if plane else "k_fx_infl", 4 * Gh * w + 8 * Gh),
        "gather": ("k_gather_ml", G * w + N * 3 * 8 + N * 3 * 8),
        "real_matvec": ("k_real_matvec", 2 * P * (4 + 2 * 8) + N * 3 * (8 + 8)),
        "plane_matvec": ("k_plane_matvec", 4 * Gh * w + 8 * Gh + 2 * P * (4 + 2 * 8) + N * 3 * (8 + 8)),
    }
    if name not in table or (name in ("real_matvec", "plane_matvec") and not P):
        return {"kernel": name, "bound": None, "achieved": None, "peak": None, "unit": None, "frac": None,
                "traffic": None, "avg_launch_us": t * 1e6}
    kname, bytes_ = table[name]
    bytes_ *= frames
    ach = bytes_ / t / 1e9
    return {"kernel": kname, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": None, "avg_launch_us": t * 1e6,
            "algorithmic_bytes_per_launch": bytes_}


if __name__ == "__main__":
    main()
