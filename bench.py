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


def evaluate(eng, D, tol=1e-6):
    import torch
    S = D["S"]
    eng.build_neighbors(D["pos"], D["box"])
    E, g, U, info = eng.compute(D["pos"], D["box"], D["Q"], D["pol"], D["thole"], S.mScales, S.pScales, S.dScales,
                                solver="pcg", tol=tol, maxiter=500, want_grad=True)
    dQ = D["qtype"](g["Q"][0])
    dpol = D["ptype"](g["pol"][0])
    dth = D["ptype"](g["thole"][0])
    return E, g, U, info, dQ, dpol, dth


def time_path(eng, D, steps, warmup):
    import torch
    for _ in range(warmup):
        out = evaluate(eng, D)
    torch.cuda.synchronize()
    eng.profile(True, True)
    t0 = time.perf_counter()
    for _ in range(steps):
        out = evaluate(eng, D)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
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


def cpu_baseline(D, n_iter):
    """FP64 oracle on the host cores: one E+grad pass and one dE/dU pass at full size."""
    import numpy as np
    import torch
    from oracle import admp_oracle as O
    S = D["S"]
    pos = S.positions_nm * 10.0
    box = S.box_nm * 10.0
    from tests.systems import reference_pairs
    t0 = time.perf_counter()
    pairs = reference_pairs(pos, box, D["rc"], S.bonded)
    t_pairs = time.perf_counter() - t0
    f = O.ADMPPmeForce(box, S.axis_types, S.axis_idx, D["rc"], 5e-4, 2, True, True, None)
    T = lambda x: torch.tensor(np.asarray(x, dtype=np.float64))
    U = torch.zeros((S.n_atoms, 3), dtype=torch.float64)
    rest = (T(S.pol), T(S.thole), S.mScales, S.pScales, S.dScales)
    t0 = time.perf_counter()
    f.grad_U(T(pos), T(box), pairs, T(S.Q_local), U, *rest)
    t_iter = time.perf_counter() - t0
    p = T(pos).requires_grad_(True)
    b = T(box).requires_grad_(True)
    q = T(S.Q_local).requires_grad_(True)
    a = T(S.pol).requires_grad_(True)
    th = T(S.thole).requires_grad_(True)
    t0 = time.perf_counter()
    e = f.energy_fn(p, b, pairs, q, U, a, th, S.mScales, S.pScales, S.dScales)
    torch.autograd.grad(e, (p, b, q, a, th))
    t_final = time.perf_counter() - t0
    t_eval = t_pairs + (n_iter + 1) * t_iter + t_final
    return {"value": 1.0 / t_eval, "unit": "evals/s", "cores": torch.get_num_threads(), "kind": "port",
            "sample": (f"FP64 oracle (torch CPU) at 12,288 atoms: brute-force pair list {t_pairs:.1f} s, one "
                       f"dE/dU pass {t_iter:.1f} s, one E+grad pass {t_final:.1f} s; one eval = pairs + "
                       f"({n_iter}+1) dE/dU passes + E+grad pass = {t_eval:.1f} s")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-json", default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (requested --gpus {args.gpus}); the ADMP path here is a "
              f"single-GPU build. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    D = setup(args.config)
    S = D["S"]
    log(f"config {args.config}: {S.n_atoms} atoms, kappa {D['kappa']:.6f}, K {D['K']}")
    res = {}
    for prec in ("fp64", "mixed"):
        eng = make_engine(D, prec)
        dt, out, prof = time_path(eng, D, args.steps, args.warmup)
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
    # dominant kernel family of the mixed path and its roofline
    prof = mix["prof"]
    cand = {k: v for k, v in kprof.items() if v[1] and k not in ("solver_vec", "solver_graph")}
    dom = max(cand.items(), key=lambda x: x[1][0])
    roof = roofline(dom[0], dom[1], D, "mixed")
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
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(D, f64["info"]["n_iter"][0])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def roofline(name, tm, D, prec):
    """Algorithmic work per launch / measured average launch duration."""
    ms, n = tm
    t = ms / n / 1e3
    S = D["S"]
    N = S.n_atoms
    w = 4 if prec == "mixed" else 8
    K = D["K"]
    G = K[0] * K[1] * K[2]
    if name in ("fft",):
        # one pass over the grid: read + write G complex values
        bytes_ = 2 * G * 2 * w
        return {"kernel": "k_fft_pass", "bound": "hbm", "achieved": bytes_ / t / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": bytes_ / t / 1e9 / HBM_PEAK_GBS, "traffic": None}
    if name in ("spread",):
        bytes_ = N * (3 * 8 + 10 * w) + G * w
        return {"kernel": "k_spread_atomic", "bound": "hbm", "achieved": bytes_ / t / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": bytes_ / t / 1e9 / HBM_PEAK_GBS, "traffic": None}
    if name in ("gather",):
        bytes_ = G * w + N * (3 * 8 + 10 * w) + N * 3 * 8
        return {"kernel": "k_gather", "bound": "hbm", "achieved": bytes_ / t / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": bytes_ / t / 1e9 / HBM_PEAK_GBS, "traffic": None}
    return {"kernel": name, "bound": "unknown", "achieved": None, "peak": None, "unit": None, "frac": None,
            "traffic": None, "avg_launch_us": t * 1e6}


if __name__ == "__main__":
    main()
