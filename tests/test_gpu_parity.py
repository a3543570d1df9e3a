"""GPU parity: the HIP library (through the C ABI) against the FP64 oracle.

Tolerances (north-star bar for the FP64 path: energy 1e-6 relative, force
relative RMS 1e-5; the FP64 kernels land near 1e-10):
  FP64   energy rel 1e-9, gradient rel RMS 1e-7 (pos, box, Q_local, pol, thole)
  mixed  energy rel 2e-5, force rel RMS 1e-4
"""
import math

import numpy as np
import pytest
import torch

from oracle import admp_oracle as O
from tests.systems import load_system, reference_pairs, small_water_box as _small_waters

pytestmark = pytest.mark.gpu
F64 = torch.float64
DEV = "cuda:0"


def T(x):
    return torch.tensor(np.asarray(x, dtype=np.float64))


def rel_rms(a, b, floor=1e-6):
    """||a - b|| / max(||b||, floor): relative RMS with an absolute floor (kJ/mol per
    unit) so gradients that vanish analytically (e.g. dE/dbox without PME and without
    boundary-crossing pairs, where both sides hold rounding noise) compare sensibly."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), floor))


def oracle_eval(s, pos, box, pairs, rc, lpme, steps_pol=None, solver="reference", tol=1e-6, idx=None, scales=None):
    n = pos.shape[0]
    at = s.axis_types if idx is None else s.axis_types
    f = O.ADMPPmeForce(box, at, s.axis_idx, rc, 5e-4, s.pme.lmax, s.pme.lpol, lpme, steps_pol)
    p = T(pos).requires_grad_(True)
    b = T(box).requires_grad_(True)
    Q = T(s.pme.Q_local[s.map_atomtype]).requires_grad_(True)
    mS, pS, dS = scales if scales is not None else (s.pme.mScales, s.pme.pScales, s.pme.dScales)
    out = {}
    if s.pme.lpol:
        a = T(s.pme.pol[s.map_poltype]).requires_grad_(True)
        t = T(s.pme.thole[s.map_poltype]).requires_grad_(True)
        aux = {}
        e, aux = f.get_energy(p, b, pairs, Q, a, t, mS, pS, dS, solver=solver, tol=tol, aux=aux)
        gp, gb, gq, ga, gt = torch.autograd.grad(e, (p, b, Q, a, t))
        out.update(pol=ga.numpy(), thole=gt.numpy(), U=aux["U_ind"].numpy() * 10.0)
    else:
        e = f.get_energy(p, b, pairs, Q, mScales=mS)
        gp, gb, gq = torch.autograd.grad(e, (p, b, Q))
    out.update(E=float(e.detach()), pos=gp.numpy(), box=gb.numpy(), Q=gq.numpy())
    return out, f


def gpu_eval(s, pos, box, pairs, rc, lpme, steps_pol=None, solver="reference", tol=None, precision="fp64",
             scales=None, maxiter=None):
    from dmff_amd.admp.pme import ADMPPmeForce
    f = ADMPPmeForce(box, s.map_atomtype, s.axis_types, s.axis_idx, rc, 5e-4, s.pme.lmax, s.pme.lpol, lpme,
                     steps_pol, precision=precision, solver=solver, tol=tol, maxiter=maxiter, device=DEV)
    p = torch.tensor(pos, dtype=F64, device=DEV, requires_grad=True)
    b = torch.tensor(box, dtype=F64, device=DEV, requires_grad=True)
    Q = torch.tensor(s.pme.Q_local[s.map_atomtype], dtype=F64, device=DEV, requires_grad=True)
    mS, pS, dS = scales if scales is not None else (s.pme.mScales, s.pme.pScales, s.pme.dScales)
    pr = torch.as_tensor(pairs, device=DEV)
    out = {}
    if s.pme.lpol:
        a = torch.tensor(s.pme.pol[s.map_poltype], dtype=F64, device=DEV, requires_grad=True)
        t = torch.tensor(s.pme.thole[s.map_poltype], dtype=F64, device=DEV, requires_grad=True)
        e, aux = f.get_energy(p, b, pr, Q, a, t, mS, pS, dS, aux={})
        gp, gb, gq, ga, gt = torch.autograd.grad(e, (p, b, Q, a, t))
        out.update(pol=ga.cpu().numpy(), thole=gt.cpu().numpy(), U=aux["U_ind"].cpu().numpy() * 10.0, aux=aux)
    else:
        e = f.get_energy(p, b, pr, Q, mS)
        gp, gb, gq = torch.autograd.grad(e, (p, b, Q))
    out.update(E=float(e.detach()), pos=gp.cpu().numpy(), box=gb.cpu().numpy(), Q=gq.cpu().numpy(),
               info=f.last_info)
    return out


def compare(g, o, etol=1e-9, gtol=1e-7, keys=("pos", "box", "Q", "pol", "thole")):
    msgs = []
    de = abs(g["E"] - o["E"]) / max(1.0, abs(o["E"]))
    msgs.append(f"E gpu={g['E']:.12f} oracle={o['E']:.12f} rel={de:.2e}")
    ok = de < etol
    for k in keys:
        if k in o:
            r = rel_rms(g[k], o[k])
            msgs.append(f"{k}: relrms={r:.2e}")
            ok = ok and r < gtol
    print("; ".join(msgs), flush=True)
    return ok, msgs


@pytest.mark.parametrize("ff,lpme,steps", [
    ("admp_nonpol.xml", False, None),
    ("admp_mono.xml", False, 5),
    ("admp.xml", False, 5),
])
def test_water_dimer_cutoff(ff, lpme, steps):
    s = load_system(ff, "water_dimer.pdb")
    pos, box = s.pos_nm * 10, s.box_nm * 10
    pairs = reference_pairs(pos, box, 4.0, s.bonded, pad=4)
    o, _ = oracle_eval(s, pos, box, pairs, 4.0, lpme, steps)
    g = gpu_eval(s, pos, box, pairs, 4.0, lpme, steps)
    ok, msgs = compare(g, o)
    assert ok, msgs


def test_fluctuating_dimer_pme_full_precision_pin():
    """examples/fluctuated_leading_term_waterff/ref_out: E = -41.437696011272465 and the
    6x3 gradient, with charges that depend on the geometry (run.py:22-56): the position
    gradient flows through dE/dQ_local of the library into the charge model (autograd)."""
    import json
    from dmff_amd.admp.pme import ADMPPmeForce
    from tests.systems import GOLD
    from tests.test_oracle_pins import fluct_charges
    en = json.load(open(f"{GOLD}/energies.json"))
    s = load_system("fluct_forcefield.xml", "fluct_water_dimer.pdb", "fluct_residues.xml")
    pos, box = s.pos_nm * 10, s.box_nm * 10
    pairs = reference_pairs(pos, box, 4.0, s.bonded)
    f = ADMPPmeForce(box, s.map_atomtype, s.axis_types, s.axis_idx, 4.0, 5e-4, 2, True, True, 5, device=DEV)
    p = torch.tensor(pos, dtype=F64, device=DEV, requires_grad=True)
    b = torch.tensor(box, dtype=F64, device=DEV)
    Q = torch.tensor(s.pme.Q_local[s.map_atomtype], dtype=F64, device=DEV)
    Q = torch.cat([fluct_charges(p, b)[:, None], Q[:, 1:]], dim=1)
    e = f.get_energy(p, b, torch.as_tensor(pairs, device=DEV), Q, s.pme.pol[s.map_poltype],
                     s.pme.thole[s.map_poltype], s.pme.mScales, s.pme.pScales, s.pme.dScales)
    (g,) = torch.autograd.grad(e, p)
    print(f"fluct: E {float(e.detach()):.12f} ref {en['fluct_E']:.12f}", flush=True)
    assert abs(float(e.detach()) - en["fluct_E"]) < 1e-8
    assert rel_rms(g.cpu().numpy(), np.asarray(en["fluct_grad"])) < 1e-7


def test_small_box_pme_polar_pcg():
    s, pos, box = _small_waters()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    o, _ = oracle_eval(s, pos, box, pairs, 6.0, True, solver="cg", tol=1e-9)
    g = gpu_eval(s, pos, box, pairs, 6.0, True, solver="pcg", tol=1e-9)
    ok, msgs = compare(g, o)
    du = np.abs(g["U"] - o["U"]).max()
    print(f"U max abs diff {du:.2e}; solver info {g['info']['n_iter']} {g['info']['residual']}")
    assert ok and du < 1e-8, msgs


def test_small_box_reference_jacobi_matches_oracle_iteration_count():
    s, pos, box = _small_waters()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    o, fo = oracle_eval(s, pos, box, pairs, 6.0, True)
    g = gpu_eval(s, pos, box, pairs, 6.0, True)
    ok, msgs = compare(g, o)
    assert ok, msgs
    assert np.abs(g["U"] - o["U"]).max() < 1e-8


@pytest.mark.parametrize("ff,key", [
    ("quadnonpol_forcefield.xml", "quadnonpol_E"),
    ("fullpol_forcefield.xml", "fullpol_E"),
])
def test_waterbox_2988(ff, key):
    """examples/water_fullpol: 2988 atoms, rc 6 A, PME; reference Jacobi (POL_CONV 1.0)."""
    import json
    from tests.systems import GOLD
    en = json.load(open(f"{GOLD}/energies.json"))
    s = load_system(ff, "waterbox_31ang.pdb", "fullpol_residues.xml")
    pos, box = s.pos_nm * 10, s.box_nm * 10
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    o, _ = oracle_eval(s, pos, box, pairs, 6.0, True)
    g = gpu_eval(s, pos, box, pairs, 6.0, True)
    ok, msgs = compare(g, o, keys=("pos", "box", "Q", "pol", "thole"))
    assert ok, msgs
    assert abs(g["E"] / en[key] - 1) < 1e-5
    if "U" in o:
        ref = np.load(f"{GOLD}/water_fullpol_uind.npy")
        assert np.abs(g["U"].ravel() - o["U"].ravel()).max() < 1e-8
        assert np.linalg.norm(g["U"].ravel() - ref[:, 1]) / np.linalg.norm(ref[:, 1]) < 5e-4


def test_neighbor_list_bit_exact():
    from dmff_amd.admp.engine import Engine
    s = load_system("fullpol_forcefield.xml", "waterbox_31ang.pdb", "fullpol_residues.xml")
    pos, box = s.pos_nm * 10, s.box_nm * 10
    for rc in (6.0, 8.0, 4.0):
        ref = reference_pairs(pos, box, rc, s.bonded)
        e = Engine(len(pos), 0, False, False, 0.0, (0, 0, 0), rc, device=DEV)
        e.set_topology(None, None, s.bonded)
        e.build_neighbors(pos, box)
        got = e.get_pairs(0).cpu().numpy().astype(np.int64)
        a = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
        b = got[np.lexsort((got[:, 1], got[:, 0]))]
        print(f"rc={rc}: oracle {len(a)} pairs, gpu {len(b)} pairs", flush=True)
        assert a.shape == b.shape and np.array_equal(a, b)
    # a row capacity of 8: the first build overflows and is rebuilt before the lists are read
    ref = reference_pairs(pos, box, 6.0, s.bonded)
    e = Engine(len(pos), 0, False, False, 0.0, (0, 0, 0), 6.0, max_neighbors=8, device=DEV)
    e.set_topology(None, None, s.bonded)
    e.build_neighbors(pos, box)
    got = e.get_pairs(0).cpu().numpy().astype(np.int64)
    a = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
    b = got[np.lexsort((got[:, 1], got[:, 0]))]
    assert a.shape == b.shape and np.array_equal(a, b)
    # random points with pairs at the cutoff to the last bit, small box (nc < 3)
    rng = np.random.default_rng(5)
    L = 11.0
    p = rng.uniform(0, L, size=(300, 3))
    p[1] = p[0] + np.array([4.0, 0, 0])
    p[3] = p[2] + np.array([0, 3.9999999999, 0])
    for rc in (4.0, 5.4):
        ref = reference_pairs(p, np.eye(3) * L, rc, {})
        e = Engine(len(p), 0, False, False, 0.0, (0, 0, 0), rc, device=DEV)
        e.set_topology(None, None, {})
        e.build_neighbors(p, np.eye(3) * L)
        got = e.get_pairs(0).cpu().numpy().astype(np.int64)
        a = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
        b = got[np.lexsort((got[:, 1], got[:, 0]))]
        assert np.array_equal(a, b), (len(a), len(b))
    # rc above half the box (single minimum image) and triclinic boxes
    from tests.systems import neighbour_cases
    for name, p, bx, rc in neighbour_cases():
        ref = reference_pairs(p, bx, rc, {})
        e = Engine(len(p), 0, False, False, 0.0, (0, 0, 0), rc, device=DEV)
        e.set_topology(None, None, {})
        e.build_neighbors(p, bx)
        got = e.get_pairs(0).cpu().numpy().astype(np.int64)
        a = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
        b = got[np.lexsort((got[:, 1], got[:, 0]))]
        print(f"{name}: oracle {len(a)} pairs, gpu {len(b)} pairs", flush=True)
        assert len(a) > 0 and a.shape == b.shape and np.array_equal(a, b), name


def test_mixed_precision_error_bounds():
    s, pos, box = _small_waters()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    o, _ = oracle_eval(s, pos, box, pairs, 6.0, True, solver="cg", tol=1e-9)
    g = gpu_eval(s, pos, box, pairs, 6.0, True, solver="pcg", tol=1e-4, precision="mixed")
    de = abs(g["E"] - o["E"]) / abs(o["E"])
    df = rel_rms(g["pos"], o["pos"])
    print(f"mixed: E rel {de:.2e}, force relrms {df:.2e}, box {rel_rms(g['box'], o['box']):.2e}", flush=True)
    assert de < 2e-5 and df < 1e-4


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_dispersion_and_slater_vs_oracle(prec):
    """Dispersion PME (pmax 10) and the Slater term; mixed runs the FP32 pair,
    spread and FFT kernels (tolerance: energy 2e-5 relative, gradients 1e-4
    relative RMS; FP64: 1e-9 / 1e-8)."""
    from dmff_amd.admp.engine import Engine
    te, tg = (1e-9, 1e-8) if prec == "fp64" else (2e-5, 1e-4)
    rng = np.random.default_rng(9)
    L = 20.0
    n = 60
    pos = rng.uniform(0, L, size=(n, 3))
    c = rng.uniform(5, 40, size=(n, 3))
    m = np.array([0.0, 0.0, 1.0, 1.0, 1.0, 1.0])
    bonded = {(0, 1): 1, (1, 0): 1, (2, 3): 2, (3, 2): 2}
    rc = 9.0
    pairs = reference_pairs(pos, np.eye(3) * L, rc, bonded)
    kappa, K1, K2, K3 = O.setup_ewald_parameters(rc, 5e-4, np.eye(3) * L)
    p = T(pos).requires_grad_(True)
    b = T(np.eye(3) * L).requires_grad_(True)
    ct = T(c).requires_grad_(True)
    e = O.energy_disp_pme(p, b, pairs, ct, m, kappa, (K1, K2, K3), 10)
    gp, gb, gc = torch.autograd.grad(e, (p, b, ct))
    eng = Engine(n, 0, False, False, 0.0, (0, 0, 0), rc, disp_kappa=kappa, dK=(K1, K2, K3), disp_pmax=10,
                 precision=prec, device=DEV)
    eng.set_topology(None, None, bonded)
    eng.set_pairs(torch.as_tensor(pairs))
    E, g = eng.disp_compute(pos, np.eye(3) * L, c, m)
    errs = [abs(float(E[0]) - float(e)) / abs(float(e)), rel_rms(g["pos"].cpu().numpy(), gp.numpy()),
            rel_rms(g["box"].cpu().numpy(), gb.numpy()), rel_rms(g["c"].cpu().numpy(), gc.numpy())]
    print(f"disp {prec}: E gpu {float(E[0]):.12f} oracle {float(e):.12f}; errors E/pos/box/c {errs}", flush=True)
    assert errs[0] < te and max(errs[1:]) < tg
    A = rng.uniform(1e2, 3e2, size=n)
    B = rng.uniform(3.0, 4.5, size=n)
    At, Bt = T(A).requires_grad_(True), T(B).requires_grad_(True)
    p2 = T(pos).requires_grad_(True)
    b2 = T(np.eye(3) * L).requires_grad_(True)
    es = O.slater_sr_energy(p2, b2, pairs, m, At, Bt)
    gp2, gb2, ga2, gbb2 = torch.autograd.grad(es, (p2, b2, At, Bt))
    E2, g2 = eng.slater_compute(pos, np.eye(3) * L, A, B, m)
    errs = [abs(float(E2[0]) - float(es)) / max(1, abs(float(es)))] + [
        rel_rms(g2[k].cpu().numpy(), v.numpy()) for k, v in (("pos", gp2), ("a", ga2), ("b", gbb2), ("box", gb2))]
    print(f"slater {prec}: E gpu {float(E2[0]):.12f} oracle {float(es):.12f}; errors {errs}", flush=True)
    assert errs[0] < te and max(errs[1:]) < tg


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_short_range_kinds_vs_oracle(prec):
    """admp_pairwise_compute for the four generate_pairwise_interaction kernels
    against the oracle (FP64: energy 1e-11, gradients 1e-10; mixed: 1e-5, 1e-4)."""
    from dmff_amd.admp.engine import Engine
    from tests.systems import pairwise_case, pairwise_oracle
    te, tg = (1e-11, 1e-10) if prec == "fp64" else (1e-5, 1e-4)
    for kind in (1, 2, 3, 4):
        pos, box, pairs, m, prm = pairwise_case(kind)
        o = pairwise_oracle(kind, pos, box, pairs, m, prm)
        eng = Engine(len(pos), 0, False, False, 0.0, (0, 0, 0), 9.0, precision=prec, device=DEV)
        eng.set_topology(None, None, {(0, 1): 1, (1, 0): 1, (2, 3): 2, (3, 2): 2})
        eng.set_pairs(torch.as_tensor(pairs))
        E, g = eng.pairwise_compute(kind, pos, box, prm, m)
        errs = [abs(float(E[0]) - o["E"]) / abs(o["E"])] + [rel_rms(g[k][0].cpu().numpy(), o[k])
                                                            for k in ("pos", "box", "prm")]
        print(f"kind {kind} {prec}: E {float(E[0]):.12e} oracle {o['E']:.12e}; errors E/pos/box/prm {errs}",
              flush=True)
        assert errs[0] < te and max(errs[1:]) < tg, (kind, errs)


def test_batched_frames_equal_single():
    from dmff_amd.admp.pme import ADMPPmeForce
    s, pos, box = _small_waters()
    rng = np.random.default_rng(3)
    frames = np.stack([pos + rng.normal(0, 0.005, pos.shape) for _ in range(3)])
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    Q = s.pme.Q_local[s.map_atomtype]
    pol = s.pme.pol[s.map_poltype]
    th = s.pme.thole[s.map_poltype]
    args = (s.pme.mScales, s.pme.pScales, s.pme.dScales)
    fb = ADMPPmeForce(box, s.map_atomtype, s.axis_types, s.axis_idx, 6.0, 5e-4, 2, True, True, None,
                      solver="pcg", tol=1e-8, n_frames=3, device=DEV)
    eng = fb.engine
    for k in range(3):
        eng.set_pairs(torch.as_tensor(pairs, device=DEV), frame=k)
    Eb = fb.get_energy(torch.tensor(frames), np.stack([box] * 3), None, Q, pol, th, *args)
    for k in range(3):
        f1 = ADMPPmeForce(box, s.map_atomtype, s.axis_types, s.axis_idx, 6.0, 5e-4, 2, True, True, None,
                          solver="pcg", tol=1e-8, device=DEV)
        e1 = f1.get_energy(frames[k], box, pairs, Q, pol, th, *args)
        assert abs(float(Eb[k]) - float(e1)) < 1e-9 * abs(float(e1))


def test_batched_frames_side_stream_product():
    """14 frames of the small box: 14 x 10 kz planes exceed the 128 that ride in the
    plane launch, so the real-space product runs on the second stream (the path of
    the 64-frame batch); every frame equals its single-frame evaluation."""
    from dmff_amd.admp.engine import Engine
    s, pos, box = _small_waters()
    rng = np.random.default_rng(21)
    F = 14
    P = np.stack([pos + rng.normal(0, 0.01, pos.shape) for _ in range(F)])
    kappa, K1, K2, K3 = O.setup_ewald_parameters(6.0, 5e-4, box)
    assert (K3 // 2 + 1) * F > 128
    args = (s.pme.Q_local[s.map_atomtype], s.pme.pol[s.map_poltype], s.pme.thole[s.map_poltype],
            s.pme.mScales, s.pme.pScales, s.pme.dScales)

    def run(Pf):
        eng = Engine(len(pos), 2, True, True, kappa, (K1, K2, K3), 6.0, n_frames=len(Pf), device=DEV)
        eng.set_topology(s.axis_types, s.axis_idx, s.bonded)
        eng.build_neighbors(Pf, np.stack([box] * len(Pf)))
        E, g, U, info = eng.compute(Pf, np.stack([box] * len(Pf)), *args, solver="pcg", tol=1e-9, maxiter=500,
                                    want_grad=True)
        assert all(info["converged"][: len(Pf)])
        return E.cpu().numpy(), g["pos"].cpu().numpy()

    Eb, Gb = run(P)
    for k in (0, 7, 13):
        E1, G1 = run(P[k:k + 1])
        assert abs(Eb[k] - E1[0]) < 1e-9 * abs(E1[0]), (k, Eb[k], E1[0])
        assert rel_rms(Gb[k], G1[0]) < 1e-7


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_block_preconditioners_same_solution_fewer_iterations(prec):
    """4096 waters (the 12,288-site system): the cluster-block, pair-block and
    diagonal preconditioners reach the same induced dipoles (max|dE/dU| < 1e-6
    each); the pair blocks take fewer iterations than the diagonal, the clusters
    no more than the pairs.  The small-box tests check them against the oracle."""
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.water import config_system
    S, c = config_system(1)
    box, rc = S.box_nm * 10.0, c["rc"] * 10.0
    kappa, K1, K2, K3 = O.setup_ewald_parameters(rc, 5e-4, box)
    eng = Engine(S.n_atoms, 2, True, True, kappa, (K1, K2, K3), rc, precision=prec, device=DEV)
    eng.set_topology(S.axis_types, S.axis_idx, S.bonded)
    pos = S.positions_nm * 10.0
    eng.build_neighbors(pos, box)
    out = {}
    for pc in ("diagonal", "pairs", "clusters"):
        E, g, U, info = eng.compute(pos, box, S.Q_local, S.pol, S.thole, S.mScales, S.pScales, S.dScales,
                                    solver="pcg", tol=1e-6, maxiter=500, want_grad=True, precond=pc)
        out[pc] = (float(E[0]), g["pos"][0].cpu().numpy(), U[0].cpu().numpy(), info)
        print(f"{prec} {pc}: E={float(E[0]):.10f} iters={info['n_iter'][0]} res={info['residual'][0]:.2e}",
              flush=True)
    (Ed, gd, Ud, idg), (Ep, gp, Up, ip) = out["diagonal"], out["pairs"]
    ic = out["clusters"][3]
    assert idg["residual"][0] < 1e-6 and ip["residual"][0] < 1e-6 and ic["residual"][0] < 1e-6
    assert ip["n_iter"][0] < idg["n_iter"][0]
    assert ic["n_iter"][0] <= ip["n_iter"][0]
    etol, ftol = (1e-11, 1e-8) if prec == "fp64" else (2e-8, 1e-6)
    for E2, g2, U2, _ in (out["pairs"], out["clusters"]):
        assert abs(E2 - Ed) < etol * abs(Ed)
        assert rel_rms(g2, gd) < ftol
        assert np.abs(U2 - Ud).max() < 1e-7


def test_small_box_diagonal_preconditioner_vs_oracle():
    from dmff_amd.admp.engine import Engine
    s, pos, box = _small_waters()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    o, _ = oracle_eval(s, pos, box, pairs, 6.0, True, solver="cg", tol=1e-9)
    kappa, K1, K2, K3 = O.setup_ewald_parameters(6.0, 5e-4, box)
    eng = Engine(len(pos), 2, True, True, kappa, (K1, K2, K3), 6.0, device=DEV)
    eng.set_topology(s.axis_types, s.axis_idx, s.bonded)
    eng.set_pairs(torch.as_tensor(pairs, device=DEV))
    E, g, U, info = eng.compute(pos, box, s.pme.Q_local[s.map_atomtype], s.pme.pol[s.map_poltype],
                                s.pme.thole[s.map_poltype], s.pme.mScales, s.pme.pScales, s.pme.dScales,
                                solver="pcg", tol=1e-9, maxiter=500, want_grad=True, precond="diagonal")
    assert abs(float(E[0]) - o["E"]) < 1e-9 * abs(o["E"])
    assert rel_rms(g["pos"][0].cpu().numpy(), o["pos"]) < 1e-7
    assert np.abs(U[0].cpu().numpy() - o["U"]).max() < 1e-8


def test_compute_without_stats_same_results_and_checks():
    """admp_compute without solver statistics (the form that need not wait for
    the device at its end): bit-identical energy, gradients and U to the call with
    them, and a non-finite position is still an argument error.  The first call
    of a context runs its set-up and final pass eagerly, the second captures them
    as graphs and later calls replay those: every call is compared with the
    first, with host synchronisations and other work between them."""
    import dmff_amd._lib as L
    from dmff_amd.admp.engine import Engine
    s, pos, box = _small_waters()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    kappa, K1, K2, K3 = O.setup_ewald_parameters(6.0, 5e-4, box)
    eng = Engine(len(pos), 2, True, True, kappa, (K1, K2, K3), 6.0, device=DEV)
    eng.set_topology(s.axis_types, s.axis_idx, s.bonded)
    eng.set_pairs(torch.as_tensor(pairs, device=DEV))
    args = (s.pme.Q_local[s.map_atomtype], s.pme.pol[s.map_poltype], s.pme.thole[s.map_poltype],
            s.pme.mScales, s.pme.pScales, s.pme.dScales)
    kw = dict(solver="pcg", tol=1e-9, maxiter=500, want_grad=True)
    E1, g1, U1, info = eng.compute(pos, box, *args, **kw)
    E2, g2, U2, none = eng.compute(pos, box, *args, stats=False, **kw)
    torch.cuda.synchronize()
    assert none is None and info["n_iter"][0] > 0
    assert torch.equal(E1, E2) and torch.equal(U1, U2)
    for k in g1:
        assert torch.equal(g1[k], g2[k]), k
    bad = pos.copy()
    bad[5, 1] = np.nan
    with pytest.raises(L.ADMPError, match="positions"):
        eng.compute(bad, box, *args, stats=False, **kw)
    E3, _, _, _ = eng.compute(pos, box, *args, stats=False, **kw)
    assert torch.equal(E1, E3)
    for k in range(4):
        torch.cuda.synchronize()
        _ = torch.randn(1000, device=DEV).sum().item()   # other work between replays
        Ek, gk, Uk, _ = eng.compute(pos, box, *args, stats=bool(k % 2), **kw)
        torch.cuda.synchronize()
        assert torch.equal(E1, Ek) and torch.equal(U1, Uk), k
        for name in g1:
            assert torch.equal(g1[name], gk[name]), (k, name)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_equal_eager_over_varied_positions(precision):
    """Twelve evaluations of displaced geometries through a context that captures
    and replays its set-up, final-pass and solver-chunk graphs (chunks of 1-4
    iterations, the last one cut to the predicted remainder, starting on either
    buffer parity) against a context created with ADMP_NO_GRAPH=1 that launches
    the same kernels eagerly: energy, U and every gradient bit-identical."""
    import os
    from dmff_amd.admp.engine import Engine
    s, pos, box = _small_waters()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    kappa, K1, K2, K3 = O.setup_ewald_parameters(6.0, 5e-4, box)

    def make():
        e = Engine(len(pos), 2, True, True, kappa, (K1, K2, K3), 6.0, precision=precision, device=DEV)
        e.set_topology(s.axis_types, s.axis_idx, s.bonded)
        e.set_pairs(torch.as_tensor(pairs, device=DEV))
        return e
    graphs = make()
    old = os.environ.get("ADMP_NO_GRAPH")
    os.environ["ADMP_NO_GRAPH"] = "1"
    try:
        eager = make()
    finally:
        if old is None:
            del os.environ["ADMP_NO_GRAPH"]
        else:
            os.environ["ADMP_NO_GRAPH"] = old
    args = (s.pme.Q_local[s.map_atomtype], s.pme.pol[s.map_poltype], s.pme.thole[s.map_poltype],
            s.pme.mScales, s.pme.pScales, s.pme.dScales)
    rng = np.random.default_rng(5)
    iters = set()
    for k in range(12):
        p = pos + rng.normal(0.0, 0.03, pos.shape)
        tol = (1e-7, 1e-9, 1e-10)[k % 3]
        kw = dict(solver="pcg", tol=tol, maxiter=500, want_grad=True)
        Eg, gg, Ug, info = graphs.compute(p, box, *args, **kw)
        Ee, ge, Ue, _ = eager.compute(p, box, *args, **kw)
        torch.cuda.synchronize()
        iters.add(info["n_iter"][0])
        assert torch.equal(Eg, Ee) and torch.equal(Ug, Ue), k
        for name in gg:
            assert torch.equal(gg[name], ge[name]), (k, name)
    assert len(iters) >= 2, iters


@pytest.mark.parametrize("fp32_products", [False, True])
def test_mixed_batched_frames_converge(fp32_products):
    """Mixed precision on a batch: frames converge at different iterations (and,
    with FP32 reciprocal products, the true-residual replacements and the switch
    to FP64 products resume them); every frame must reach the FP64 residual
    tolerance and agree with the FP64 batch."""
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import setup_ewald_parameters
    s, pos, box = _small_waters()
    rng = np.random.default_rng(11)
    F = 4
    frames = np.stack([pos + rng.normal(0, 0.02 * (k + 1), pos.shape) for k in range(F)])
    kappa, K1, K2, K3 = setup_ewald_parameters(6.0, 5e-4, box)
    Q = s.pme.Q_local[s.map_atomtype]
    pol = s.pme.pol[s.map_poltype]
    th = s.pme.thole[s.map_poltype]
    out = {}
    for prec in ("mixed", "fp64"):
        eng = Engine(len(pos), 2, True, True, kappa, (K1, K2, K3), 6.0, n_frames=F, precision=prec, device=DEV)
        eng.set_topology(s.axis_types, s.axis_idx, s.bonded)
        eng.build_neighbors(frames, np.stack([box] * F))
        E, g, U, info = eng.compute(frames, np.stack([box] * F), Q, pol, th, s.pme.mScales, s.pme.pScales,
                                    s.pme.dScales, solver="pcg", tol=1e-6, maxiter=400,
                                    fp32_products=fp32_products and prec == "mixed")
        print(prec, info["n_iter"], info["residual"], flush=True)
        assert all(r < 1e-6 for r in info["residual"]), info
        assert max(info["n_iter"]) < 200
        out[prec] = (E.cpu().numpy(), g["pos"].cpu().numpy())
    Em, Gm = out["mixed"]
    Ed, Gd = out["fp64"]
    assert np.all(np.abs(Em - Ed) < 2e-5 * np.abs(Ed))
    for k in range(F):
        assert np.linalg.norm(Gm[k] - Gd[k]) / np.linalg.norm(Gd[k]) < 1e-4


@pytest.mark.parametrize("K", [(16, 16, 16), (20, 24, 15), (32, 27, 25), (21, 22, 26), (6, 7, 8), (104, 104, 104),
                               (50, 30, 40), (10, 70, 50)])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_grid_sizes_vs_oracle(K, prec):
    """Energies and forces on grids whose sizes exercise every FFT radix path
    (4, 2, 3, 5, 7, 11, 13), the scatter path of grids below 9 points per axis,
    and (104^3) the line-major FFT passes, against the oracle on the same grid."""
    from dmff_amd.admp.engine import Engine
    s, pos, box = _small_waters()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    kappa = O.setup_ewald_parameters(6.0, 5e-4, box)[0]
    Ql = s.pme.Q_local[s.map_atomtype]
    mS = s.pme.mScales
    p = T(pos).requires_grad_(True)
    e = O.energy_pme(p, T(box), T(pairs).long(), T(Ql), None, None, None, mS, None, None,
                     s.axis_types, s.axis_idx, kappa, K, 2, False, True)
    (gp,) = torch.autograd.grad(e, (p,))
    eng = Engine(len(pos), 2, False, True, kappa, K, 6.0, precision=prec, device=DEV)
    eng.set_topology(s.axis_types, s.axis_idx, s.bonded)
    eng.set_pairs(torch.as_tensor(pairs, device=DEV))
    E, g, U, info = eng.compute(pos, box, Ql, mScales=mS, want_grad=True)
    de = abs(float(E[0]) - float(e)) / abs(float(e))
    df = rel_rms(g["pos"][0].cpu().numpy(), gp.numpy())
    print(f"K={K} {prec}: E rel {de:.2e} force relrms {df:.2e}", flush=True)
    etol, ftol = (1e-9, 1e-7) if prec == "fp64" else (2e-5, 1e-4)
    assert de < etol and df < ftol


def test_per_frame_pairs_survive_capacity_growth():
    """Pair lists set frame by frame, the second one longer (capacity grows):
    the first frame's list must survive the reallocation (dmff/utils.py:65-84
    semantics per frame), so each frame equals a single-frame evaluation."""
    from dmff_amd.admp.engine import Engine
    s, pos, box = _small_waters()
    kappa, K1, K2, K3 = O.setup_ewald_parameters(6.0, 5e-4, box)
    Ql = s.pme.Q_local[s.map_atomtype]
    mS = s.pme.mScales
    lists = [reference_pairs(pos, box, 4.0, s.bonded), reference_pairs(pos, box, 6.0, s.bonded)]
    eng = Engine(len(pos), 2, False, True, kappa, (K1, K2, K3), 6.0, n_frames=2, device=DEV)
    eng.set_topology(s.axis_types, s.axis_idx, s.bonded)
    for k in range(2):
        eng.set_pairs(torch.as_tensor(lists[k], device=DEV), frame=k)
    E, g, U, info = eng.compute(np.stack([pos, pos]), np.stack([box, box]), Ql, mScales=mS, want_grad=True)
    for k in range(2):
        e1 = Engine(len(pos), 2, False, True, kappa, (K1, K2, K3), 6.0, device=DEV)
        e1.set_topology(s.axis_types, s.axis_idx, s.bonded)
        e1.set_pairs(torch.as_tensor(lists[k], device=DEV))
        Ek, gk, _, _ = e1.compute(pos, box, Ql, mScales=mS, want_grad=True)
        assert abs(float(E[k]) - float(Ek[0])) <= 1e-12 * abs(float(Ek[0])), (k, float(E[k]), float(Ek[0]))
        assert rel_rms(g["pos"][k].cpu().numpy(), gk["pos"][0].cpu().numpy()) < 1e-12


def test_backward_central_differences_12288_atoms():
    """Hand-written backward at full size (4096 polarisable waters, 12,288 sites,
    K = 50^3): dE/dpos on 50 random coordinates and dE/d(Q_local, pol, thole) per
    type on 20 random parameters against central differences (relative step 1e-5,
    FP64, induced dipoles solved to 1e-10 so the Hellmann-Feynman gradient is exact
    to the tolerance written here).  Tolerance: relative error 1e-5 per checked
    component (or 1e-4 kJ/mol per unit absolute for near-zero components)."""
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import TypeReducer, setup_ewald_parameters
    from dmff_amd.admp.water import config_system
    S, c = config_system(1)
    box = S.box_nm * 10.0
    rc = c["rc"] * 10.0
    kappa, K1, K2, K3 = setup_ewald_parameters(rc, 5e-4, box)
    eng = Engine(S.n_atoms, 2, True, True, kappa, (K1, K2, K3), rc, device=DEV)
    eng.set_topology(S.axis_types, S.axis_idx, S.bonded)
    pos0 = torch.as_tensor(S.positions_nm * 10.0, dtype=F64, device=DEV)
    boxt = torch.as_tensor(box, dtype=F64, device=DEV)
    eng.build_neighbors(pos0, boxt)   # the pair set stays fixed for the differences
    qt, pt = TypeReducer(S.map_atomtype, device=DEV), TypeReducer(S.map_poltype, device=DEV)
    Qt = torch.as_tensor(S.Q_local_types, dtype=F64, device=DEV)
    Pt = torch.as_tensor(S.pol_types, dtype=F64, device=DEV)
    Tt = torch.as_tensor(S.thole_types, dtype=F64, device=DEV)
    mq = torch.as_tensor(S.map_atomtype, device=DEV)
    mp = torch.as_tensor(S.map_poltype, device=DEV)

    def energy(pos, Q, P, Th, grad=False):
        E, g, U, info = eng.compute(pos, boxt, Q[mq], P[mp], Th[mp], S.mScales, S.pScales, S.dScales,
                                    solver="pcg", tol=1e-10, maxiter=1000, want_grad=grad)
        return float(E[0]), g

    E0, g = energy(pos0, Qt, Pt, Tt, grad=True)
    gpos = g["pos"][0]
    gQ, gP, gT = qt(g["Q"][0]), pt(g["pol"][0]), pt(g["thole"][0])
    rng = np.random.default_rng(20)

    def check(analytic, fd, what):
        err = abs(analytic - fd)
        ok = err <= 1e-5 * abs(fd) or err <= 1e-4
        assert ok, (what, analytic, fd)

    for _ in range(50):
        a, d = int(rng.integers(S.n_atoms)), int(rng.integers(3))
        h = 1e-5 * max(1.0, abs(float(pos0[a, d])))
        pp, pm = pos0.clone(), pos0.clone()
        pp[a, d] += h
        pm[a, d] -= h
        fd = (energy(pp, Qt, Pt, Tt)[0] - energy(pm, Qt, Pt, Tt)[0]) / (2 * h)
        check(float(gpos[a, d]), fd, ("pos", a, d))
    params = [("Q", t, k) for t in range(Qt.shape[0]) for k in range(Qt.shape[1])] + \
             [("pol", t, 0) for t in range(len(Pt))] + [("thole", t, 0) for t in range(len(Tt))]
    for idx in rng.choice(len(params), size=20, replace=False):
        kind, t, k = params[idx]
        base = {"Q": Qt, "pol": Pt, "thole": Tt}[kind]
        v = float(base[t, k] if kind == "Q" else base[t])
        h = 1e-5 * max(abs(v), 1e-3)
        vals = []
        for sgn in (1, -1):
            Q, P, Th = Qt.clone(), Pt.clone(), Tt.clone()
            tgt = {"Q": Q, "pol": P, "thole": Th}[kind]
            if kind == "Q":
                tgt[t, k] += sgn * h
            else:
                tgt[t] += sgn * h
            vals.append(energy(pos0, Q, P, Th)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        an = float(gQ[t, k]) if kind == "Q" else float((gP if kind == "pol" else gT)[t])
        check(an, fd, (kind, t, k))
    # virial: dE/dbox with Cartesian positions held fixed (the reference's box
    # gradient, tests/test_admp/test_gradDisp.py:57-58), all nine components
    gbox = g["box"][0]
    for p_ in range(3):
        for q_ in range(3):
            h = 1e-5 * float(boxt[0, 0])
            vals = []
            for sgn in (1, -1):
                bx = boxt.clone()
                bx[p_, q_] += sgn * h
                E, _, _, _ = eng.compute(pos0, bx, Qt[mq], Pt[mp], Tt[mp], S.mScales, S.pScales, S.dScales,
                                         solver="pcg", tol=1e-10, maxiter=1000, want_grad=False)
                vals.append(float(E[0]))
            fd = (vals[0] - vals[1]) / (2 * h)
            print(f"dE/dbox[{p_},{q_}] analytic {float(gbox[p_, q_]):.9f} fd {fd:.9f}", flush=True)
            check(float(gbox[p_, q_]), fd, ("box", p_, q_))
    # Thole widths enter only where pScale = 0 and dScale != 0: with dScales = 1 the
    # intramolecular induced-induced pairs use a = thole_i + thole_j (pme.py:794-795)
    dS1 = np.ones_like(S.dScales)

    def e_d1(Th, grad=False):
        E, gg, U, info = eng.compute(pos0, boxt, Qt[mq], Pt[mp], Th[mp], S.mScales, S.pScales, dS1,
                                     solver="pcg", tol=1e-10, maxiter=1000, want_grad=grad)
        return float(E[0]), gg

    _, gd = e_d1(Tt, grad=True)
    gT1 = pt(gd["thole"][0])
    assert float(torch.linalg.norm(gT1)) > 1e-3
    for t in range(len(Tt)):
        h = 1e-5 * abs(float(Tt[t]))
        Tp, Tm = Tt.clone(), Tt.clone()
        Tp[t] += h
        Tm[t] -= h
        fd = (e_d1(Tp)[0] - e_d1(Tm)[0]) / (2 * h)
        print(f"dE/dthole[{t}] (dScales 1) analytic {float(gT1[t]):.9f} fd {fd:.9f}", flush=True)
        check(float(gT1[t]), fd, ("thole dS=1", t))


def test_type_reduce_deterministic():
    from dmff_amd.admp.pme import TypeReducer
    rng = np.random.default_rng(12)
    tmap = rng.integers(0, 4, size=12288)
    a = torch.as_tensor(rng.normal(size=(12288, 9)), device=DEV)
    tr = TypeReducer(tmap, device=DEV)
    r1, r2 = tr(a), tr(a)
    assert torch.equal(r1, r2)
    an = a.cpu().numpy()
    ref = np.stack([an[tmap == t].sum(0) for t in range(4)])
    assert np.allclose(r1.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)
