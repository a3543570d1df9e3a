"""The library's kernels run on the CPU under an emulation of the HIP runtime
(tests/cpu_emu): one fibre per GPU thread, wave operations resolved per
64-lane wave, and every device buffer placed against a guard page, so that a
read or write past the end of any buffer faults at the access and the run
ends with status 86 after naming the kernel, block and thread.

Each case runs in a subprocess through the same Python interface as the GPU
parity tests and is checked against the oracle to the same tolerances:
  FP64   energy rel 1e-9, gradient rel RMS 1e-7
  mixed  energy rel 2e-5, force rel RMS 1e-4
A clean exit shows that no kernel touched memory outside its buffers on these
inputs (solver partials, neighbour rows, grids, FFT scratch included).
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest
import torch

from oracle import admp_oracle as O
from tests.systems import reference_pairs, small_water_box
from tests.test_gpu_parity import compare, oracle_eval, rel_rms

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def T(x):
    return torch.tensor(np.asarray(x, dtype=np.float64))


@pytest.fixture(scope="module")
def emu_lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is needed to build the CPU emulation")
    from tests.cpu_emu.build import build
    return build(str(tmp_path_factory.mktemp("emu")))


def run(lib, *args, timeout=900, extra_env=None):
    env = dict(os.environ, PYTHONPATH=REPO + os.pathsep + os.environ.get("PYTHONPATH", ""))
    env.update(extra_env or {})
    p = subprocess.run([sys.executable, "-m", "tests.cpu_emu.runner", lib, *[str(a) for a in args]], cwd=REPO,
                       env=env, capture_output=True, text=True, timeout=timeout)
    tail = (p.stdout[-2000:] + p.stderr[-3000:])
    assert p.returncode == 0, f"{args}: exit status {p.returncode} (86 = guard-page fault)\n{tail}"
    lines = [l for l in p.stdout.splitlines() if l.startswith("RESULT ")]
    assert lines, tail
    return json.loads(lines[-1][len("RESULT "):])


def sorted_pairs(a):
    a = np.asarray(a, dtype=np.int64).reshape(-1, 3)
    return a[np.lexsort((a[:, 1], a[:, 0]))]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_small_box_polarisable_pcg(emu_lib, prec):
    s, pos, box = small_water_box()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    o, _ = oracle_eval(s, pos, box, pairs, 6.0, True, solver="cg", tol=1e-9)
    g = run(emu_lib, "small_box", prec)
    if prec == "fp64":
        ok, msgs = compare(g, o)
        assert ok and np.abs(np.asarray(g["U"]) - o["U"]).max() < 1e-8, msgs
    else:
        ok, msgs = compare(g, o, etol=2e-5, gtol=1e-4, keys=("pos",))
        assert ok, msgs


@pytest.mark.parametrize("n_frames", [3, 14])
def test_batched_frames_with_library_neighbour_lists(emu_lib, n_frames):
    """3 frames: the real-space product rides in the plane launch; 14 frames (140
    planes): it runs on the second stream."""
    r = run(emu_lib, "frames", "fp64", n_frames)
    for k in range(3):
        assert abs(r["E_batch"][k] - r["E_single"][k]) < 1e-9 * abs(r["E_single"][k])
        assert rel_rms(r["G_batch"][k], r["G_single"][k]) < 1e-8


def test_dispersion_and_slater(emu_lib):
    rng = np.random.default_rng(9)
    L, n, rc = 20.0, 60, 9.0
    pos = rng.uniform(0, L, size=(n, 3))
    c = rng.uniform(5, 40, size=(n, 3))
    m = np.array([0.0, 0.0, 1.0, 1.0, 1.0, 1.0])
    bonded = {(0, 1): 1, (1, 0): 1, (2, 3): 2, (3, 2): 2}
    pairs = reference_pairs(pos, np.eye(3) * L, rc, bonded)
    kappa, K1, K2, K3 = O.setup_ewald_parameters(rc, 5e-4, np.eye(3) * L)
    p, b, ct = T(pos).requires_grad_(True), T(np.eye(3) * L).requires_grad_(True), T(c).requires_grad_(True)
    e = O.energy_disp_pme(p, b, pairs, ct, m, kappa, (K1, K2, K3), 10)
    gp, gb, gc = torch.autograd.grad(e, (p, b, ct))
    e = e.detach()
    A, B = rng.uniform(1e2, 3e2, size=n), rng.uniform(3.0, 4.5, size=n)
    At, Bt = T(A).requires_grad_(True), T(B).requires_grad_(True)
    p2, b2 = T(pos).requires_grad_(True), T(np.eye(3) * L).requires_grad_(True)
    es = O.slater_sr_energy(p2, b2, pairs, m, At, Bt)
    gp2, gb2, ga2, gbb2 = torch.autograd.grad(es, (p2, b2, At, Bt))
    es = es.detach()
    r = run(emu_lib, "disp_slater", "fp64")
    assert abs(r["E_disp"] - float(e)) < 1e-9 * abs(float(e))
    assert rel_rms(r["disp_pos"], gp.numpy()) < 1e-8
    assert rel_rms(r["disp_box"], gb.numpy()) < 1e-8
    assert rel_rms(r["disp_c"], gc.numpy()) < 1e-8
    assert abs(r["E_slater"] - float(es)) < 1e-10 * max(1.0, abs(float(es)))
    assert rel_rms(r["slater_pos"], gp2.numpy()) < 1e-9
    assert rel_rms(r["slater_box"], gb2.numpy()) < 1e-9
    assert rel_rms(r["slater_a"], ga2.numpy()) < 1e-9
    assert rel_rms(r["slater_b"], gbb2.numpy()) < 1e-9


def test_neighbour_lists_bit_exact(emu_lib):
    r = run(emu_lib, "neighbors", "fp64")
    s, pos, box = small_water_box()
    for rc in (6.0, 4.0):
        ref = sorted_pairs(reference_pairs(pos, box, rc, s.bonded))
        got = sorted_pairs(r[f"water_{rc}"])
        assert ref.shape == got.shape and np.array_equal(ref, got), rc
    ref = sorted_pairs(reference_pairs(pos, box, 6.0, s.bonded))
    got = sorted_pairs(r["water_6.0_cap8"])
    assert ref.shape == got.shape and np.array_equal(ref, got)
    rng = np.random.default_rng(5)
    p = rng.uniform(0, 11.0, size=(300, 3))
    p[1] = p[0] + np.array([4.0, 0, 0])
    p[3] = p[2] + np.array([0, 3.9999999999, 0])
    for rc in (4.0, 5.4):
        ref = sorted_pairs(reference_pairs(p, np.eye(3) * 11.0, rc, {}))
        got = sorted_pairs(r[f"random_{rc}"])
        assert ref.shape == got.shape and np.array_equal(ref, got), rc
    from tests.systems import neighbour_cases
    for name, p, b, rc in neighbour_cases():
        ref = sorted_pairs(reference_pairs(p, b, rc, {}))
        got = sorted_pairs(r[name])
        assert len(ref) > 0 and ref.shape == got.shape and np.array_equal(ref, got), name
    for name, sc in (("seq_1.0", 1.0), ("seq_1.0b", 1.0), ("seq_0.7", 0.7), ("seq_1.3", 1.3)):
        ref = sorted_pairs(reference_pairs(pos * sc, box * sc, 4.0, s.bonded))
        got = sorted_pairs(r[name])
        assert len(ref) > 0 and ref.shape == got.shape and np.array_equal(ref, got), name


@pytest.mark.parametrize("K,prec", [((104, 104, 104), "fp64"), ((104, 104, 104), "mixed"), ((6, 7, 8), "fp64"),
                                    ((20, 24, 15), "fp64"), ((32, 27, 25), "mixed"), ((50, 30, 40), "fp64"),
                                    ((10, 70, 50), "mixed")])
def test_grid_paths(emu_lib, K, prec):
    """104^3 in FP64 takes the line-major FFT passes (its plane exceeds the LDS
    limit); in mixed precision the fused plane pass.  50, 30, 40, 10 and 70 have
    radix-10 stages (with 5, 3, 4 and 7 beside them)."""
    s, pos, box = small_water_box()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    kappa = O.setup_ewald_parameters(6.0, 5e-4, box)[0]
    p = T(pos).requires_grad_(True)
    e = O.energy_pme(p, T(box), T(pairs).long(), T(s.pme.Q_local[s.map_atomtype]), None, None, None,
                     s.pme.mScales, None, None, s.axis_types, s.axis_idx, kappa, K, 2, False, True)
    (gp,) = torch.autograd.grad(e, (p,))
    e = e.detach()
    r = run(emu_lib, "grid", prec, ",".join(str(k) for k in K))
    etol, ftol = (1e-9, 1e-7) if prec == "fp64" else (2e-5, 1e-4)
    assert abs(r["E"] - float(e)) < etol * abs(float(e))
    assert rel_rms(r["pos"], gp.numpy()) < ftol


def test_water_box_several_solver_blocks(emu_lib):
    """1,536 atoms: six solver blocks, the fused gather epilogue and a warm start."""
    r64 = run(emu_lib, "waters", "fp64", 512)
    rmx = run(emu_lib, "waters", "mixed", 512)
    assert r64["residual"] < 1e-6 and r64["n_iter"] > 5
    assert r64["n_iter_warm"] <= 1 and abs(r64["E_warm"] - r64["E"]) < 1e-9 * abs(r64["E"])
    assert abs(rmx["E"] - r64["E"]) < 2e-5 * abs(r64["E"])


def test_block_preconditioners_same_solution_fewer_iterations(emu_lib):
    """512 waters: the pair-block and the cluster-block preconditioners converge
    to the same induced dipoles as the diagonal one (all to max|dE/dU| < 1e-6),
    the pairs in fewer iterations than the diagonal, the clusters in no more
    than the pairs."""
    r = run(emu_lib, "precond", "fp64", 512)
    d, p, c = r["diagonal"], r["pairs"], r["clusters"]
    print({k: v["n_iter"] for k, v in r.items()})
    assert d["residual"] < 1e-6 and p["residual"] < 1e-6 and c["residual"] < 1e-6
    assert p["n_iter"] < d["n_iter"], (p["n_iter"], d["n_iter"])
    assert c["n_iter"] <= p["n_iter"], (c["n_iter"], p["n_iter"])
    for x in (p, c):
        assert abs(x["E"] - d["E"]) < 1e-12 * abs(d["E"])
        assert rel_rms(x["pos"], d["pos"]) < 1e-8
        assert np.abs(np.asarray(x["U"]) - np.asarray(d["U"])).max() < 1e-8


def test_type_reduce(emu_lib):
    r = run(emu_lib, "type_reduce", "fp64")
    m, a, b = np.asarray(r["map"]), np.asarray(r["a"]), np.asarray(r["b"])
    ra = np.stack([a[m == t].sum(0) for t in range(3)])
    rb = np.stack([b[:, m == t].sum(1) for t in range(3)], axis=1)
    assert np.allclose(r["ra"], ra, rtol=1e-12, atol=1e-12)
    assert np.allclose(r["rb"], rb, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("variant", ["list", "col"])
def test_small_box_other_spread_variants(emu_lib, variant):
    """The one-point-per-lane and four-bricks-per-wavefront spreads (not the
    default) against the oracle on the polarisable small box."""
    s, pos, box = small_water_box()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    o, _ = oracle_eval(s, pos, box, pairs, 6.0, True, solver="cg", tol=1e-9)
    g = run(emu_lib, "small_box", "fp64", extra_env={"ADMP_SPREAD": variant})
    ok, msgs = compare(g, o)
    assert ok, msgs


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_bad_geometry_is_an_argument_error(emu_lib, prec):
    r = run(emu_lib, "bad_geometry", prec)
    for k in ("compute_nan_pos", "compute_nan_pos_nostats", "compute_far_pos", "slater_nan_pos", "after_bad_build"):
        assert "status 1" in r[k] and "positions" in r[k], (k, r[k])
    for k in ("compute_flat_box", "compute_inf_box", "build_flat_box"):
        assert "status 1" in r[k] and "box" in r[k], (k, r[k])
    s, pos, box = small_water_box()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    o, _ = oracle_eval(s, pos, box, pairs, 6.0, True, solver="cg", tol=1e-9)
    tol_e, tol_g = (1e-9, 1e-6) if prec == "fp64" else (2e-5, 1e-4)
    assert abs(r["E_good"] - o["E"]) <= tol_e * abs(o["E"])
    assert rel_rms(np.asarray(r["pos_good"]), o["pos"]) <= tol_g


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_short_range_kinds_vs_oracle(emu_lib, prec):
    from tests.systems import pairwise_case, pairwise_oracle
    r = run(emu_lib, "pairwise", prec)
    tol_e, tol_g = (1e-11, 1e-10) if prec == "fp64" else (1e-5, 1e-4)
    for kind in (1, 2, 3, 4):
        o = pairwise_oracle(kind, *pairwise_case(kind))
        g = r[str(kind)]
        assert abs(g["E"] - o["E"]) <= tol_e * abs(o["E"]), (kind, g["E"], o["E"])
        for k in ("pos", "box", "prm"):
            assert rel_rms(np.asarray(g[k]), o[k]) <= tol_g, (kind, k, rel_rms(np.asarray(g[k]), o[k]))


def test_generators_peg2(emu_lib):
    """Hamiltonian -> createPotential -> potential_fn for the nine PEG2 generators
    (the reference's test_slater values to decimal 4, the oracle to 1e-9)."""
    import json as _json
    from tests.systems import GOLD
    from tests.test_oracle_pins import peg2_oracle_terms
    r = run(emu_lib, "generators", "fp64")
    ref = _json.load(open(f"{GOLD}/energies.json"))["peg2"]
    for name, want in zip(ref["forces"], ref["values"]):
        assert abs(r[name]["E"] - want) < 1.5e-4, (name, r[name]["E"], want)
    s_terms = peg2_oracle_terms()
    for name in ref["forces"]:
        o = float(s_terms[name])
        assert abs(r[name]["E"] - o) <= 1e-9 * abs(o) + 1e-12, (name, r[name]["E"], o)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_every_axis_type_vs_oracle(emu_lib, prec):
    from tests.systems import axis_type_system, oracle_pme, reference_pairs, rel_err
    s = axis_type_system()
    pairs = reference_pairs(s["pos"], s["box"], 7.0, s["bonded"])
    o = oracle_pme(s["pos"], s["box"], pairs, s["axis_types"], s["axis_idx"], s["Q"], s["pol"], s["thole"],
                   (s["mScales"], s["pScales"], s["dScales"]), 7.0, True)
    g = run(emu_lib, "axis_types", prec)
    te, tg = (1e-9, 1e-7) if prec == "fp64" else (2e-5, 1e-4)
    assert abs(g["E"] - o["E"]) < te * abs(o["E"])
    keys = ("pos", "box", "Q", "pol", "thole") if prec == "fp64" else ("pos",)
    for k in keys:
        assert rel_err(g[k], o[k]) < tg, (k, rel_err(g[k], o[k]))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_triclinic_box_vs_oracle(emu_lib, prec):
    from tests.systems import oracle_pme, reference_pairs, rel_err, triclinic_water_system
    S, pos, box = triclinic_water_system()
    ref = reference_pairs(pos, box, 6.0, S.bonded)
    g = run(emu_lib, "triclinic", prec)
    got = sorted_pairs(g["pairs"])
    assert np.array_equal(got, sorted_pairs(ref))
    o = oracle_pme(pos, box, ref, S.axis_types, S.axis_idx, S.Q_local, S.pol, S.thole,
                   (S.mScales, S.pScales, S.dScales), 6.0, True)
    te, tg = (1e-9, 1e-7) if prec == "fp64" else (2e-5, 1e-4)
    assert abs(g["E"] - o["E"]) < te * abs(o["E"])
    keys = ("pos", "box", "Q", "pol", "thole") if prec == "fp64" else ("pos",)
    for k in keys:
        assert rel_err(g[k], o[k]) < tg, (k, rel_err(g[k], o[k]))


def test_coulomb_pme_vs_oracle(emu_lib):
    from tests.test_oracle_pins import coul_pme_system, EN as _EN
    s = coul_pme_system()
    r = run(emu_lib, "coul", "fp64")
    p = torch.tensor(s["pos_nm"] * 10, requires_grad=True)
    q = torch.tensor(s["charges"], requires_grad=True)
    e = O.energy_pme(p, T(s["box_nm"] * 10), s["pairs"], q.reshape(-1, 1), None, None, None, s["mscales"], None,
                     None, None, None, s["kappa_nm"] / 10, s["K"], 0, False)
    gp, gq = torch.autograd.grad(e, (p, q))
    assert abs(r["E"] - float(e)) < 1e-9 * abs(float(e))
    assert abs(r["E"] - _EN["coul_pme"]["value"]) < _EN["coul_pme"]["atol"]
    assert rel_rms(np.asarray(r["pos"]), gp.numpy() * 10) < 1e-8
    assert rel_rms(np.asarray(r["q"]), gq.numpy()) < 1e-8


def test_esp_consistency(emu_lib):
    r = run(emu_lib, "esp", "fp64")
    assert abs(r["dE"] - r["esp"]) < 1e-3 * abs(r["esp"]), r


@pytest.mark.parametrize("variant", ["pbc", "nopbc"])
def test_qeq_vs_oracle(emu_lib, variant):
    """ADMPQeqForce (MINRES on the library's lmax-0 Coulomb) against the oracle's
    dense solve of the same stationary point: periodic with damp_mod 3 and the
    slab correction, and the plain pair sum with damp_mod 2.  Energy 1e-12,
    charges 1e-8, gradients (positions, box, chi, J, eta) 1e-7 relative."""
    from tests.systems import qeq_small_system
    pbc = variant == "pbc"
    r = run(emu_lib, "qeq", variant)
    s = qeq_small_system(pbc)
    f = O.ADMPQeqForce(s["init_q"], s["rc_nm"], s["kappa"], s["K"], 3 if pbc else 2, s["const_list"],
                       s["const_vals"], neutral_flag=True, slab_flag=pbc, constQ=True, pbc_flag=pbc)
    P, B = T(s["pos_nm"]).requires_grad_(True), T(s["box_nm"]).requires_grad_(True)
    chi, J, eta = (T(s[k]).requires_grad_(True) for k in ("chi", "J", "eta"))
    e, q, lag = f.get_energy(P, B, s["pairs"], s["mscales"], eta, chi, J)
    g = torch.autograd.grad(e, (P, B, chi, J, eta), allow_unused=True)
    assert abs(r["E"] - float(e.detach())) < 1e-12 * abs(float(e.detach()))
    assert np.abs(np.asarray(r["q"]) - q.numpy()).max() < 1e-8
    for k, gg in zip(("pos", "box", "chi", "J", "eta"), g):
        if gg is None:
            assert np.abs(np.asarray(r[k])).max() == 0.0, k
        else:
            assert rel_rms(np.asarray(r[k]), gg.numpy()) < 1e-7, k
