"""Subprocess entry: run one case through the CPU-emulated library and print its
results as one JSON line prefixed "RESULT ".  A guard-page fault ends the
process with status 86 after the handler has named the kernel, block and thread.

    python -m tests.cpu_emu.runner <libadmp_emu.so> <case> [precision] [grid K1,K2,K3 | waters/precond count]

The cases call the same Python interface as the GPU parity tests; only the
Engine's device set-up is replaced (CPU tensors, the null stream).
"""
import ctypes
import json
import sys

import numpy as np
import torch

F64 = torch.float64


def patch_engine(lib_path):
    import dmff_amd._lib as L
    from dmff_amd.admp import engine as E
    L.LIB_PATH = lib_path
    L.HOST_MEMORY = True  # the emulated library's device memory is host memory
    base_init = E.Engine.__init__

    def init(self, *args, **kwargs):
        kwargs["device"] = "cpu"
        torch_cuda = torch.cuda
        saved = (torch_cuda.device, torch_cuda.init)

        class _NoDevice:
            def __init__(self, dev):
                pass

            def __enter__(self):
                return self

            def __exit__(self, *exc):
                return False

        torch_cuda.device, torch_cuda.init = _NoDevice, (lambda: None)
        try:
            base_init(self, *args, **kwargs)
        finally:
            torch_cuda.device, torch_cuda.init = saved

    E.Engine.__init__ = init
    E.Engine._stream = lambda self: self.lib.admp_set_stream(self.h, ctypes.c_void_p(0))


def _lst(x):
    return np.asarray(x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x, dtype=np.float64).tolist()


def small_box(precision):
    """Polarisable PME on the 13.2 A water box, PCG to 1e-9, with every gradient."""
    from dmff_amd.admp.pme import ADMPPmeForce
    from tests.systems import reference_pairs, small_water_box
    s, pos, box = small_water_box()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    f = ADMPPmeForce(box, s.map_atomtype, s.axis_types, s.axis_idx, 6.0, 5e-4, s.pme.lmax, s.pme.lpol, True, None,
                     precision=precision, solver="pcg", tol=1e-9, device="cpu")
    T = lambda x: torch.tensor(np.asarray(x, dtype=np.float64), requires_grad=True)
    p, b, Q = T(pos), T(box), T(s.pme.Q_local[s.map_atomtype])
    a, t = T(s.pme.pol[s.map_poltype]), T(s.pme.thole[s.map_poltype])
    e, aux = f.get_energy(p, b, torch.as_tensor(pairs), Q, a, t, s.pme.mScales, s.pme.pScales, s.pme.dScales,
                          aux={})
    gp, gb, gq, ga, gt = torch.autograd.grad(e, (p, b, Q, a, t))
    return {"E": float(e.detach()), "pos": _lst(gp), "box": _lst(gb), "Q": _lst(gq), "pol": _lst(ga),
            "thole": _lst(gt), "U": _lst(aux["U_ind"] * 10.0), "n_iter": f.last_info["n_iter"]}


def frames(precision, n_frames=3):
    """Perturbed frames of the small box in one call (3 by default; 14 put the
    real-space product on the second stream), each frame's neighbour list built by
    the library, and frames 0, middle and last again alone."""
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import setup_ewald_parameters
    from tests.systems import small_water_box
    s, pos, box = small_water_box()
    rng = np.random.default_rng(3)
    P = np.stack([pos + rng.normal(0, 0.005, pos.shape) for _ in range(n_frames)])
    kappa, K1, K2, K3 = setup_ewald_parameters(6.0, 5e-4, box)
    Q = s.pme.Q_local[s.map_atomtype]
    pol, th = s.pme.pol[s.map_poltype], s.pme.thole[s.map_poltype]

    def run(Pf):
        F = len(Pf)
        eng = Engine(len(pos), 2, True, True, kappa, (K1, K2, K3), 6.0, n_frames=F, precision=precision)
        eng.set_topology(s.axis_types, s.axis_idx, s.bonded)
        eng.build_neighbors(Pf, np.stack([box] * F))
        E, g, U, info = eng.compute(Pf, np.stack([box] * F), Q, pol, th, s.pme.mScales, s.pme.pScales,
                                    s.pme.dScales, solver="pcg", tol=1e-8, maxiter=500, want_grad=True)
        return _lst(E), _lst(g["pos"])

    Eb, Gb = run(P)
    idx = sorted({0, n_frames // 2, n_frames - 1})
    single = [run(P[k:k + 1]) for k in idx]
    return {"E_batch": [Eb[k] for k in idx], "G_batch": [Gb[k] for k in idx], "E_single": [x[0][0] for x in single],
            "G_single": [x[1][0] for x in single]}


def disp_slater(precision):
    """Dispersion PME (pmax 10) and the Slater term on 60 random sites in a 20 A box."""
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import setup_ewald_parameters
    from tests.systems import reference_pairs
    rng = np.random.default_rng(9)
    L, n, rc = 20.0, 60, 9.0
    pos = rng.uniform(0, L, size=(n, 3))
    c = rng.uniform(5, 40, size=(n, 3))
    m = np.array([0.0, 0.0, 1.0, 1.0, 1.0, 1.0])
    bonded = {(0, 1): 1, (1, 0): 1, (2, 3): 2, (3, 2): 2}
    pairs = reference_pairs(pos, np.eye(3) * L, rc, bonded)
    kappa, K1, K2, K3 = setup_ewald_parameters(rc, 5e-4, np.eye(3) * L)
    eng = Engine(n, 0, False, False, 0.0, (0, 0, 0), rc, precision=precision, disp_kappa=kappa, dK=(K1, K2, K3),
                 disp_pmax=10)
    eng.set_topology(None, None, bonded)
    eng.set_pairs(torch.as_tensor(pairs))
    Ed, gd = eng.disp_compute(pos, np.eye(3) * L, c, m)
    A, B = rng.uniform(1e2, 3e2, size=n), rng.uniform(3.0, 4.5, size=n)
    Es, gs = eng.slater_compute(pos, np.eye(3) * L, A, B, m)
    out = {"E_disp": float(Ed[0]), "E_slater": float(Es[0])}
    out.update({"disp_" + k: _lst(v[0]) for k, v in gd.items()})
    out.update({"slater_" + k: _lst(v[0]) for k, v in gs.items()})
    return out


def neighbors(precision):
    """Library-built neighbour lists: the small water box at three cutoffs and 300
    random points in an 11 A box (fewer than three cells a side)."""
    from dmff_amd.admp.engine import Engine
    from tests.systems import small_water_box
    s, pos, box = small_water_box()
    out = {}
    for rc in (6.0, 4.0):
        e = Engine(len(pos), 0, False, False, 0.0, (0, 0, 0), rc)
        e.set_topology(None, None, s.bonded)
        e.build_neighbors(pos, box)
        out[f"water_{rc}"] = e.get_pairs(0).numpy().astype(np.int64).tolist()
    # capacity 8 per row: the first build overflows and is rebuilt before the lists are read
    e = Engine(len(pos), 0, False, False, 0.0, (0, 0, 0), 6.0, max_neighbors=8)
    e.set_topology(None, None, s.bonded)
    e.build_neighbors(pos, box)
    out["water_6.0_cap8"] = e.get_pairs(0).numpy().astype(np.int64).tolist()
    rng = np.random.default_rng(5)
    p = rng.uniform(0, 11.0, size=(300, 3))
    p[1] = p[0] + np.array([4.0, 0, 0])
    p[3] = p[2] + np.array([0, 3.9999999999, 0])
    for rc in (4.0, 5.4):
        e = Engine(len(p), 0, False, False, 0.0, (0, 0, 0), rc)
        e.set_topology(None, None, {})
        e.build_neighbors(p, np.eye(3) * 11.0)
        out[f"random_{rc}"] = e.get_pairs(0).numpy().astype(np.int64).tolist()
    from tests.systems import neighbour_cases
    for name, p, b, rc in neighbour_cases():
        e = Engine(len(p), 0, False, False, 0.0, (0, 0, 0), rc)
        e.set_topology(None, None, {})
        e.build_neighbors(p, b)
        out[name] = e.get_pairs(0).numpy().astype(np.int64).tolist()
    # one context, successive builds: the second takes the first cell grid over,
    # the shrunken box needs a coarser one (rebuilt at the next read), the
    # grown box keeps it (still valid)
    e = Engine(len(pos), 0, False, False, 0.0, (0, 0, 0), 4.0)
    e.set_topology(None, None, s.bonded)
    for name, sc in (("seq_1.0", 1.0), ("seq_1.0b", 1.0), ("seq_0.7", 0.7), ("seq_1.3", 1.3)):
        e.build_neighbors(pos * sc, box * sc)
        out[name] = e.get_pairs(0).numpy().astype(np.int64).tolist()
    return out


def grid(precision, K):
    """Non-polarisable PME on the small box on a given grid (FFT radix paths, the
    scatter path of grids under 9 points a side, line-major and plane passes)."""
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import setup_ewald_parameters
    from tests.systems import reference_pairs, small_water_box
    s, pos, box = small_water_box()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    kappa = setup_ewald_parameters(6.0, 5e-4, box)[0]
    eng = Engine(len(pos), 2, False, True, kappa, K, 6.0, precision=precision)
    eng.set_topology(s.axis_types, s.axis_idx, s.bonded)
    eng.set_pairs(torch.as_tensor(pairs))
    E, g, U, info = eng.compute(pos, box, s.pme.Q_local[s.map_atomtype], mScales=s.pme.mScales, want_grad=True)
    return {"E": float(E[0]), "pos": _lst(g["pos"][0])}


def waters(precision, n_waters):
    """A polarisable water box of n_waters (several solver blocks), neighbour list
    built by the library, PCG to 1e-6, then a warm-started second evaluation."""
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import setup_ewald_parameters
    from dmff_amd.admp.serving import water_spec
    spec, pos, box = water_spec("w", n_waters, 7)
    kappa, K1, K2, K3 = setup_ewald_parameters(spec.rc, 5e-4, np.asarray(box))
    eng = Engine(spec.n_atoms, 2, True, True, kappa, (K1, K2, K3), spec.rc, precision=precision)
    eng.set_topology(spec.axis_types, spec.axis_idx, spec.exclusions)
    eng.build_neighbors(pos, box)
    args = (spec.Q_local, spec.pol, spec.thole, *spec.scales)
    E, g, U, info = eng.compute(pos, box, *args, solver="pcg", tol=1e-6, maxiter=500, want_grad=True)
    E2, g2, U2, info2 = eng.compute(pos, box, *args, U_init=U, solver="pcg", tol=1e-6, maxiter=500,
                                    want_grad=True)
    return {"E": float(E[0]), "E_warm": float(E2[0]), "n_iter": info["n_iter"][0],
            "n_iter_warm": info2["n_iter"][0], "residual": info["residual"][0], "K": [K1, K2, K3],
            "rc": spec.rc, "n_atoms": spec.n_atoms}


def precond(precision, n_waters):
    """The same solve with the diagonal, the pair-block and the cluster-block preconditioner."""
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import setup_ewald_parameters
    from dmff_amd.admp.serving import water_spec
    spec, pos, box = water_spec("w", n_waters, 7)
    kappa, K1, K2, K3 = setup_ewald_parameters(spec.rc, 5e-4, np.asarray(box))
    eng = Engine(spec.n_atoms, 2, True, True, kappa, (K1, K2, K3), spec.rc, precision=precision)
    eng.set_topology(spec.axis_types, spec.axis_idx, spec.exclusions)
    eng.build_neighbors(pos, box)
    out = {}
    for pc in ("diagonal", "pairs", "clusters"):
        E, g, U, info = eng.compute(pos, box, spec.Q_local, spec.pol, spec.thole, *spec.scales, solver="pcg",
                                    tol=1e-6, maxiter=500, want_grad=True, precond=pc)
        out[pc] = {"E": float(E[0]), "pos": _lst(g["pos"][0]), "U": _lst(U[0]), "n_iter": info["n_iter"][0],
                   "residual": info["residual"][0]}
    return out


def type_reduce(precision):
    """Per-type sums of per-site arrays through the library (atoms on axis 0 and 1)."""
    from dmff_amd.admp.pme import TypeReducer
    rng = np.random.default_rng(11)
    tmap = rng.integers(0, 3, size=1000)
    a = rng.normal(size=(1000, 9))
    b = rng.normal(size=(2, 1000))
    tr = TypeReducer(tmap)
    return {"map": tmap.tolist(), "a": a.tolist(), "b": b.tolist(), "ra": _lst(tr(torch.as_tensor(a))),
            "rb": _lst(tr(torch.as_tensor(b), axis=1))}


def bad_geometry(precision):
    """Non-finite positions and degenerate boxes: every entry point returns an
    argument error (no kernel derives an index from them) and the context still
    evaluates the next good call correctly."""
    import dmff_amd._lib as L
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import setup_ewald_parameters
    from tests.systems import reference_pairs, small_water_box
    s, pos, box = small_water_box()
    pairs = reference_pairs(pos, box, 6.0, s.bonded)
    kappa, K1, K2, K3 = setup_ewald_parameters(6.0, 5e-4, box)
    eng = Engine(len(pos), 2, True, True, kappa, (K1, K2, K3), 6.0, precision=precision)
    eng.set_topology(s.axis_types, s.axis_idx, s.bonded)
    eng.set_pairs(torch.as_tensor(pairs))
    args = (s.pme.Q_local[s.map_atomtype], s.pme.pol[s.map_poltype], s.pme.thole[s.map_poltype],
            s.pme.mScales, s.pme.pScales, s.pme.dScales)
    out = {}

    def attempt(name, fn):
        try:
            fn()
            out[name] = "ok"
        except L.ADMPError as e:
            out[name] = str(e)

    bad_nan = pos.copy()
    bad_nan[5, 1] = np.nan
    bad_far = pos.copy()
    bad_far[7, 2] = 3.0e6
    flat = box.copy()
    flat[2] = flat[0] + flat[1] * 1e-9
    nanbox = box.copy()
    nanbox[1, 1] = np.inf
    kw = dict(solver="pcg", tol=1e-8, maxiter=200, want_grad=True)
    attempt("compute_nan_pos", lambda: eng.compute(bad_nan, box, *args, **kw))
    attempt("compute_nan_pos_nostats", lambda: eng.compute(bad_nan, box, *args, stats=False, **kw))
    attempt("compute_far_pos", lambda: eng.compute(bad_far, box, *args, **kw))
    attempt("compute_flat_box", lambda: eng.compute(pos, flat, *args, **kw))
    attempt("compute_inf_box", lambda: eng.compute(pos, nanbox, *args, **kw))
    m = np.array([0.0, 0.0, 1.0, 1.0, 1.0, 1.0])
    attempt("slater_nan_pos", lambda: eng.slater_compute(bad_nan, box, np.ones(len(pos)), np.ones(len(pos)), m))
    attempt("build_flat_box", lambda: eng.build_neighbors(pos, flat))
    # a neighbour build from bad positions reports at the next call that reads its lists
    eng.build_neighbors(bad_nan, box)
    attempt("after_bad_build", lambda: eng.get_pairs(0))
    eng.set_pairs(torch.as_tensor(pairs))
    E, g, U, info = eng.compute(pos, box, *args, **kw)
    out["E_good"] = float(E[0])
    out["pos_good"] = _lst(g["pos"][0])
    return out


def pairwise(precision):
    """admp_pairwise_compute for every short-range kind on its test system."""
    from dmff_amd.admp.engine import Engine
    from tests.systems import pairwise_case
    out = {}
    for kind in (1, 2, 3, 4):
        pos, box, pairs, m, prm = pairwise_case(kind)
        eng = Engine(len(pos), 0, False, False, 0.0, (0, 0, 0), 9.0, precision=precision)
        eng.set_topology(None, None, {(0, 1): 1, (1, 0): 1, (2, 3): 2, (3, 2): 2})
        eng.set_pairs(torch.as_tensor(pairs))
        E, g = eng.pairwise_compute(kind, pos, box, prm, m)
        out[str(kind)] = {"E": float(E[0]), "pos": _lst(g["pos"][0]), "box": _lst(g["box"][0]),
                          "prm": _lst(g["prm"][0])}
    return out


def generators(precision):
    """PEG2 through Hamiltonian -> createPotential -> potential_fn (all nine
    generators) and the water-dimer neighbour list, on the emulated library."""
    from dmff_amd.admp.hamiltonian import Hamiltonian
    from dmff_amd.admp.nblist import NeighborList
    from dmff_amd.admp import frontend as fe
    from tests.systems import golden_path
    H = Hamiltonian(golden_path("peg.xml"))
    pots = H.createPotential(golden_path("peg2.pdb"), nonbondedMethod="CutoffPeriodic", nonbondedCutoff=1.49,
                             ethresh=1e-4, device="cpu", precision=precision)
    pos_nm, top = fe.read_pdb(golden_path("peg2.pdb"))
    nbl = NeighborList(top.box_nm, 1.49, pots.meta["cov_map"], device="cpu")
    nbl.allocate(pos_nm)
    pairs = nbl.pairs
    out = {"pairs": _lst(pairs)}
    P = torch.tensor(pos_nm, dtype=F64, requires_grad=True)
    B = torch.tensor(top.box_nm, dtype=F64)
    for name, fn in pots.dmff_potentials.items():
        e = fn(P, B, pairs, H.paramset)
        (g,) = torch.autograd.grad(e, P)
        out[name] = {"E": float(e.detach()), "pos": _lst(g)}
    return out


def _arr(d):
    return {k: (np.asarray(v).tolist() if isinstance(v, np.ndarray) else v) for k, v in d.items()}


def axis_types(precision):
    """Every local-frame type (tests.systems.axis_type_system), polarisable PME."""
    from tests.systems import axis_type_system, library_pme, reference_pairs
    s = axis_type_system()
    pairs = reference_pairs(s["pos"], s["box"], 7.0, s["bonded"])
    return _arr(library_pme(s["pos"], s["box"], s["axis_types"], s["axis_idx"], s["bonded"], s["Q"], s["pol"],
                            s["thole"], (s["mScales"], s["pScales"], s["dScales"]), 7.0, True, precision, "cpu",
                            pairs=pairs, tol=1e-9))


def triclinic(precision):
    """Polarisable waters in a triclinic cell, neighbour list built by the library."""
    from tests.systems import library_pme, triclinic_water_system
    S, pos, box = triclinic_water_system()
    return _arr(library_pme(pos, box, S.axis_types, S.axis_idx, S.bonded, S.Q_local, S.pol, S.thole,
                            (S.mScales, S.pScales, S.dScales), 6.0, True, precision, "cpu", tol=1e-9))


def coul(precision):
    """CoulombPMEForce (lmax 0, gromacs kappa) on the methane + water fixture."""
    from dmff_amd.admp.coul import CoulombPMEForce
    from tests.test_oracle_pins import coul_pme_system
    s = coul_pme_system()
    f = CoulombPMEForce(0.5, s["charges"], s["kappa_nm"], s["K"], precision=precision, device="cpu")
    P = torch.tensor(s["pos_nm"], dtype=F64, requires_grad=True)
    q = torch.tensor(s["charges"], dtype=F64, requires_grad=True)
    f.init_charges = q
    e = f.generate_get_energy()(P, torch.tensor(s["box_nm"], dtype=F64), torch.as_tensor(s["pairs"]), s["mscales"])
    gp, gq = torch.autograd.grad(e, (P, q))
    return {"E": float(e.detach()), "pos": _lst(gp), "q": _lst(gq)}


def esp(precision):
    """generate_esp of the water (admp_charge.xml) at the ion position, and the
    library's energy difference E(water + ion) - E(water)."""
    from dmff_amd.admp.hamiltonian import Hamiltonian
    from dmff_amd.admp import frontend as fe
    from tests.systems import golden_path
    out = {}
    E = {}
    for name, pdb in (("water", "water.pdb"), ("probe", "water_ion.pdb")):
        H = Hamiltonian(golden_path("admp_charge.xml"))
        pot = H.createPotential(golden_path(pdb), nonbondedMethod="NoCutoff", nonbondedCutoff=10.0, ethresh=5e-4,
                                step_pol=1, device="cpu")
        pos, top = fe.read_pdb(golden_path(pdb))
        n = len(pos)
        cm = pot.meta["cov_map"]
        pairs = np.array([[i, j, cm[i, j]] for i in range(n) for j in range(i + 1, n)])
        E[name] = float(pot.getPotentialFunc()(torch.tensor(pos), torch.eye(3, dtype=F64) * 100.0, pairs,
                                               H.paramset))
        if name == "water":
            es = H.generators["ADMPPmeForce"].pme_force.generate_esp()
            water_pos, water_H = pos, H
        else:
            grid = pos[3:4]
    out["esp"] = float(es(torch.tensor(water_pos), torch.tensor(grid), water_H.paramset["ADMPPmeForce"]["Q_local"])[0])
    out["dE"] = E["probe"] - E["water"]
    return out


def cfg3s(precision):
    """Configuration 3's recipe at 384 waters: electrostatic energy and its
    components (real, reciprocal, self, polarisation penalty) in one precision."""
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import setup_ewald_parameters
    from tests.systems import cfg3_recipe_system
    S, pos, box, cl, A, B = cfg3_recipe_system(n_waters=384)
    kappa, K1, K2, K3 = setup_ewald_parameters(10.0, 5e-4, box)
    eng = Engine(S.n_atoms, 2, True, True, kappa, (K1, K2, K3), 10.0, precision=precision)
    eng.set_topology(S.axis_types, S.axis_idx, S.bonded)
    eng.build_neighbors(pos, box)
    E, g, U, info = eng.compute(pos, box, S.Q_local, S.pol, S.thole, S.mScales, S.pScales, S.dScales, solver="pcg",
                                tol=1e-8, maxiter=1000, want_grad=True)
    return {"E": float(E[0]), "e_real": info["e_real"], "e_recip": info["e_recip"], "e_self": info["e_self"],
            "e_pol": info["e_pol"], "n_iter": info["n_iter"][0], "pos": _lst(g["pos"][0])}


def qeq(variant):
    """ADMPQeqForce on the small charge-equilibration case: periodic (damp_mod 3,
    slab correction) or not (damp_mod 2): energy, charges, and gradients with
    respect to positions, box, chi, J and eta."""
    from dmff_amd.admp.qeq import ADMPQeqForce
    from tests.systems import qeq_small_system
    pbc = variant == "pbc"
    s = qeq_small_system(pbc)
    f = ADMPQeqForce(s["init_q"], s["rc_nm"], s["kappa"], s["K"], damp_mod=3 if pbc else 2,
                     const_list=s["const_list"], const_vals=s["const_vals"], slab_flag=pbc, pbc_flag=pbc,
                     has_aux=True, device="cpu")
    P = torch.tensor(s["pos_nm"], dtype=F64, requires_grad=True)
    B = torch.tensor(s["box_nm"], dtype=F64, requires_grad=True)
    chi, J, eta = (torch.tensor(s[k], dtype=F64, requires_grad=True) for k in ("chi", "J", "eta"))
    e, aux = f.generate_get_energy()(P, B, torch.as_tensor(s["pairs"]), s["mscales"], eta, chi, J, {
        "q": torch.zeros(len(s["init_q"]), dtype=F64), "lagmt": torch.ones(len(s["const_list"]), dtype=F64)})
    g = torch.autograd.grad(e, (P, B, chi, J, eta))
    return {"E": float(e.detach()), "q": _lst(aux["q"]), "iters": f.last_iterations,
            "pos": _lst(g[0]), "box": _lst(g[1]), "chi": _lst(g[2]), "J": _lst(g[3]), "eta": _lst(g[4])}


CASES = {"qeq": qeq, "cfg3s": cfg3s, "axis_types": axis_types, "triclinic": triclinic, "coul": coul, "esp": esp, "generators": generators, "pairwise": pairwise, "bad_geometry": bad_geometry, "type_reduce": type_reduce, "small_box": small_box, "frames": frames, "disp_slater": disp_slater, "neighbors": neighbors}

if __name__ == "__main__":
    lib, case = sys.argv[1], sys.argv[2]
    prec = sys.argv[3] if len(sys.argv) > 3 else "fp64"
    patch_engine(lib)
    if case == "grid":
        out = grid(prec, tuple(int(k) for k in sys.argv[4].split(",")))
    elif case == "waters":
        out = waters(prec, int(sys.argv[4]))
    elif case == "frames" and len(sys.argv) > 4:
        out = frames(prec, int(sys.argv[4]))
    elif case == "precond":
        out = precond(prec, int(sys.argv[4]))
    else:
        out = CASES[case](prec)
    print("RESULT " + json.dumps(out), flush=True)
