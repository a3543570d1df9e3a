"""Test-side helpers: load the reference's golden systems through the host
frontend and build reference-format pair lists with the oracle's brute force."""
import math
import os

import numpy as np

from dmff_amd.admp import frontend as fe

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DATA = os.path.join(GOLD, "data")


def golden_path(name):
    return os.path.join(DATA, name)


class System:
    pass


def load_system(ffxml, pdb, residues=None):
    ff = fe.ForceFieldXML(golden_path(ffxml))
    pos_nm, top = fe.read_pdb(golden_path(pdb))
    extra = fe.load_residue_bonds(golden_path(residues)) if residues else None
    fe.assign_types(top, ff, extra)
    pme = fe.ADMPPmeParams(ff)
    bonded = fe.bonded_distances(top.n_atoms, top.bonds)
    axis_types, axis_idx = fe.assign_axes(pme, top, ff, bonded)
    map_atomtype, map_poltype = fe.type_maps(pme, top, ff)
    s = System()
    s.ff, s.top, s.pme = ff, top, pme
    s.pos_nm = pos_nm
    s.box_nm = top.box_nm
    s.bonded = bonded
    s.axis_types, s.axis_idx = axis_types, axis_idx
    s.map_atomtype, s.map_poltype = map_atomtype, map_poltype
    s.n = top.n_atoms
    return s


def nbond_column(pairs, bonded):
    """cov_map[i, j] for each pair (dmff/common/nblist.py:22-25); 0 if not within 6 bonds."""
    return np.array([bonded.get((int(i), int(j)), 0) for i, j in pairs], dtype=np.int64)


def reference_pairs(pos, box, rc, bonded, pad=0):
    """(Np,3) [i, j, nbond] from the oracle's brute force, optionally padded with
    (N, N, 0) rows as the reference NeighborList does (dmff/common/nblist.py:42-55)."""
    from oracle.admp_oracle import brute_force_pairs
    p = brute_force_pairs(pos, box, rc)
    nb = nbond_column(p, bonded)
    out = np.concatenate([p, nb[:, None]], axis=1)
    if pad:
        n = pos.shape[0]
        out = np.concatenate([out, np.tile(np.array([[n, n, 0]]), (pad, 1))], axis=0)
    return out


def small_water_box(ff="fullpol_forcefield.xml"):
    """The waters of the 31 A test box that lie wholly inside a 13.2 A corner,
    shifted into a 13.2 A periodic box (44 molecules, 132 atoms)."""
    s = load_system(ff, "waterbox_31ang.pdb", "fullpol_residues.xml")
    p = s.pos_nm * 10
    mol = p.reshape(-1, 3, 3)
    ok = np.all((mol > 1.0) & (mol < 13.0), axis=(1, 2))
    idx = np.nonzero(np.repeat(ok, 3))[0]
    remap = {int(a): k for k, a in enumerate(idx)}
    pos = p[idx] - 0.4
    s.bonded = {(remap[i], remap[j]): v for (i, j), v in s.bonded.items() if i in remap and j in remap}
    s.axis_types = s.axis_types[idx]
    s.axis_idx = np.array([[remap.get(int(a), -1) if a >= 0 else -1 for a in row] for row in s.axis_idx[idx]])
    s.map_atomtype = s.map_atomtype[idx]
    s.map_poltype = s.map_poltype[idx]
    return s, pos, np.eye(3) * 13.2


TRICLINIC_BOX = np.array([[20.0, 0.0, 0.0], [5.0, 19.0, 0.0], [-4.0, 3.0, 18.0]])


def neighbour_cases():
    """Extra neighbour-list cases beyond the water boxes: (name, positions, box, rc).
      small_rc8   192 random points in a 12.4 A cube at rc 8 A (rc > L/2: the
                  reference's single minimum image, one cell a side)
      tric_*      400 points in a triclinic box (rows a, b, c), rc 6 A and
                  rc 9.4 A (above half the smallest height, 17.1 A), with a pair at
                  rc to 1e-10 A."""
    rng = np.random.default_rng(17)
    out = []
    p = rng.uniform(0, 12.4, size=(192, 3))
    out.append(("small_rc8", p, np.eye(3) * 12.4, 8.0))
    s = rng.uniform(0, 1, size=(400, 3))
    p = s @ TRICLINIC_BOX
    d = np.array([1.0, 2.0, -2.0]) / 3.0
    p[1] = p[0] + d * 5.9999999999
    for rc in (6.0, 9.4):
        out.append((f"tric_{rc}", p, TRICLINIC_BOX.copy(), rc))
    return out


def pairwise_case(kind, seed=21):
    """60 sites on a jittered 4x4x4 lattice (5 A spacing, +-1 A) in a 20 A box,
    pairs within 9 A with two bonded pairs, scales (0.1, 0.3, 1, 1, 1, 1), and
    per-site parameters of a short-range kind (admp_sr_kind) in realistic ranges
    of each kernel's own units.  Returns (pos, box, pairs, mScales, prm (N, n_prm))."""
    rng = np.random.default_rng(seed + 100 * kind)
    g = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), axis=-1).reshape(-1, 3)[:60]
    pos = g * 5.0 + 2.5 + rng.uniform(-1.0, 1.0, size=(60, 3))
    box = np.eye(3) * 20.0
    bonded = {(0, 1): 1, (1, 0): 1, (2, 3): 2, (3, 2): 2}
    pairs = reference_pairs(pos, box, 9.0, bonded)
    m = np.array([0.1, 0.3, 1.0, 1.0, 1.0, 1.0])
    n = len(pos)
    if kind == 1:    # slater: A, B (1/A)
        prm = np.stack([rng.uniform(1e2, 3e2, n), rng.uniform(3.0, 4.5, n)], axis=1)
    elif kind == 2:  # TT qq c6: A (hartree), B (1/bohr), Q (e), C (sqrt(kJ/mol A^6))
        prm = np.stack([rng.uniform(1.0, 20.0, n), rng.uniform(1.5, 2.5, n), rng.uniform(-0.8, 0.8, n),
                        rng.uniform(5.0, 40.0, n)], axis=1)
    elif kind == 3:  # TT qq: B (1/A), Q (e)
        prm = np.stack([rng.uniform(3.0, 4.5, n), rng.uniform(-0.8, 0.8, n)], axis=1)
    else:            # slater damping: B (1/A), c6, c8, c10 (sqrt of kJ/mol A^n)
        prm = np.stack([rng.uniform(3.0, 4.5, n), rng.uniform(5.0, 40.0, n), rng.uniform(10.0, 100.0, n),
                        rng.uniform(50.0, 300.0, n)], axis=1)
    return pos, box, pairs, m, prm


def pairwise_oracle(kind, pos, box, pairs, m, prm):
    """The oracle's generate_pairwise_interaction for `kind`: energy and gradients
    with respect to positions, box and the per-site parameters."""
    import torch
    from oracle import admp_oracle as O
    T = lambda x: torch.tensor(np.asarray(x, dtype=np.float64), requires_grad=True)
    p, b, q = T(pos), T(box), T(prm)
    e = O.generate_pairwise_interaction(O.SR_KERNELS[kind])(p, b, pairs, m, *[q[:, k] for k in range(prm.shape[1])])
    gp, gb, gq = torch.autograd.grad(e, (p, b, q))
    return {"E": float(e.detach()), "pos": gp.numpy(), "box": gb.numpy(), "prm": gq.numpy()}


def axis_type_system(seed=8, n_units=12, L=16.0):
    """Pyramidal XH3 units (C3v, X-H 1.05 A, H-X-H 106 deg) on a jittered lattice in
    an L A box, every local-frame type in use: the centres cycle through ThreeFold,
    ZBisect, Bisector, ZThenX, ZOnly and NoAxisType; hydrogens are ZThenX (z = X,
    x = the next H) or ZOnly (z = X).  Random multipoles (harmonic, e A^l),
    polarisabilities ~1.0 A^3, Thole 0.39; bonds X-H (1-2) and H-H (1-3).
    Returns a dict with pos, box, axis_types, axis_idx, bonded, Q (N,9), pol,
    thole and m/p/d scales."""
    rng = np.random.default_rng(seed)
    th = np.radians(106.0)
    # angle beta between each X-H and the C3 axis: cos(th) = cos^2 b + sin^2 b cos(120)
    b = math.acos(math.sqrt((2 * math.cos(th) + 1) / 3))
    Hdir = np.array([[math.sin(b) * math.cos(p), math.sin(b) * math.sin(p), math.cos(b)]
                     for p in (0.0, 2 * math.pi / 3, 4 * math.pi / 3)])
    n_side = int(math.ceil(n_units ** (1 / 3)))
    sites = np.array([[i, j, k] for i in range(n_side) for j in range(n_side) for k in range(n_side)])[:n_units]
    pos, types, idx, bonded = [], [], [], {}
    centre_types = [3, 2, 1, 0, 4, 5]
    for u, s in enumerate(sites):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        c = (s + 0.5) * (L / n_side) + rng.uniform(-0.5, 0.5, 3)
        base = len(pos)
        pos.append(c)
        for h in Hdir:
            pos.append(c + 1.05 * (R @ h))
        ct = centre_types[u % 6]
        h1, h2, h3 = base + 1, base + 2, base + 3
        types.append(ct)
        idx.append({3: [h1, h2, h3], 2: [h1, h2, h3], 1: [h1, h2, -1], 0: [h1, h2, -1], 4: [h1, -1, -1],
                    5: [-1, -1, -1]}[ct])
        for k, hh in enumerate((h1, h2, h3)):
            if (u + k) % 2 == 0:
                types.append(0)
                idx.append([base, (h1, h2, h3)[(k + 1) % 3], -1])
            else:
                types.append(4)
                idx.append([base, -1, -1])
        for hh in (h1, h2, h3):
            bonded[(base, hh)] = bonded[(hh, base)] = 1
        for a_, b_ in ((h1, h2), (h1, h3), (h2, h3)):
            bonded[(a_, b_)] = bonded[(b_, a_)] = 2
    pos = np.asarray(pos)
    n = len(pos)
    Q = np.zeros((n, 9))
    Q[:, 0] = rng.normal(0, 0.3, n)
    Q[:, 1:4] = rng.normal(0, 0.1, (n, 3))
    Q[:, 4:9] = rng.normal(0, 0.1, (n, 5))
    sc = np.array([0.0, 0.0, 1.0, 1.0, 1.0, 1.0])
    return {"pos": pos, "box": np.eye(3) * L, "axis_types": np.asarray(types), "axis_idx": np.asarray(idx),
            "bonded": bonded, "Q": Q, "pol": rng.uniform(0.8, 1.2, n), "thole": np.full(n, 0.39),
            "mScales": sc, "pScales": sc.copy(), "dScales": np.ones(6)}


def triclinic_water_system(n_waters=96, box=None, seed=31):
    """Polarisable waters (the benchmark recipe's multipoles, polarisabilities,
    Thole and scales) in a triclinic cell (rows a, b, c; TRICLINIC_BOX scaled to
    about liquid density by default): O sites drawn uniformly in fractional
    coordinates, redrawn while any minimum-image O-O distance is below 2.5 A,
    random orientations.  Returns (S, positions A, box A)."""
    from dmff_amd.admp.water import HOH, OH, _random_rotation, make_water_system
    B = np.asarray(box if box is not None else TRICLINIC_BOX * (n_waters / 216.0) ** (1 / 3), dtype=np.float64)
    S = make_water_system(n_waters, 1.0, seed)           # types, frames, multipoles; positions replaced below
    rng = np.random.default_rng(seed)
    Bi = np.linalg.inv(B)
    h = np.array([[OH * math.sin(HOH / 2), 0.0, OH * math.cos(HOH / 2)],
                  [-OH * math.sin(HOH / 2), 0.0, OH * math.cos(HOH / 2)]]) * 10.0
    O = []
    pos = []
    for _ in range(n_waters):
        for _t in range(5000):
            o = rng.uniform(0, 1, 3) @ B
            if not O:
                break
            d = np.asarray(O) - o
            s = d @ Bi
            s -= np.floor(s + 0.5)
            d = s @ B
            if np.min(np.einsum("ij,ij->i", d, d)) >= 2.5 ** 2:
                break
        else:
            raise RuntimeError("could not place a water")
        R = _random_rotation(rng)
        O.append(o)
        pos += [o, o + R @ h[0], o + R @ h[1]]
    S.positions_nm = np.asarray(pos) / 10.0
    S.box_nm = B / 10.0
    return S, np.asarray(pos), B


def cfg3_recipe_system(n_waters=512, seed=1003):
    """Configuration 3's recipe at a reduced size: polarisable waters at liquid
    density (L = 2.48 nm for 512 waters, so 2 rc <= L at rc 1.0 nm) with per-atom
    sqrt(C6, C8, C10) (+-10% jitter of the O/H values, A units) and per-type Slater
    A (prefactor A_i A_j in [1e4, 1e5] kJ/mol) and B (30-45 1/nm, in 1/A).
    Returns (S, positions A, box A, c_list, A, B)."""
    from dmff_amd.admp.water import make_water_system
    L = (n_waters / 33.43) ** (1 / 3)
    n_side = int(math.ceil(n_waters ** (1 / 3) - 1e-9))
    S = make_water_system(n_waters, L, seed, polarizable=True, n_side=n_side)
    rng = np.random.default_rng(seed + 1)
    C = np.array([[2.0e-3, 1.0e-4, 6.0e-6], [1.0e-4, 3.0e-6, 1.0e-7]])
    t = S.map_atomtype
    Ca = C[t] * rng.uniform(0.9, 1.1, size=(S.n_atoms, 3)) * np.array([1e6, 1e8, 1e10])
    A_t = np.sqrt(rng.uniform(1e4, 1e5, size=2))
    B_t = rng.uniform(30.0, 45.0, size=2) / 10.0
    return S, S.positions_nm * 10.0, S.box_nm * 10.0, np.sqrt(Ca), A_t[t], B_t[t]


def oracle_pme(pos, box, pairs, axis_types, axis_idx, Q, pol, thole, scales, rc, lpol, lpme=True, tol=1e-10,
               ethresh=5e-4, lmax=2):
    """Energy and gradients (pos, box, Q, pol, thole) and U by the oracle, CG to tol."""
    import torch
    from oracle import admp_oracle as O
    T = lambda x: torch.tensor(np.asarray(x, dtype=np.float64), requires_grad=True)
    f = O.ADMPPmeForce(np.asarray(box), axis_types, axis_idx, rc, ethresh, lmax, lpol, lpme, None)
    p, b, q = T(pos), T(box), T(Q)
    mS, pS, dS = scales
    out = {}
    if lpol:
        a, t = T(pol), T(thole)
        e, aux = f.get_energy(p, b, pairs, q, a, t, mS, pS, dS, solver="cg", tol=tol, aux={})
        gp, gb, gq, ga, gt = torch.autograd.grad(e, (p, b, q, a, t))
        out.update(pol=ga.numpy(), thole=gt.numpy(), U=aux["U_ind"].detach().numpy() * 10.0)
    else:
        e = f.get_energy(p, b, pairs, q, mScales=mS)
        gp, gb, gq = torch.autograd.grad(e, (p, b, q))
    out.update(E=float(e.detach()), pos=gp.numpy(), box=gb.numpy(), Q=gq.numpy())
    return out


def library_pme(pos, box, axis_types, axis_idx, bonded, Q, pol, thole, scales, rc, lpol, precision="fp64",
                device="cuda:0", pairs=None, tol=1e-10, ethresh=5e-4, lmax=2):
    """The same through the library's Engine: pairs given (reference format) or
    built by the library (pairs=None, then also returns them)."""
    from dmff_amd.admp.engine import Engine
    from dmff_amd.admp.pme import setup_ewald_parameters
    import torch
    kappa, K1, K2, K3 = setup_ewald_parameters(rc, ethresh, np.asarray(box))
    eng = Engine(len(pos), lmax, lpol, True, kappa, (K1, K2, K3), rc, precision=precision, device=device)
    eng.set_topology(axis_types, axis_idx, bonded)
    out = {}
    if pairs is None:
        eng.build_neighbors(pos, box)
        out["pairs"] = eng.get_pairs(0).cpu().numpy().astype(np.int64)
    else:
        eng.set_pairs(torch.as_tensor(pairs))
    E, g, U, info = eng.compute(pos, box, Q, pol if lpol else None, thole if lpol else None, *scales,
                                solver="pcg", tol=tol, maxiter=1000, want_grad=True)
    out.update(E=float(E[0]), pos=g["pos"][0].cpu().numpy(), box=g["box"][0].cpu().numpy(),
               Q=g["Q"][0].cpu().numpy(), n_iter=info["n_iter"][0], K=(K1, K2, K3))
    if lpol:
        out.update(pol=g["pol"][0].cpu().numpy(), thole=g["thole"][0].cpu().numpy(), U=U[0].cpu().numpy())
    return out


def rel_err(a, b, floor=1e-6):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), floor))


def qeq_system(case="qeq", water_bonds=True):
    """The charge-equilibration cases of tests/test_admp/test_qeq.py: an electrode
    of 144 atoms (qeq: one neutral group, template charges as the start; qeq2: plus
    one water, two neutral groups, zero start).  Types follow the XML residue
    particles in order, as the reference test assigns them; the water's O-H bonds
    come from OpenMM's standard template for H2O (HOH).  Returns positions and box
    in nm, per-atom chi/J/eta, constraint groups, the PME setup and the pairs."""
    import json
    import xml.etree.ElementTree as ET
    from dmff_amd.admp import frontend as fe
    from oracle.admp_oracle import setup_ewald_parameters
    c = json.load(open(os.path.join(GOLD, "energies.json")))[case]
    xml = golden_path(f"{case}.xml")
    pos_nm, top = fe.read_pdb(golden_path(f"{case}.pdb"))
    root = ET.parse(xml).getroot()
    parts = [(a.get("type"), float(a.get("charge"))) for r in root.find("Residues") for a in r if a.tag == "Atom"]
    n = top.n_atoms
    assert len(parts) == n
    node = root.find("ADMPQeqForce")
    rows = [a.attrib for a in node if a.tag == "Atom"]
    key = "type" if "type" in rows[0] else "class"
    keys = [r[key] for r in rows]
    tmap = np.array([keys.index(t) for t, _ in parts], dtype=np.int64)
    col = lambda a: np.array([float(r[a]) for r in rows], dtype=np.float64)
    charges = np.array([q for _, q in parts], dtype=np.float64)
    bonds = []
    if case == "qeq2":
        charges[:] = 0.0
        const_list, const_vals = c["const_list"], c["const_vals"]
        if water_bonds:
            bonds = [(144, 145), (144, 146)]
    else:
        groups = {}
        for i, rid in enumerate(top.res_ids):
            groups.setdefault(rid, []).append(i)
        const_list = list(groups.values())
        const_vals = [float(np.sum(charges[g])) for g in const_list]
    bonded = fe.bonded_distances(n, bonds)
    box_nm = np.asarray(top.box_nm, dtype=np.float64)
    kappa, K1, K2, K3 = setup_ewald_parameters(c["rc_nm"], c["ethresh"], box_nm)
    s14 = float(node.get("coulomb14scale"))
    return dict(pos_nm=pos_nm, box_nm=box_nm, tmap=tmap, chi_t=col("chi"), J_t=col("J"), eta_t=col("eta"),
                init_q=charges, const_list=const_list, const_vals=const_vals, kappa=kappa, K=(K1, K2, K3),
                rc_nm=c["rc_nm"], damp_mod=int(node.get("DampMod", 3)),
                mscales=np.array([0.0, 0.0, s14, 1.0, 1.0, 1.0]), bonded=bonded,
                pairs=reference_pairs(pos_nm, box_nm, c["rc_nm"], bonded), value=c["value"], decimal=c["decimal"])


def qeq_small_system(pbc=True, seed=11):
    """A small charge-equilibration case with every term live: 18 sites (six
    3-site groups, neutral) in a 1.0 nm box, eta 0.9-1.6 (so the erfc pair and
    self terms matter at Angstrom distances), J > 0, chi of both signs, 1-2 and
    1-3 exclusions inside each group; rc 0.45 nm and K 10 when periodic."""
    from oracle.admp_oracle import setup_ewald_parameters
    rng = np.random.default_rng(seed)
    n = 18
    L = 1.0
    pos = []
    while len(pos) < n:
        p = rng.uniform(0.05, 0.95, size=3) * L
        if all(np.linalg.norm(((p - q) + L / 2) % L - L / 2) > 0.15 for q in pos):
            pos.append(p)
    pos = np.array(pos)
    box = np.eye(3) * L
    groups = [list(range(3 * k, 3 * k + 3)) for k in range(6)]
    bonds = [(g[0], g[1]) for g in groups] + [(g[0], g[2]) for g in groups]
    bonded = fe.bonded_distances(n, bonds)
    chi = rng.uniform(-50.0, 150.0, size=n)
    J = rng.uniform(0.2, 0.6, size=n)
    eta = rng.uniform(0.9, 1.6, size=n)
    rc = 0.45
    if pbc:
        kappa, K1, K2, K3 = setup_ewald_parameters(rc, 5e-4, box)
        K = (max(K1, 10), max(K2, 10), max(K3, 10))
        pairs = reference_pairs(pos, box, rc, bonded)
    else:
        kappa, K = 0.0, (1, 1, 1)
        p = np.array([[i, j] for i in range(n) for j in range(i + 1, n)])
        pairs = np.concatenate([p, nbond_column(p, bonded)[:, None]], axis=1)
    return dict(pos_nm=pos, box_nm=box, chi=chi, J=J, eta=eta, const_list=groups, const_vals=[0.0] * 6,
                init_q=np.zeros(n), rc_nm=rc, kappa=kappa, K=K, pairs=pairs,
                mscales=np.array([0.0, 0.0, 0.5, 1.0, 1.0, 1.0]), bonded=bonded)
