"""Synthetic rigid-water systems for the ADMP PME benchmarks.

Recipe (nm, kJ/mol; seeds numpy default_rng(1000 + config)):
  geometry  O-H 0.09572 nm, HOH 104.52 deg; O on a jittered n^3 lattice of
            spacing L/n (+-0.02 nm uniform jitter), uniformly random orientation
            (random unit quaternion); a molecule is redrawn while any O-O
            distance is below 0.25 nm.
  frames    O bisector (kz = -H, kx = -H), H z-then-x (kz = O, kx = other H).
  multipoles (local frame, per type) q_O = -0.8, q_H = +0.4;
            mu_O = (0, 0, 0.006) e nm, mu_H ~ N(0, 0.002) per component;
            Theta traceless, components ~ N(0, 0.0005) e nm^2.
            Converted like the XML loader: dipole x10, quadrupole x300, harmonic.
  polarisation alpha_O = 0.00088 nm^3, alpha_H = 0.00050 nm^3 (x1000 -> A^3),
            Thole 0.39, m/p/d scales (0,0,0,0,1) + trailing 1.0.
Atom order is (O, H1, H2) per molecule.
"""
import math

import numpy as np

from .frontend import Bisector, ZThenX, cart2harm

OH = 0.09572
HOH = math.radians(104.52)


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_positions(n_side, L, rng, jitter=0.02, min_oo=0.25, n_waters=None):
    """Positions (nm), shape (3*nw, 3); O on a jittered n_side^3 lattice.

    An oxygen within min_oo of a new one sits at a lattice site at most
    floor((min_oo + 2 jitter) / cell) + 1 cells away along each axis, so only
    those sites are checked; the accept/redraw decisions, and with them the
    random stream, are those of a check against every placed oxygen."""
    h = np.array([[OH * math.sin(HOH / 2), 0.0, OH * math.cos(HOH / 2)],
                  [-OH * math.sin(HOH / 2), 0.0, OH * math.cos(HOH / 2)]])
    ijk = np.array([[i, j, k] for i in range(n_side) for j in range(n_side) for k in range(n_side)],
                   dtype=np.int64)
    if n_waters is not None and n_waters < len(ijk):
        ijk = ijk[np.sort(rng.choice(len(ijk), n_waters, replace=False))]
    sites = ijk.astype(np.float64) * (L / n_side)
    cell = L / n_side
    m = int(math.floor((min_oo + 2 * jitter) / cell)) + 1
    if 2 * m + 1 >= n_side:
        offs = None     # the window wraps the whole lattice: check every placed oxygen
    else:
        r = range(-m, m + 1)
        offs = [(a, b, c) for a in r for b in r for c in r]
    placed = {}
    out = []
    for s, key in zip(sites, map(tuple, ijk)):
        if offs is None:
            cand = list(placed.values())
        else:
            cand = [placed[q] for q in (((key[0] + a) % n_side, (key[1] + b) % n_side, (key[2] + c) % n_side)
                                        for a, b, c in offs) if q in placed]
        cand = np.asarray(cand) if cand else None
        for _ in range(1000):
            o = s + rng.uniform(-jitter, jitter, size=3) + cell * 0.25
            ok = True
            if cand is not None:
                d = cand - o
                d -= L * np.round(d / L)
                ok = np.min(np.einsum("ij,ij->i", d, d)) >= min_oo * min_oo
            if ok:
                break
        R = _random_rotation(rng)
        placed[key] = o
        out.append(o)
        out.append(o + R @ h[0])
        out.append(o + R @ h[1])
    return np.asarray(out)


class WaterSystem:
    pass


def make_water_system(n_waters, L_nm, seed, polarizable=True, n_side=None, Q_local_types=None):
    """Q_local_types: per-type harmonic multipoles to use instead of the seeded draw."""
    rng = np.random.default_rng(seed)
    if n_side is None:
        n_side = int(round(n_waters ** (1.0 / 3.0)))
        if n_side ** 3 < n_waters:
            n_side += 1
    pos = water_positions(n_side, L_nm, rng, n_waters=n_waters if n_side ** 3 != n_waters else None)
    nw = pos.shape[0] // 3
    # per-type local multipoles (Cartesian nm units: c, dX dY dZ, qXX qYY qZZ qXY qXZ qYZ)
    cart = np.zeros((2, 10))
    cart[0, 0], cart[1, 0] = -0.8, 0.4
    cart[0, 1:4] = [0.0, 0.0, 0.006]
    cart[1, 1:4] = rng.normal(0, 0.002, size=3)
    for t in range(2):
        xx, yy, xy, xz, yz = rng.normal(0, 0.0005, size=5)
        cart[t, 4:10] = [xx, yy, -(xx + yy), xy, xz, yz]
    cart[:, 1:4] *= 10.0
    cart[:, 4:10] *= 300.0
    S = WaterSystem()
    S.positions_nm = pos
    S.box_nm = np.eye(3) * L_nm
    S.n_atoms = 3 * nw
    S.map_atomtype = np.tile(np.array([0, 1, 1]), nw)
    S.map_poltype = S.map_atomtype.copy()
    S.Q_local_types = cart2harm(cart, 2) if Q_local_types is None else np.array(Q_local_types, dtype=np.float64)
    S.Q_local = S.Q_local_types[S.map_atomtype]
    S.pol_types = np.array([0.00088, 0.00050]) * 1000.0
    S.thole_types = np.array([0.39, 0.39])
    S.pol = S.pol_types[S.map_poltype]
    S.thole = S.thole_types[S.map_poltype]
    S.polarizable = polarizable
    sc = np.array([0.0, 0.0, 0.0, 0.0, 1.0, 1.0])
    S.mScales, S.pScales, S.dScales = sc.copy(), sc.copy(), sc.copy()
    axis_types = np.zeros(S.n_atoms, dtype=np.int64)
    axis_idx = np.full((S.n_atoms, 3), -1, dtype=np.int64)
    bonded = {}
    for w in range(nw):
        o, h1, h2 = 3 * w, 3 * w + 1, 3 * w + 2
        axis_types[o] = Bisector
        axis_idx[o] = [h1, h2, -1]
        axis_types[h1] = ZThenX
        axis_idx[h1] = [o, h2, -1]
        axis_types[h2] = ZThenX
        axis_idx[h2] = [o, h1, -1]
        for a, b, d in ((o, h1, 1), (o, h2, 1), (h1, h2, 2)):
            bonded[(a, b)] = d
            bonded[(b, a)] = d
    S.axis_types, S.axis_idx, S.bonded = axis_types, axis_idx, bonded
    return S


# the benchmark configurations (nm)
CONFIGS = {
    0: dict(n_waters=216, L=1.864, rc=0.8, polarizable=False),
    1: dict(n_waters=4096, L=4.967, rc=0.8, polarizable=True),
    3: dict(n_waters=32768, L=9.93, rc=1.0, polarizable=True),
}


def config_system(cfg):
    c = CONFIGS[cfg]
    return make_water_system(c["n_waters"], c["L"], 1000 + cfg, polarizable=c["polarizable"]), c


_CFG1_Q = []


def config1_multipoles():
    """Configuration 1's per-type local multipoles (harmonic, e*A units)."""
    if not _CFG1_Q:
        _CFG1_Q.append(config_system(1)[0].Q_local_types)
    return _CFG1_Q[0]
