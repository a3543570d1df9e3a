"""Test-side helpers: load the reference's golden systems through the host
frontend and build reference-format pair lists with the oracle's brute force."""
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
