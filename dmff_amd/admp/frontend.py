"""Host-side frontend for the ADMP PME path: force-field XML, PDB and topology.

Mirrors what DMFF's generator layer does before it reaches the backend:
  * XML parsing of <ADMPPmeForce>, <ADMPDispPmeForce>, <SlaterExForce>
    (dmff/generators/admp.py:866-1067, 221-270, 672-726),
  * unit conversion: dipoles x10, Cartesian quadrupoles x300, polarisability
    x1000 (nm^3 -> A^3) (dmff/generators/admp.py:1028,1055,1063),
  * Cartesian -> spherical-harmonic multipoles (dmff/admp/multipole.py:34-74),
  * the bonded-distance matrix (dmff/api/topology.py:391-418),
  * OpenMM-style local-axis assignment (dmff/generators/admp.py:1117-1160,1209-1336).

Everything here runs once per topology on the host with numpy; no GPU work.
"""
import math
import xml.etree.ElementTree as ET
from collections import deque

import numpy as np

# dmff/admp/multipole.py:12 (10-digit literal)
RT3 = 1.73205080757
INV_RT3 = 1.0 / RT3

ZThenX, Bisector, ZBisect, ThreeFold, ZOnly, NoAxisType = 0, 1, 2, 3, 4, 5


class DMFFException(Exception):
    """Same role as dmff.utils.DMFFException."""


# ---------------------------------------------------------------------------
# Multipole representation conversions (Cartesian <-> harmonic)
# ---------------------------------------------------------------------------
def cart2harm(theta, lmax):
    """(n, 10) Cartesian (c, dX, dY, dZ, qXX, qYY, qZZ, qXY, qXZ, qYZ) ->
    (n, (lmax+1)^2) harmonic in the order 00,10,11c,11s,20,21c,21s,22c,22s
    (dmff/admp/multipole.py:15-74)."""
    if lmax > 2:
        raise ValueError("l > 2 (beyond quadrupole) not supported")
    theta = np.asarray(theta, dtype=np.float64)
    n = theta.shape[0]
    out = np.zeros((n, (lmax + 1) ** 2))
    out[:, 0] = theta[:, 0]
    if lmax >= 1:
        out[:, 1] = theta[:, 3]
        out[:, 2] = theta[:, 1]
        out[:, 3] = theta[:, 2]
    if lmax >= 2:
        xx, yy, zz, xy, xz, yz = (theta[:, 4 + k] for k in range(6))
        out[:, 4] = zz
        out[:, 5] = 2 * INV_RT3 * xz
        out[:, 6] = 2 * INV_RT3 * yz
        out[:, 7] = INV_RT3 * (xx - yy)
        out[:, 8] = 2 * INV_RT3 * xy
    return out


def harm2cart(Q, lmax):
    """Inverse of cart2harm (dmff/admp/multipole.py:77-114)."""
    Q = np.asarray(Q, dtype=np.float64)
    n = Q.shape[0]
    nc = {0: 1, 1: 4, 2: 10}[lmax]
    out = np.zeros((n, nc))
    out[:, 0] = Q[:, 0]
    if lmax >= 1:
        out[:, 1] = Q[:, 2]
        out[:, 2] = Q[:, 3]
        out[:, 3] = Q[:, 1]
    if lmax >= 2:
        q20, q21c, q21s, q22c, q22s = (Q[:, 4 + k] for k in range(5))
        out[:, 4] = -0.5 * q20 + RT3 / 2 * q22c
        out[:, 5] = -0.5 * q20 - RT3 / 2 * q22c
        out[:, 6] = q20
        out[:, 7] = RT3 / 2 * q22s
        out[:, 8] = RT3 / 2 * q21c
        out[:, 9] = RT3 / 2 * q21s
    return out


# ---------------------------------------------------------------------------
# PDB + topology
# ---------------------------------------------------------------------------
class Topology:
    """Atoms (name, residue name, residue index, element, type) and bonds."""

    def __init__(self):
        self.atom_names = []
        self.res_names = []
        self.res_ids = []
        self.elements = []
        self.types = []
        self.bonds = []
        self.conect = []    # (i, j) bonds from CONECT records, 0-based atom indices
        self.box_nm = None

    @property
    def n_atoms(self):
        return len(self.atom_names)


def read_pdb(path):
    """Minimal PDB reader: ATOM/HETATM coordinates (returned in nm) and CRYST1 box (nm)."""
    top = Topology()
    pos = []
    serial = {}
    model_done = False
    res_key_prev = None
    res_counter = -1
    with open(path) as f:
        for line in f:
            rec = line[:6].strip()
            if rec == "CONECT":
                # serial numbers in columns 7-11, 12-16, ... (bonds are added once per pair)
                fields = [line[k:k + 5].strip() for k in range(6, min(len(line.rstrip("\n")), 31), 5)]
                ids = [int(x) for x in fields if x]
                if ids:
                    a = ids[0]
                    for b in ids[1:]:
                        if a in serial and b in serial:
                            i, j = serial[a], serial[b]
                            if i != j:
                                top.conect.append((min(i, j), max(i, j)))
                continue
            if rec == "CRYST1":
                a, b, c = float(line[6:15]), float(line[15:24]), float(line[24:33])
                al, be, ga = float(line[33:40]), float(line[40:47]), float(line[47:54])
                top.box_nm = _box_from_lengths(a, b, c, al, be, ga) / 10.0
            elif rec in ("ATOM", "HETATM") and not model_done:
                name = line[12:16].strip()
                resn = line[17:20].strip()
                key = (line[21], line[22:27])
                if key != res_key_prev:
                    res_counter += 1
                    res_key_prev = key
                x, y, z = float(line[30:38]), float(line[38:46]), float(line[46:54])
                el = line[76:78].strip() if len(line) >= 78 else name[0]
                top.atom_names.append(name)
                top.res_names.append(resn)
                top.res_ids.append(res_counter)
                top.elements.append(el)
                try:
                    serial[int(line[6:11])] = len(pos)
                except ValueError:
                    pass
                pos.append((x / 10.0, y / 10.0, z / 10.0))
            elif rec == "ENDMDL":
                model_done = model_done or bool(pos)   # atoms of later models are skipped
            elif rec == "END":
                break
    return np.asarray(pos, dtype=np.float64), top


def _box_from_lengths(a, b, c, al, be, ga):
    if abs(al - 90) < 1e-6 and abs(be - 90) < 1e-6 and abs(ga - 90) < 1e-6:
        return np.diag([a, b, c]).astype(np.float64)
    al, be, ga = (math.radians(x) for x in (al, be, ga))
    ax = np.array([a, 0, 0])
    bx = np.array([b * math.cos(ga), b * math.sin(ga), 0])
    cx = c * math.cos(be)
    cy = c * (math.cos(al) - math.cos(be) * math.cos(ga)) / math.sin(ga)
    cz = math.sqrt(c * c - cx * cx - cy * cy)
    return np.array([ax, bx, [cx, cy, cz]])


def load_residue_bonds(path):
    """Residue bond definitions in OpenMM's residues.xml format (Bond from/to atom names)."""
    out = {}
    root = ET.parse(path).getroot()
    for res in root.iter("Residue"):
        out.setdefault(res.get("name"), []).extend(
            (b.get("from"), b.get("to")) for b in res.iter("Bond"))
    return out


# Water residue and atom names the PDB reader of OpenMM maps to HOH / O, H1, H2
# (its residue and atom name replacement tables); applied when the force field
# has no template under the PDB's own name.
WATER_RESIDUES = ("SOL", "WAT", "TIP3", "T3P", "TIP4", "T4P", "SPC", "H2O")
WATER_ATOMS = {"OW": "O", "OH2": "O", "HW1": "H1", "HW2": "H2", "H1W": "H1", "H2W": "H2"}


def _alias_names(top, tmpl):
    for i in range(top.n_atoms):
        if top.res_names[i] not in tmpl and top.res_names[i] in WATER_RESIDUES and "HOH" in tmpl:
            top.res_names[i] = "HOH"
            top.atom_names[i] = WATER_ATOMS.get(top.atom_names[i], top.atom_names[i])


def assign_types(top, ff, extra_bonds=None):
    """Atom typing from the force field's <Residues> templates (residue and atom
    names), bonds from the template or from extra residue bond definitions, and
    per-atom charges where the template carries them (top.charges, else None)."""
    tmpl = ff.residues
    _alias_names(top, tmpl)
    charges = []
    for i in range(top.n_atoms):
        r = tmpl.get(top.res_names[i])
        if r is None or top.atom_names[i] not in r["types"]:
            raise DMFFException("Atom %d not matched in forcefield!" % i)
        top.types.append(r["types"][top.atom_names[i]])
        charges.append(r.get("charges", {}).get(top.atom_names[i]))
    top.charges = None if any(c is None for c in charges) else np.asarray(charges, dtype=np.float64)
    by_res = {}
    for i, rid in enumerate(top.res_ids):
        by_res.setdefault(rid, {})[top.atom_names[i]] = i
    bonds = set(getattr(top, "conect", []))
    for rid, names in by_res.items():
        rname = top.res_names[next(iter(names.values()))]
        pairs = list(tmpl.get(rname, {}).get("bonds", []))
        if extra_bonds and rname in extra_bonds:
            pairs += extra_bonds[rname]
        for a, b in pairs:
            if a in names and b in names:
                i, j = names[a], names[b]
                bonds.add((min(i, j), max(i, j)))
    top.bonds = sorted(bonds)
    return top


def bonded_distances(n_atoms, bonds, nmax=6):
    """Sparse form of DMFFTopology.buildCovMat (dmff/api/topology.py:391-418):
    returns {(i, j): nbond} for i != j within nmax bonds (BFS graph distance)."""
    adj = [[] for _ in range(n_atoms)]
    for i, j in bonds:
        adj[i].append(j)
        adj[j].append(i)
    out = {}
    for s in range(n_atoms):
        if not adj[s]:
            continue
        dist = {s: 0}
        dq = deque([s])
        while dq:
            u = dq.popleft()
            if dist[u] >= nmax:
                continue
            for v in adj[u]:
                if v not in dist:
                    dist[v] = dist[u] + 1
                    dq.append(v)
        for v, d in dist.items():
            if v != s:
                out[(s, v)] = d
    return out


def cov_map_dense(n_atoms, bonds, nmax=6):
    """Dense (N,N) bonded-distance matrix, as buildCovMat returns it."""
    cm = np.zeros((n_atoms, n_atoms), dtype=np.int64)
    for (i, j), d in bonded_distances(n_atoms, bonds, nmax).items():
        cm[i, j] = d
    return cm


# ---------------------------------------------------------------------------
# Force-field XML
# ---------------------------------------------------------------------------
class ForceFieldXML:
    def __init__(self, path):
        self.root = ET.parse(path).getroot()
        self.types = {}
        for t in self.root.iter("Type"):
            self.types[t.get("name")] = t.attrib
        self.residues = {}
        for r in self.root.iter("Residue"):
            names = [a.get("name") for a in r.iter("Atom")]
            types = {a.get("name"): a.get("type") for a in r.iter("Atom")}
            bonds = [(names[int(b.get("from"))], names[int(b.get("to"))]) for b in r.iter("Bond")
                     if b.get("from") is not None and b.get("from").lstrip("-").isdigit()]
            bonds += [(b.get("atomName1"), b.get("atomName2")) for b in r.iter("Bond")
                      if b.get("atomName1") is not None and b.get("atomName2") is not None]
            charges = {a.get("name"): float(a.get("charge")) for a in r.iter("Atom") if a.get("charge") is not None}
            self.residues[r.get("name")] = {"types": types, "bonds": bonds, "charges": charges}

    def force(self, name):
        return self.root.find(name)


def _scales(node, prefix, defaults):
    """mScale12..16 etc. + trailing forced 1.0 (dmff/generators/admp.py:883-918)."""
    s = [float(node.get(f"{prefix}1{i}", defaults[i - 2])) for i in range(2, 7)]
    s.append(1.0)
    return np.asarray(s, dtype=np.float64)


def set_axis_type(kz, kx, ky):
    """dmff/generators/admp.py:1117-1160."""
    kyN = ky.startswith("-")
    ky = ky[1:] if kyN else ky
    kxN = kx.startswith("-")
    kx = kx[1:] if kxN else kx
    kzN = kz.startswith("-")
    kz = kz[1:] if kzN else kz
    t = ZThenX
    if not kz:
        t = NoAxisType
    if kz and not kx:
        t = ZOnly
    if (kz and kzN) or (kx and kxN):
        t = Bisector
    if kx and kxN and ky and kyN:
        t = ZBisect
    if kz and kzN and kx and kxN and ky and kyN:
        t = ThreeFold
    return t, (kz, kx, ky)


class ADMPPmeParams:
    """Parsed <ADMPPmeForce>: per-type Q_local (harmonic), pol (A^3), thole, scales."""

    def __init__(self, ff):
        node = ff.force("ADMPPmeForce")
        if node is None:
            raise DMFFException("no ADMPPmeForce in force field")
        self.lmax = int(node.get("lmax"))
        self.mScales = _scales(node, "mScale", [0.0, 0.0, 1.0, 1.0, 1.0])
        self.pScales = _scales(node, "pScale", [0.0, 0.0, 1.0, 1.0, 1.0])
        self.dScales = _scales(node, "dScale", [1.0, 1.0, 1.0, 1.0, 1.0])
        self.key_type = "type"
        self.multipole_types, self.kz, self.kx, self.ky = [], [], [], []
        cart = []
        for a in node:
            if a.tag not in ("Atom", "Multipole"):
                continue
            self.key_type = "type" if "type" in a.attrib else "class"
            self.multipole_types.append(a.get(self.key_type))
            self.kz.append(a.get("kz", ""))
            self.kx.append(a.get("kx", ""))
            self.ky.append(a.get("ky", ""))
            row = [float(a.get("c0"))]
            if self.lmax >= 1:
                row += [float(a.get("dX")) * 10, float(a.get("dY")) * 10, float(a.get("dZ")) * 10]
            else:
                row += [0.0, 0.0, 0.0]
            if self.lmax >= 2:
                row += [float(a.get(k)) * 300 for k in ("qXX", "qYY", "qZZ", "qXY", "qXZ", "qYZ")]
            else:
                row += [0.0] * 6
            cart.append(row)
        cart = np.asarray(cart, dtype=np.float64)
        self.Q_local = cart2harm(cart, self.lmax)
        pol_nodes = [a for a in node if a.tag == "Polarize"]
        self.lpol = len(pol_nodes) > 0
        self.polarize_types, pol, thole = [], [], []
        for a in pol_nodes:
            self.polarize_types.append(a.get(self.key_type))
            pxx = float(a.get("polarizabilityXX"))
            pyy = float(a.get("polarizabilityYY"))
            pzz = float(a.get("polarizabilityZZ"))
            pol.append((pxx + pyy + pzz) / 3.0 * 1000.0)
            thole.append(float(a.get("thole")))
        self.pol = np.asarray(pol, dtype=np.float64)
        self.thole = np.asarray(thole, dtype=np.float64)

    def params(self):
        d = {"Q_local": self.Q_local.copy()}
        if self.lpol:
            d["pol"] = self.pol.copy()
            d["thole"] = self.thole.copy()
        return d


def atom_type_keys(top, ff, key_type):
    if key_type == "type":
        return list(top.types)
    return [ff.types[t]["class"] for t in top.types]


def assign_axes(pme, top, ff, bonded):
    """Local-axis assignment, dmff/generators/admp.py:1209-1336.
    bonded: {(i,j): nbond} (1 = directly bonded)."""
    n = top.n_atoms
    keys = atom_type_keys(top, ff, pme.key_type)
    nbr = [[] for _ in range(n)]
    for (i, j), d in bonded.items():
        if d == 1:
            nbr[i].append(j)
    for lst in nbr:
        lst.sort()
    if pme.lmax == 0:
        return np.zeros(n, dtype=np.int64), np.full((n, 3), -1, dtype=np.int64)
    axis_types = {}
    axis_idx = {}
    patched = np.zeros(n, dtype=bool)
    ntyp = len(pme.multipole_types)
    for rev in range(ntyp):
        it = ntyp - rev - 1
        t = pme.multipole_types[it]
        for ia in [i for i in range(n) if keys[i] == t]:
            if patched[ia]:
                continue
            atype, (kz, kx, ky) = set_axis_type(pme.kz[it], pme.kx[it], pme.ky[it])
            z = x = y = -1
            neighbors = nbr[ia]
            if len(neighbors) < 1 and atype != NoAxisType:
                continue
            if atype == ZThenX:
                z = next((j for j in neighbors if keys[j] == kz), -1)
                if z < 0:
                    continue
                x = next((j for j in neighbors if j != z and keys[j] == kx), -1)
                if x < 0:
                    x = next((j for j in nbr[z] if j != ia and keys[j] == kx), -1)
                if x < 0:
                    continue
            elif atype == ZOnly:
                z = next((j for j in neighbors if keys[j] == kz), -1)
                if z < 0:
                    continue
            elif atype == Bisector:
                z = next((j for j in neighbors if keys[j] == kz), -1)
                if z < 0:
                    continue
                x = next((j for j in neighbors if j != z and keys[j] == kx), -1)
                if x < 0:
                    continue
            elif atype in (ZBisect, ThreeFold):
                z = next((j for j in neighbors if keys[j] == kz), -1)
                if z < 0:
                    continue
                x = next((j for j in neighbors if j != z and keys[j] == kx), -1)
                if x < 0:
                    continue
                y = next((j for j in neighbors if j not in (z, x) and keys[j] == ky), -1)
                if y < 0:
                    continue
            elif atype != NoAxisType:
                continue
            axis_types[ia] = atype
            axis_idx[ia] = [z, x, y]
            patched[ia] = True
    for ia in range(n):
        if not patched[ia]:
            raise DMFFException("Atom %d not matched in forcefield!" % ia)
    return (np.array([axis_types[i] for i in range(n)], dtype=np.int64),
            np.array([axis_idx[i] for i in range(n)], dtype=np.int64))


def type_maps(pme, top, ff):
    keys = atom_type_keys(top, ff, pme.key_type)
    map_atomtype = np.array([pme.multipole_types.index(k) for k in keys], dtype=np.int64)
    if pme.lpol:
        map_poltype = np.array([pme.polarize_types.index(k) for k in keys], dtype=np.int64)
    else:
        map_poltype = np.zeros(len(keys), dtype=np.int64)
    return map_atomtype, map_poltype


class ADMPDispPmeParams:
    """Parsed <ADMPDispPmeForce> (dmff/generators/admp.py:221-270): per-type C6/C8/C10
    in kJ/mol*nm^p and mScales."""

    def __init__(self, ff, name="ADMPDispPmeForce"):
        node = ff.force(name)
        if node is None:
            raise DMFFException(f"no {name} in force field")
        self.mScales = _scales(node, "mScale", [0.0, 0.0, 0.0, 1.0, 1.0])
        self.keys, c6, c8, c10 = [], [], [], []
        self.key_type = "type"
        for a in node.iter("Atom"):
            self.key_type = "type" if "type" in a.attrib else "class"
            self.keys.append(a.get(self.key_type))
            c6.append(float(a.get("C6")))
            c8.append(float(a.get("C8")))
            c10.append(float(a.get("C10")))
        self.C6, self.C8, self.C10 = (np.asarray(v, dtype=np.float64) for v in (c6, c8, c10))


class SlaterExParams:
    """Parsed <SlaterExForce> (dmff/generators/admp.py:672-726): per-type A (kJ/mol), B (1/nm)."""

    def __init__(self, ff, name="SlaterExForce"):
        node = ff.force(name)
        if node is None:
            raise DMFFException(f"no {name} in force field")
        self.mScales = _scales(node, "mScale", [0.0, 0.0, 0.0, 1.0, 1.0])
        self.keys, A, B = [], [], []
        self.key_type = "type"
        for a in node.iter("Atom"):
            self.key_type = "type" if "type" in a.attrib else "class"
            self.keys.append(a.get(self.key_type))
            A.append(float(a.get("A")))
            B.append(float(a.get("B")))
        self.A, self.B = np.asarray(A), np.asarray(B)
