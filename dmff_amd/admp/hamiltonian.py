"""Hamiltonian for the ADMP force fields: the subset of dmff/api/hamiltonian.py
(74-136) that the ADMP path needs.

    H = Hamiltonian("peg.xml")
    pots = H.createPotential("peg2.pdb", nonbondedMethod="CutoffPeriodic",
                             nonbondedCutoff=1.49, ethresh=1e-4)
    E = pots.getPotentialFunc()(positions_nm, box_nm, pairs, H.paramset)
    E_es = pots.dmff_potentials["ADMPPmeForce"](positions_nm, box_nm, pairs, H.paramset)

Topologies come from this package's PDB reader and the force field's residue
templates (atom typing by residue and atom name, bonds from the templates and
CONECT records); `potential.meta["cov_map"]` is the bonded-distance matrix of
dmff/api/topology.py:391-418 (dense for up to 20,000 atoms, otherwise the
sparse {(i, j): nbond} map under meta["bonded"] only).  Forces in the XML
that are not ADMP generators (classical terms) are ignored.
"""
import numpy as np

from . import frontend as fe
from .generators import GENERATORS, NoCutoff, ParamSet


class Potential:
    """dmff/api/hamiltonian.py:26-71: named potential functions and metadata."""

    def __init__(self):
        self.dmff_potentials = {}
        self.meta = {}

    def addDmffPotential(self, name, potential):
        self.dmff_potentials[name] = potential

    def getPotentialFunc(self, names=None):
        """The sum of the named potentials (all when names is None)."""
        if names is None:
            names = list(self.dmff_potentials)
        if isinstance(names, str):
            names = [names]
        fns = [self.dmff_potentials[n] for n in names]

        def total(positions, box, pairs, params, aux=None):
            E = 0.0
            for fn in fns:
                e = fn(positions, box, pairs, params) if aux is None else fn(positions, box, pairs, params, aux)
                if isinstance(e, tuple):
                    e, aux = e
                E = E + e
            return E if aux is None else (E, aux)

        return total


class Hamiltonian:
    """Force field from one or more XML files (the first one's residues and atom
    types type the topology); one generator per ADMP force element."""

    def __init__(self, *xmlnames):
        if not xmlnames:
            raise ValueError("at least one force-field XML is needed")
        self.ffs = [fe.ForceFieldXML(x) for x in xmlnames]
        self.ff = self.ffs[0]
        for extra in self.ffs[1:]:
            self.ff.types.update(extra.types)
            self.ff.residues.update(extra.residues)
        self.paramset = ParamSet()
        self.generators = {}
        for ff in self.ffs:
            for child in ff.root:
                cls = GENERATORS.get(child.tag)
                if cls is not None and child.tag not in self.generators:
                    self.generators[child.tag] = cls(ff, self.paramset)

    def getGenerators(self):
        return list(self.generators.values())

    def topology(self, pdb, residues=None):
        """(positions nm, Topology) of a PDB file, typed by the force field."""
        pos, top = fe.read_pdb(pdb)
        extra = fe.load_residue_bonds(residues) if residues else None
        fe.assign_types(top, self.ff, extra)
        return pos, top

    def createPotential(self, topology, nonbondedMethod=NoCutoff, nonbondedCutoff=1.0, **kwargs):
        """dmff/api/hamiltonian.py:104-128.  topology: a PDB path or a typed
        frontend.Topology.  Keywords (ethresh, step_pol, has_aux, ...) go to every
        generator."""
        top = topology
        if isinstance(topology, str):
            _, top = self.topology(topology, kwargs.pop("residues", None))
        pot = Potential()
        top._meta = pot.meta
        n = top.n_atoms
        bonded = fe.bonded_distances(n, top.bonds)
        pot.meta["bonded"] = bonded
        if n <= 20000:
            cm = np.zeros((n, n), dtype=np.int64)
            for (i, j), d in bonded.items():
                cm[i, j] = d
            pot.meta["cov_map"] = cm
        for name, gen in self.generators.items():
            pot.addDmffPotential(name, gen.createPotential(top, nonbondedMethod, nonbondedCutoff, **kwargs))
        return pot
