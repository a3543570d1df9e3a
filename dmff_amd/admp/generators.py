"""Generator layer of the ADMP path: force-field XML -> createPotential ->
potential_fn(positions, box, pairs, params, aux=None), as dmff/generators/admp.py
builds it, on the MI355X library.

Frontend units are the reference's: positions and box in nm (converted to
Angstrom inside potential_fn, admp.py:1370-1371), parameters per atom type in
the XML's units.  `params` is a ParamSet or a dict {force name: {parameter:
tensor}}; the per-type tensors are expanded per atom by indexing with the type
maps, so torch autograd returns per-type gradients (the role jax.grad plays in
the reference).

Generators (reference file:line):
  ADMPPmeForce       ADMPPmeGenerator        dmff/generators/admp.py:866-1414
  ADMPDispPmeForce   ADMPDispPmeGenerator    :221-397   (returns -E)
  ADMPDispForce      ADMPDispGenerator       :22-218    (TT damping - dispersion PME)
  QqTtDampingForce   QqTtDampingGenerator    :400-526
  SlaterDampingForce SlaterDampingGenerator  :529-669
  SlaterExForce      SlaterExGenerator       :672-792
  SlaterSrEsForce, SlaterSrPolForce, SlaterSrDispForce, SlaterDhfForce
                     SlaterSrEsGenerator and subclasses :795-863 (returns -E)
  ADMPQeqForce       ADMPQeqGenerator        dmff/generators/qeq.py:17-223
"""
import numpy as np
import torch

from . import frontend as fe
from ._autograd import as_t
from .disp_pme import ADMPDispPmeForce
from .pairwise import (TT_damping_qq_c6_kernel, TT_damping_qq_kernel, generate_pairwise_interaction,
                       slater_disp_damping_kernel, slater_sr_kernel)
from .pme import ADMPPmeForce, setup_ewald_parameters
from .qeq import ADMPQeqForce

F64 = torch.float64

# nonbonded methods (openmm.app names); strings or objects whose str() is the name
PME, CutoffPeriodic, NoCutoff, CutoffNonPeriodic = "PME", "CutoffPeriodic", "NoCutoff", "CutoffNonPeriodic"


def _method_name(m):
    name = getattr(m, "name", None) or getattr(m, "__name__", None) or str(m)
    return name.split(".")[-1]


def _cutoff_angstrom(c):
    """nonbondedCutoff in nm (a number) or a quantity with value_in_unit."""
    if hasattr(c, "value_in_unit"):
        import importlib
        unit = importlib.import_module("openmm.unit")
        return float(c.value_in_unit(unit.angstrom))
    return float(c) * 10.0


class ParamSet:
    """Per-force parameter tensors: paramset[force][name] (dmff/api/paramset.py:6-163,
    without masks).  `parameters` is the plain dict view the reference exposes."""

    def __init__(self):
        self.parameters = {}

    def addField(self, field):
        self.parameters.setdefault(field, {})

    def addParameter(self, values, name, field):
        self.parameters.setdefault(field, {})[name] = torch.as_tensor(np.asarray(values), dtype=F64).clone()

    def __getitem__(self, k):
        return self.parameters[k]

    def __contains__(self, k):
        return k in self.parameters

    def to(self, device):
        for f in self.parameters.values():
            for k in list(f):
                f[k] = f[k].to(device)
        return self

    def requires_grad_(self, flag=True):
        for f in self.parameters.values():
            for v in f.values():
                v.requires_grad_(flag)
        return self


def _params(params, name):
    p = params.parameters if hasattr(params, "parameters") else params
    return p[name]


class _AtomTable:
    """<Atom .../> rows of one force: key attribute (type or class), per-row
    attributes, and the force node's mScale12..16 with the forced trailing 1.0."""

    def __init__(self, ff, name, defaults=(0.0, 0.0, 0.0, 1.0, 1.0)):
        node = ff.force(name)
        if node is None:
            raise fe.DMFFException(f"no {name} in force field")
        self.node = node
        self.rows = [a.attrib for a in node if a.tag == "Atom"]
        self.key_type = "type"
        for r in self.rows:
            self.key_type = "type" if "type" in r else "class"
        self.keys = [r[self.key_type] for r in self.rows]
        self.mScales = fe._scales(node, "mScale", list(defaults))

    def column(self, attr):
        return np.asarray([float(r[attr]) for r in self.rows], dtype=np.float64)

    def type_map(self, top, ff):
        keys = fe.atom_type_keys(top, ff, self.key_type)
        try:
            return np.array([self.keys.index(k) for k in keys], dtype=np.int64)
        except ValueError as e:
            raise fe.DMFFException(f"atom type not found in {self.node.tag}: {e}")


class _Base:
    name = None

    def getName(self):
        return self.name

    def getJaxPotential(self):
        return self._potential

    getPotential = getJaxPotential


# ---------------------------------------------------------------------------
class ADMPPmeGenerator(_Base):
    """dmff/generators/admp.py:866-1414."""
    name = "ADMPPmeForce"

    def __init__(self, ff, paramset):
        self.ff = ff
        self.pme = fe.ADMPPmeParams(ff)
        self.lmax, self.lpol = self.pme.lmax, self.pme.lpol
        self.mScales, self.pScales, self.dScales = self.pme.mScales, self.pme.pScales, self.pme.dScales
        self.ethresh = 5e-4
        self.step_pol = None
        paramset.addField(self.name)
        paramset.addParameter(self.pme.Q_local, "Q_local", field=self.name)
        if self.lpol:
            paramset.addParameter(self.pme.pol, "pol", field=self.name)
            paramset.addParameter(self.pme.thole, "thole", field=self.name)
        self._potential = None

    def createPotential(self, top, nonbondedMethod, nonbondedCutoff, **kwargs):
        """admp.py:1162-1408.  Extra keywords (not in the reference) go to the
        ADMPPmeForce drop-in: precision, solver, tol, maxiter, device."""
        m = _method_name(nonbondedMethod)
        if m not in (PME, CutoffPeriodic, NoCutoff, CutoffNonPeriodic):
            raise ValueError("Illegal nonbonded method for ADMPPmeForce")
        lpme = m == PME
        n = top.n_atoms
        map_atomtype, map_poltype = fe.type_maps(self.pme, top, self.ff)
        box = np.asarray(top.box_nm, dtype=np.float64) * 10.0 if top.box_nm is not None else np.eye(3)
        rc = _cutoff_angstrom(nonbondedCutoff)
        bonded = fe.bonded_distances(n, top.bonds)
        axis_types, axis_idx = fe.assign_axes(self.pme, top, self.ff, bonded)
        has_aux = bool(kwargs.get("has_aux", False))
        self.ethresh = kwargs.get("ethresh", self.ethresh)
        self.step_pol = kwargs.get("step_pol", self.step_pol)
        extra = {k: kwargs[k] for k in ("precision", "solver", "tol", "maxiter", "device") if k in kwargs}
        pme_force = ADMPPmeForce(box, map_atomtype, axis_types, axis_idx, rc, self.ethresh, self.lmax, self.lpol,
                                 lpme, self.step_pol, has_aux, **extra)
        self.pme_force = pme_force
        meta = getattr(top, "_meta", None)
        if meta is not None:
            meta[self.name + "_map_atomtype"] = map_atomtype
            meta[self.name + "_map_poltype"] = map_poltype
        ma = torch.as_tensor(map_atomtype)
        mp = torch.as_tensor(map_poltype)
        lpol = self.lpol
        mS, pS, dS = self.mScales, self.pScales, self.dScales

        def potential_fn(positions, box, pairs, params, aux=None):
            positions = as_t(positions) * 10.0
            box = as_t(box) * 10.0
            p = _params(params, "ADMPPmeForce")
            Q_local = p["Q_local"][ma.to(p["Q_local"].device)]
            if lpol:
                pol = p["pol"][mp.to(p["pol"].device)]
                tholes = p["thole"][mp.to(p["thole"].device)]
                if has_aux:
                    U0 = aux["U_ind"] if aux is not None else torch.zeros((n, 3), dtype=F64)
                    return pme_force.get_energy(positions, box, pairs, Q_local, pol, tholes, mS, pS, dS,
                                                U_init=U0, aux=aux if aux is not None else {})
                return pme_force.get_energy(positions, box, pairs, Q_local, pol, tholes, mS, pS, dS)
            energy = pme_force.get_energy(positions, box, pairs, Q_local, mS)
            return (energy, aux) if has_aux else energy

        self._potential = potential_fn
        return potential_fn


# ---------------------------------------------------------------------------
class ADMPDispPmeGenerator(_Base):
    """dmff/generators/admp.py:221-397: undamped C6/C8/C10, returns -E."""
    name = "ADMPDispPmeForce"

    def __init__(self, ff, paramset):
        self.ff = ff
        self.tab = _AtomTable(ff, self.name)
        self.mScales = self.tab.mScales
        self.ethresh, self.pmax = 5e-4, 10
        paramset.addField(self.name)
        for k in ("C6", "C8", "C10"):
            paramset.addParameter(self.tab.column(k), k, field=self.name)
        self._potential = None

    def createPotential(self, top, nonbondedMethod, nonbondedCutoff, **kwargs):
        m = _method_name(nonbondedMethod)
        if m not in (CutoffPeriodic, NoCutoff, PME):
            raise ValueError(f"Illegal nonbonded method for {self.name}")
        lpme = m != CutoffPeriodic   # admp.py:344-347: only CutoffPeriodic switches PME off
        map_atomtype = self.tab.type_map(top, self.ff)
        box = np.asarray(top.box_nm, dtype=np.float64) * 10.0 if lpme else np.ones((3, 3))
        rc = _cutoff_angstrom(nonbondedCutoff)
        self.ethresh = kwargs.get("ethresh", self.ethresh)
        extra = {k: kwargs[k] for k in ("precision", "device") if k in kwargs}
        disp = ADMPDispPmeForce(box, rc, self.ethresh, self.pmax, lpme, **extra)
        self.disp_force = disp
        ma = torch.as_tensor(map_atomtype)
        has_aux = bool(kwargs.get("has_aux", False))
        mS = self.mScales

        def potential_fn(positions, box, pairs, params, aux=None):
            positions = as_t(positions) * 10.0
            box = as_t(box) * 10.0
            p = _params(params, self.name)
            idx = ma.to(p["C6"].device)
            c_list = torch.stack([torch.sqrt(p["C6"][idx] * 1e6), torch.sqrt(p["C8"][idx] * 1e8),
                                  torch.sqrt(p["C10"][idx] * 1e10)], dim=1)
            E = -disp.get_energy(positions, box, pairs, c_list, mS)
            return (E, aux) if has_aux else E

        self._potential = potential_fn
        return potential_fn


# ---------------------------------------------------------------------------
class _PairGenerator(_Base):
    """Common shape of the short-range generators: per-type XML columns, a pair
    kernel, a conversion of the per-atom columns to the kernel's arguments, and
    the sign of the result."""
    kernel = None
    columns = ()
    sign = 1.0

    def __init__(self, ff, paramset, name=None):
        if name is not None:
            self.name = name
        self.ff = ff
        self.tab = _AtomTable(ff, self.name)
        self.mScales = self.tab.mScales
        paramset.addField(self.name)
        for k in self.columns:
            paramset.addParameter(self.tab.column(k), k, field=self.name)
        self._potential = None

    def convert(self, cols):
        raise NotImplementedError

    def createPotential(self, top, nonbondedMethod, nonbondedCutoff, **kwargs):
        map_atomtype = self.tab.type_map(top, self.ff)
        extra = {k: kwargs[k] for k in ("precision", "device") if k in kwargs}
        pot = generate_pairwise_interaction(self.kernel, static_args={}, **extra)
        ma = torch.as_tensor(map_atomtype)
        has_aux = bool(kwargs.get("has_aux", False))
        mS, sign = self.mScales, self.sign

        def potential_fn(positions, box, pairs, params, aux=None):
            positions = as_t(positions) * 10.0
            box = as_t(box) * 10.0
            p = _params(params, self.name)
            idx = ma.to(p[self.columns[0]].device)
            args = self.convert({k: p[k][idx] for k in self.columns})
            E = pot(positions, box, pairs, mS, *args)
            if sign != 1.0:
                E = sign * E
            return (E, aux) if has_aux else E

        self._potential = potential_fn
        return potential_fn


class QqTtDampingGenerator(_PairGenerator):
    """admp.py:400-526: Tang-Toennies damped charge-charge, B nm^-1 -> A^-1."""
    name, kernel, columns = "QqTtDampingForce", TT_damping_qq_kernel, ("B", "Q")

    def convert(self, c):
        return c["B"] / 10.0, c["Q"]


class SlaterDampingGenerator(_PairGenerator):
    """admp.py:529-669: Slater-ISA damped C6/C8/C10 (nm -> A)."""
    name, kernel, columns = "SlaterDampingForce", slater_disp_damping_kernel, ("B", "C6", "C8", "C10")

    def convert(self, c):
        return (c["B"] / 10.0, torch.sqrt(c["C6"] * 1e6), torch.sqrt(c["C8"] * 1e8), torch.sqrt(c["C10"] * 1e10))


class SlaterExGenerator(_PairGenerator):
    """admp.py:672-792: Slater-ISA exchange A (kJ/mol), B nm^-1 -> A^-1."""
    name, kernel, columns = "SlaterExForce", slater_sr_kernel, ("A", "B")

    def convert(self, c):
        return c["A"], c["B"] / 10.0


class SlaterSrEsGenerator(SlaterExGenerator):
    """admp.py:795-842: the charge-penetration terms, exchange form with a minus sign."""
    name, sign = "SlaterSrEsForce", -1.0


class SlaterSrPolGenerator(SlaterSrEsGenerator):
    name = "SlaterSrPolForce"


class SlaterSrDispGenerator(SlaterSrEsGenerator):
    name = "SlaterSrDispForce"


class SlaterDhfGenerator(SlaterSrEsGenerator):
    name = "SlaterDhfForce"


class ADMPDispGenerator(_Base):
    """admp.py:22-218: TT-damped short range (A, B, Q, C6) minus dispersion PME."""
    name = "ADMPDispForce"

    def __init__(self, ff, paramset):
        self.ff = ff
        self.tab = _AtomTable(ff, self.name)
        self.mScales = self.tab.mScales
        self.ethresh, self.pmax = 5e-4, 10
        paramset.addField(self.name)
        for k in ("A", "B", "Q", "C6", "C8", "C10"):
            paramset.addParameter(self.tab.column(k), k, field=self.name)
        self._potential = None

    def createPotential(self, top, nonbondedMethod, nonbondedCutoff, **kwargs):
        m = _method_name(nonbondedMethod)
        if m not in (CutoffPeriodic, NoCutoff, PME):
            raise ValueError("Illegal nonbonded method for ADMPDispForce")
        lpme = m != CutoffPeriodic
        map_atomtype = self.tab.type_map(top, self.ff)
        box = np.asarray(top.box_nm, dtype=np.float64) * 10.0 if lpme else np.ones((3, 3))
        rc = _cutoff_angstrom(nonbondedCutoff)
        self.ethresh = kwargs.get("ethresh", self.ethresh)
        extra = {k: kwargs[k] for k in ("precision", "device") if k in kwargs}
        disp = ADMPDispPmeForce(box, rc, self.ethresh, self.pmax, lpme, **extra)
        self.disp_pme_force = disp
        sr = generate_pairwise_interaction(TT_damping_qq_c6_kernel, static_args={}, **extra)
        ma = torch.as_tensor(map_atomtype)
        has_aux = bool(kwargs.get("has_aux", False))
        mS = self.mScales

        def potential_fn(positions, box, pairs, params, aux=None):
            positions = as_t(positions) * 10.0
            box = as_t(box) * 10.0
            p = _params(params, self.name)
            idx = ma.to(p["A"].device)
            a = p["A"][idx] / 2625.5              # kJ/mol -> hartree
            b = p["B"][idx] * 0.0529177249        # nm^-1 -> bohr^-1
            q = p["Q"][idx]
            c6 = torch.sqrt(p["C6"][idx] * 1e6)
            c8 = torch.sqrt(p["C8"][idx] * 1e8)
            c10 = torch.sqrt(p["C10"][idx] * 1e10)
            E_sr = sr(positions, box, pairs, mS, a, b, q, c6)
            E_lr = disp.get_energy(positions, box, pairs, torch.stack([c6, c8, c10], dim=1), mS)
            E = E_sr - E_lr
            return (E, aux) if has_aux else E

        self._potential = potential_fn
        return potential_fn


# ---------------------------------------------------------------------------
class ADMPQeqGenerator(_Base):
    """dmff/generators/qeq.py:17-223: per-type chi, J, eta; charge groups per
    residue unless const_list/const_vals are given."""
    name = "ADMPQeqForce"

    def __init__(self, ff, paramset):
        self.ff = ff
        self.table = _AtomTable(ff, self.name)
        node = self.table.node
        self.coulomb14scale = float(node.get("coulomb14scale"))
        self.damp_mod = int(node.get("DampMod", 3))
        paramset.addField(self.name)
        for attr in ("chi", "J", "eta"):
            paramset.addParameter(self.table.column(attr), attr, field=self.name)
        self._potential = None

    def createPotential(self, top, nonbondedMethod, nonbondedCutoff, **kwargs):
        """qeq.py:119-222.  Extra keywords (not in the reference) go to the
        ADMPQeqForce drop-in: precision, device, tol, maxiter."""
        m = _method_name(nonbondedMethod)
        if m not in (PME, NoCutoff):
            raise fe.DMFFException("Illegal nonbonded method for NonbondedForce")
        no_cut = m == NoCutoff
        r_cut = _cutoff_angstrom(nonbondedCutoff) / 10.0
        if not no_cut:
            box = np.asarray(top.box_nm, dtype=np.float64)
            kappa, K1, K2, K3 = setup_ewald_parameters(r_cut, kwargs.get("ethresh", 1e-5), box,
                                                       kwargs.get("PmeSpacing", 0.1),
                                                       kwargs.get("PmeCoeffMethod", "openmm"))
        else:
            kappa, K1, K2, K3 = 1.0, 1, 1, 1
        n = top.n_atoms
        init_q = np.asarray(getattr(top, "charges", None) if getattr(top, "charges", None) is not None
                            else np.zeros(n), dtype=np.float64)
        map_idx = torch.as_tensor(self.table.type_map(top, self.ff))
        if "const_list" in kwargs and "const_vals" in kwargs:
            const_list, const_vals = kwargs["const_list"], kwargs["const_vals"]
        else:
            groups = {}
            for i, rid in enumerate(top.res_ids):
                groups.setdefault(rid, []).append(i)
            const_list = list(groups.values())
            const_vals = [float(np.sum(init_q[g])) for g in const_list]
        has_aux = bool(kwargs.get("has_aux", False))
        extra = {k: kwargs[k] for k in ("precision", "device", "tol", "maxiter") if k in kwargs}
        force = ADMPQeqForce(init_q, r_cut, kappa, (K1, K2, K3), damp_mod=self.damp_mod, const_list=const_list,
                             const_vals=const_vals, neutral_flag=kwargs.get("neutral", True),
                             slab_flag=kwargs.get("slab", False), constQ=kwargs.get("constQ", True),
                             pbc_flag=not no_cut, has_aux=has_aux, **extra)
        self.qeq_force = force
        qeq_energy = force.generate_get_energy()
        mscales = np.array([0.0, 0.0, self.coulomb14scale, 1.0, 1.0, 1.0])

        def potential_fn(positions, box, pairs, params, aux=None):
            p = _params(params, self.name)
            eta = p["eta"][map_idx.to(p["eta"].device)]
            chi = p["chi"][map_idx.to(p["chi"].device)]
            J = p["J"][map_idx.to(p["J"].device)]
            if has_aux:
                return qeq_energy(positions, box, pairs, mscales, eta, chi, J, aux)
            return qeq_energy(positions, box, pairs, mscales, eta, chi, J)

        self._potential = potential_fn
        return potential_fn


GENERATORS = {
    "ADMPPmeForce": ADMPPmeGenerator,
    "ADMPDispPmeForce": ADMPDispPmeGenerator,
    "ADMPDispForce": ADMPDispGenerator,
    "QqTtDampingForce": QqTtDampingGenerator,
    "SlaterDampingForce": SlaterDampingGenerator,
    "SlaterExForce": SlaterExGenerator,
    "SlaterSrEsForce": SlaterSrEsGenerator,
    "SlaterSrPolForce": SlaterSrPolGenerator,
    "SlaterSrDispForce": SlaterSrDispGenerator,
    "SlaterDhfForce": SlaterDhfGenerator,
    "ADMPQeqForce": ADMPQeqGenerator,
}
