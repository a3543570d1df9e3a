"""Drop-in counterpart of DMFF's CoulombPMEForce (dmff/classical/inter.py:269-330),
the charge-only (lmax 0) caller of the ADMP PME path, on the MI355X library.

    force = CoulombPMEForce(r_cut, init_charges, kappa, (K1, K2, K3), topology_matrix=None)
    get_energy = force.generate_get_energy()
    E = get_energy(positions, box, pairs, mscales)        # topology_matrix None
    E = get_energy(positions, box, pairs, bcc, mscales)   # charges = init + top_mat @ bcc

Units as the reference: r_cut and positions/box in nm, kappa in 1/nm (the
generator's gromacs or openmm rule, dmff/admp/pme.py:426-492); the energy in
kJ/mol carries gradients with respect to positions, box, and charges or bcc.
"""
import numpy as np
import torch

from ._autograd import EnergyWithGrads, as_t
from .engine import Engine

F64 = torch.float64


class CoulombPMEForce:
    def __init__(self, r_cut, init_charges, kappa, K, pme_order=6, topology_matrix=None, *, precision="fp64",
                 device="cuda"):
        if pme_order != 6:
            raise AssertionError("PME order other than 6 is not supported")
        self.r_cut = float(r_cut)
        self.init_charges = as_t(init_charges)
        self.lmax = 0
        self.kappa = float(kappa)
        self.K1, self.K2, self.K3 = (int(k) for k in K)
        self.pme_order = pme_order
        self.top_mat = None if topology_matrix is None else as_t(topology_matrix)
        self.precision = precision
        self.device = device
        self._engine = None

    def _eng(self, n):
        if self._engine is None or self._engine.n_atoms != n:
            e = Engine(n, 0, False, True, self.kappa / 10.0, (self.K1, self.K2, self.K3), self.r_cut * 10.0,
                       precision=self.precision, device=self.device)
            e.set_topology(np.full(n, 5), np.full((n, 3), -1))
            self._engine = e
        return self._engine

    def _energy(self, positions, box, pairs, charges, mscales):
        pos = as_t(positions)
        bx = as_t(box)
        q = as_t(charges).reshape(-1)
        eng = self._eng(pos.shape[0])
        eng.set_pairs(pairs)

        def run(want_grad):
            E, g, _, _ = eng.compute(pos * 10.0, bx * 10.0, q.reshape(-1, 1), mScales=mscales, want_grad=want_grad)
            if not want_grad:
                return E.reshape(()), (None, None, None)
            # d/d(nm) = 10 d/d(A) for positions and box
            return E.reshape(()), (g["pos"][0] * 10.0, g["box"][0] * 10.0, g["Q"][0][:, 0])

        E = EnergyWithGrads.apply(run, pos, bx, q)
        return E.to(pos.device) if E.device != pos.device else E

    def generate_get_energy(self):
        """inter.py:290-330."""
        if self.top_mat is None:
            def get_energy(positions, box, pairs, mscales):
                return self._energy(positions, box, pairs, self.init_charges, mscales)
        else:
            def get_energy(positions, box, pairs, bcc, mscales):
                q = self.init_charges.to(as_t(bcc).device) + (self.top_mat.to(as_t(bcc).device) @ as_t(bcc)).reshape(-1)
                return self._energy(positions, box, pairs, q, mscales)
        return get_energy
