"""Drop-in counterpart of DMFF's ADMPDispPmeForce (dmff/admp/disp_pme.py:13-77) on
the MI355X HIP library.

Same constructor (box, rc, ethresh, pmax, lpme) and get_energy(positions, box,
pairs, c_list, mScales) as the reference backend (Angstrom, kJ/mol; c_list (N,3)
holds sqrt(C6), sqrt(C8), sqrt(C10) per atom, the generator's convention).  The
energy is the backend's positive sum (the generator negates it, see
generators.py) and carries gradients with respect to positions, box and c_list
through the library's hand-written backward.
"""
import torch

from ._autograd import EnergyWithGrads, as_t
from .engine import Engine
from .pme import setup_ewald_parameters

F64 = torch.float64


class ADMPDispPmeForce:
    """Dispersion PME calculator (dmff/admp/disp_pme.py:20-50).

    Extra keyword arguments (not in the reference): precision ("fp64" | "mixed"),
    device.  The library context is created on the first call (its size follows
    the number of atoms) and kept while that size is unchanged."""

    def __init__(self, box, rc, ethresh, pmax, lpme=True, *, precision="fp64", device="cuda"):
        self.rc = float(rc)
        self.ethresh = ethresh
        self.pmax = int(pmax)
        self.lpme = lpme
        self.box0 = box
        if lpme:
            self.kappa, self.K1, self.K2, self.K3 = setup_ewald_parameters(rc, ethresh, box)
        else:
            self.kappa, self.K1, self.K2, self.K3 = 0.0, 0, 0, 0
        self.pme_order = 6
        self.precision = precision
        self.device = device
        self._engine = None

    def update_env(self, attr, val):
        """disp_pme.py:53-58: change an environment attribute and rebuild."""
        setattr(self, attr, val)
        if attr in ("rc", "ethresh", "lpme"):
            if self.lpme:
                self.kappa, self.K1, self.K2, self.K3 = setup_ewald_parameters(self.rc, self.ethresh, self.box0)
            else:
                self.kappa, self.K1, self.K2, self.K3 = 0.0, 0, 0, 0
        self._engine = None

    def _eng(self, n):
        if self._engine is None or self._engine.n_atoms != n:
            e = Engine(n, 0, False, False, 0.0, (0, 0, 0), self.rc, precision=self.precision,
                       disp_kappa=self.kappa if self.lpme else 0.0,
                       dK=(self.K1, self.K2, self.K3) if self.lpme else (0, 0, 0), disp_pmax=self.pmax,
                       device=self.device)
            e.set_topology(None, None, {})
            self._engine = e
        return self._engine

    def get_energy(self, positions, box, pairs, c_list, mScales):
        """energy_disp_pme (disp_pme.py:80-130): real + reciprocal (p = 6, 8, 10 up to
        pmax) + self; real space only when lpme is False."""
        pos = as_t(positions)
        bx = as_t(box)
        c = as_t(c_list)
        eng = self._eng(pos.shape[0])
        eng.set_pairs(pairs)

        def run(want_grad):
            E, g = eng.disp_compute(pos, bx, c, mScales, want_grad=want_grad)
            grads = (g["pos"][0], g["box"][0], g["c"][0]) if want_grad else (None, None, None)
            return E.reshape(()), grads

        E = EnergyWithGrads.apply(run, pos, bx, c)
        return E.to(pos.device) if E.device != pos.device else E

    def get_forces(self, positions, *args):
        """value_and_grad of get_energy with respect to positions (disp_pme.py:75)."""
        pos = as_t(positions).detach().clone().requires_grad_(True)
        E = self.get_energy(pos, *args)
        (g,) = torch.autograd.grad(E, pos)
        return E, g
