"""MD integration of the ADMP path, mirroring the reference's OpenMM plugin
(backend/openmm_dmff_plugin/platforms/cuda/src/CudaDMFFKernels.cpp:118-220,
kernels/DMFFForce.cu:1-10) on the MI355X library:

  * positions, box, parameters, the neighbour list and the induced dipoles stay
    on the device; the list is rebuilt on the GPU every step (the plugin builds
    it on the host each step, computeNeighborListVoxelHash);
  * the induced dipoles of the previous step warm-start the next solve (the
    plugin's last_U_ind);
  * forces (-dE/dpos, kJ/mol/nm) are added to OpenMM-style fixed-point force
    buffers (long long, 2^32 units, layout [x[padded], y[padded], z[padded]])
    by admp_add_forces_fixed, the counterpart of the plugin's addForces kernel.

VelocityVerlet is a minimal NVE integrator (nm, ps, amu, kJ/mol) on the same
buffers, for testing and examples.
"""
import ctypes

import numpy as np
import torch

from .. import _lib as L
from .engine import Engine
from .pme import setup_ewald_parameters

F64 = torch.float64
FIXED_POINT = 4294967296.0   # 2^32, OpenMM's force-buffer scale


class DMFFForce:
    """ADMP multipolar PME (+ induced dipoles) as an MD force.

    system: any object with n_atoms, axis_types, axis_idx, bonded ({(i, j): nbond}),
    Q_local (N, 9), pol, thole (per atom), mScales, pScales, dScales; rc in A."""

    def __init__(self, system, box_nm, rc, ethresh=5e-4, precision="fp64", tol=1e-8, device="cuda", nblist_every=1):
        """nblist_every: rebuild the pair list every that many evaluations (1, the
        plugin's behaviour, every step; 0 once only: a fixed list, on which the
        energy is smooth because listed pairs carry no cutoff, pme.py:1052-1173)."""
        self.dev = torch.device(device)
        self.nblist_every = int(nblist_every)
        self.n_eval = 0
        box_A = np.asarray(box_nm, dtype=np.float64) * 10.0
        kappa, K1, K2, K3 = setup_ewald_parameters(rc, ethresh, box_A)
        self.eng = Engine(system.n_atoms, 2, True, True, kappa, (K1, K2, K3), rc, precision=precision, device=device)
        self.eng.set_topology(system.axis_types, system.axis_idx, system.bonded)
        t = lambda x: torch.as_tensor(np.asarray(x), dtype=F64, device=self.dev)
        self.Q, self.pol, self.thole = t(system.Q_local), t(system.pol), t(system.thole)
        self.scales = (system.mScales, system.pScales, system.dScales)
        self.n_atoms = system.n_atoms
        self.tol = tol
        self.last_U = None
        self.last_info = None
        self.lib = L.load()

    def execute(self, positions_nm, box_nm, force_buffers=None, padded=None, atom_index=None, include_energy=True):
        """One force evaluation (CudaCalcDMFFForceKernel::execute): returns the
        energy (kJ/mol, a Python float; 0 when include_energy is False) and adds the
        forces to force_buffers (int64 device tensor of 3 * padded) when given."""
        pos = positions_nm.to(self.dev, F64) * 10.0
        box = torch.as_tensor(box_nm, dtype=F64, device=self.dev) * 10.0
        if self.n_eval == 0 or (self.nblist_every > 0 and self.n_eval % self.nblist_every == 0):
            self.eng.build_neighbors(pos, box)
        self.n_eval += 1
        E, g, U, info = self.eng.compute(pos, box, self.Q, self.pol, self.thole, *self.scales, U_init=self.last_U,
                                         solver="pcg", tol=self.tol, maxiter=1000, want_grad=force_buffers is not None)
        self.last_U = U
        self.last_info = info
        if force_buffers is not None:
            padded = self.n_atoms if padded is None else int(padded)
            dE = g["pos"][0].contiguous()
            ai = ctypes.c_void_p(atom_index.data_ptr()) if atom_index is not None else ctypes.c_void_p(0)
            stream = torch.cuda.current_stream(self.dev).cuda_stream if self.dev.type == "cuda" else 0
            st = self.lib.admp_add_forces_fixed(ctypes.c_void_p(dE.data_ptr()), 10.0, ai,
                                                ctypes.c_void_p(force_buffers.data_ptr()), self.n_atoms, padded,
                                                ctypes.c_void_p(stream))
            if st != L.ADMP_OK:
                raise L.ADMPError(f"admp_add_forces_fixed failed with status {st}")
        return float(E[0]) if include_energy else 0.0


class VelocityVerlet:
    """NVE velocity Verlet on the device (nm, ps, amu, kJ/mol) with forces taken
    from fixed-point buffers, as an OpenMM integrator would read them."""

    def __init__(self, force, positions_nm, masses_amu, box_nm, dt_ps, velocities=None):
        self.force = force
        self.dev = force.dev
        self.x = torch.as_tensor(np.asarray(positions_nm), dtype=F64, device=self.dev).clone()
        self.m = torch.as_tensor(np.asarray(masses_amu), dtype=F64, device=self.dev).reshape(-1, 1)
        self.v = torch.zeros_like(self.x) if velocities is None else \
            torch.as_tensor(np.asarray(velocities), dtype=F64, device=self.dev).clone()
        self.box = box_nm
        self.dt = float(dt_ps)
        n = self.x.shape[0]
        self.buf = torch.zeros(3 * n, dtype=torch.int64, device=self.dev)
        self.epot = self._forces()

    def _forces(self):
        self.buf.zero_()
        e = self.force.execute(self.x, self.box, self.buf)
        n = self.x.shape[0]
        self.f = self.buf.to(F64).reshape(3, n).t().contiguous() / FIXED_POINT
        return e

    def kinetic(self):
        return float(0.5 * torch.sum(self.m * self.v * self.v))

    def step(self, n=1):
        for _ in range(n):
            self.v += 0.5 * self.dt * self.f / self.m
            self.x += self.dt * self.v
            self.epot = self._forces()
            self.v += 0.5 * self.dt * self.f / self.m
        return self.epot + self.kinetic()
