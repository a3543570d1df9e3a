"""Drop-in counterpart of DMFF's ADMPPmeForce (dmff/admp/pme.py:36-423) on the
MI355X HIP library.

Same constructor arguments and get_energy signatures as the reference backend
(Angstrom, e, kJ/mol; U_init and aux["U_ind"] in e*nm as the reference).
Energies are torch scalars that carry gradients with respect to positions,
box, Q_local, pol and tholes through hand-written backward kernels (the
induced dipoles are held fixed, as the reference's stop_gradient does).
"""
import ctypes
import math

import numpy as np
import torch

from .. import _lib as L
from .engine import Engine, _dev

F64 = torch.float64
# dmff/admp/settings.py:2-3
POL_CONV = 1.0
MAX_N_POL = 30


def setup_ewald_parameters(rc, ethresh, box=None, spacing=None, method="openmm"):
    """kappa (1/A) and K1..K3 (dmff/admp/pme.py:426-492)."""
    if method == "openmm":
        box = np.asarray(box.detach().cpu() if isinstance(box, torch.Tensor) else box, dtype=np.float64)
        kappa = math.sqrt(-math.log(2 * ethresh)) / rc
        K = [int(math.ceil(2 * kappa * box[d, d] / 3 / ethresh ** 0.2)) for d in range(3)]
        return kappa, K[0], K[1], K[2]
    if method == "gromacs":
        kappa = 5.0
        i = 0
        while math.erfc(kappa * rc) > ethresh:
            i += 1
            kappa *= 2
        low, high = 0.0, kappa
        for _ in range(i + 60):
            kappa = (low + high) / 2
            if math.erfc(kappa * rc) > ethresh:
                low = kappa
            else:
                high = kappa
        box = np.asarray(box, dtype=np.float64)
        K = [int(math.ceil(box[d, d] / spacing)) for d in range(3)]
        return kappa, K[0], K[1], K[2]
    raise ValueError(f"Invalid method: {method}.Valid methods: 'openmm', 'gromacs'")


class TypeReducer:
    """Sums per-atom gradients into per-type gradients, the adjoint of the
    ``Q_local[map_atomtype]`` expansion (dmff/generators/admp.py:1372-1375).

    One library launch (``admp_type_reduce``): a workgroup per (type, component)
    sums its atoms in a fixed order, so results are deterministic.  Inputs are
    gradients (no autograd through the reduction)."""

    def __init__(self, type_map, n_types=None, device=None):
        m = np.asarray(type_map, dtype=np.int64)
        self.n_types = int(m.max()) + 1 if n_types is None else int(n_types)
        self._map_host = torch.as_tensor(m.astype(np.int32))
        self._maps = {}
        if device is not None:
            self._map_on(torch.device(device))
        self.lib = L.load()

    @property
    def map(self):
        return self._map_host

    def _map_on(self, dev):
        key = str(dev)
        if key not in self._maps:
            self._maps[key] = self._map_host.to(dev)
        return self._maps[key]

    def __call__(self, per_atom, axis=0):
        """per_atom: (..., N, ...) with atoms on `axis`; returns the same with types on `axis`.
        per_atom must live in device memory (a CUDA tensor): the reduction is a GPU launch."""
        x = per_atom.detach().to(torch.float64)
        shape = list(x.shape)
        axis = axis % len(shape)
        n = shape[axis]
        if n != self._map_host.numel():
            raise ValueError(f"{n} atoms on axis {axis}, type map has {self._map_host.numel()}")
        if not x.is_cuda and not L.HOST_MEMORY:
            raise ValueError("TypeReducer: per_atom must be a CUDA tensor (the reduction runs on the GPU)")
        tmap = self._map_on(x.device)
        pre = int(np.prod(shape[:axis])) if axis else 1
        d = int(np.prod(shape[axis + 1:])) if axis + 1 < len(shape) else 1
        x = x.contiguous()
        out = torch.empty([pre, self.n_types, d], dtype=torch.float64, device=x.device)
        stream = torch.cuda.current_stream(x.device).cuda_stream if x.is_cuda else 0
        st = self.lib.admp_type_reduce(ctypes.c_void_p(x.data_ptr()), pre, n, d, ctypes.c_void_p(tmap.data_ptr()),
                                       self.n_types, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
        if st != L.ADMP_OK:
            raise L.ADMPError(f"admp_type_reduce failed with status {st}")
        shape[axis] = self.n_types
        return out.reshape(shape)


class _ADMPEnergy(torch.autograd.Function):
    """E(pos, box, Q_local, pol, thole) with the gradients computed by the library
    in the same call (value-and-grad), returned on backward."""

    @staticmethod
    def forward(ctx, pos, box, Q_local, pol, thole, run):
        need = any(ctx.needs_input_grad[:5])
        E, g = run(need)
        ctx.g = g
        ctx.shapes = (pos.shape, box.shape, Q_local.shape,
                      pol.shape if pol is not None else None, thole.shape if thole is not None else None)
        ctx.devs = tuple(x.device if x is not None else None for x in (pos, box, Q_local, pol, thole))
        return E

    @staticmethod
    def backward(ctx, gE):
        g = ctx.g
        sp, sb, sq, sa, st = ctx.shapes
        if not g:
            return None, None, None, None, None, None
        F = gE.reshape(-1).shape[0]
        gE = gE.to(g["pos"].device)
        w = gE.reshape(F, 1, 1)
        gp = (w * g["pos"]).reshape(sp) if ctx.needs_input_grad[0] else None
        gb = (w * g["box"]).reshape(sb) if ctx.needs_input_grad[1] else None
        gq = None
        if ctx.needs_input_grad[2]:
            q = (w * g["Q"]).sum(0)
            gq = torch.zeros(sq, dtype=q.dtype, device=q.device)
            gq[:, : q.shape[1]] = q
        ga = (gE.reshape(F, 1) * g["pol"]).sum(0).reshape(sa) if (sa is not None and ctx.needs_input_grad[3]) else None
        gt = (gE.reshape(F, 1) * g["thole"]).sum(0).reshape(st) if (st is not None and ctx.needs_input_grad[4]) else None
        out = [x.to(d) if x is not None else None for x, d in zip((gp, gb, gq, ga, gt), ctx.devs)]
        return (*out, None)


class ADMPPmeForce:
    """Multipolar PME calculator (dmff/admp/pme.py:36-114 constructor arguments).

    Extra keyword arguments (not in the reference): precision ("fp64" | "mixed"),
    solver ("reference" | "pcg") and tol for the induced-dipole solve, n_frames
    for batched evaluation, device.  With the defaults the solver is the
    reference's: Jacobi to POL_CONV = 1.0 / MAX_N_POL = 30 sweeps, or exactly
    steps_pol sweeps when steps_pol is set.
    """

    def __init__(self, box, map_atomtype, axis_type, axis_indices, rc, ethresh, lmax, lpol=False, lpme=True,
                 steps_pol=None, has_aux=False, *, precision="fp64", solver="reference", tol=None, maxiter=None,
                 n_frames=1, device="cuda"):
        self.map_atomtype = map_atomtype
        self.axis_type = np.asarray(axis_type)
        self.axis_indices = None if axis_indices is None else np.asarray(axis_indices)
        self.rc = float(rc)
        self.ethresh = ethresh
        self.lmax = int(lmax)
        self.lpme = lpme
        if lpme:
            self.kappa, self.K1, self.K2, self.K3 = setup_ewald_parameters(rc, ethresh, box)
        else:
            self.kappa, self.K1, self.K2, self.K3 = 0.0, 0, 0, 0
        self.pme_order = 6
        self.lpol = lpol
        self.steps_pol = steps_pol
        self.n_atoms = len(axis_type)
        self.has_aux = has_aux
        self.precision = precision
        self.solver = solver
        self.tol = tol
        self.maxiter = maxiter
        self.n_frames = n_frames
        self.device = device
        self._engine = None
        self.last_info = None

    # -------------------------------------------------------------- engine
    @property
    def engine(self):
        if self._engine is None:
            e = Engine(self.n_atoms, self.lmax, self.lpol, self.lpme, self.kappa, (self.K1, self.K2, self.K3),
                       self.rc, n_frames=self.n_frames, precision=self.precision, device=self.device)
            ai = self.axis_indices if self.axis_indices is not None else np.full((self.n_atoms, 3), -1)
            e.set_topology(self.axis_type, ai)
            self._engine = e
        return self._engine

    def update_env(self, attr, val):
        """pme.py:313-319: change an environment attribute and rebuild."""
        setattr(self, attr, val)
        if attr in ("rc", "ethresh", "lpme") and self.lpme and getattr(self, "_box0", None) is not None:
            self.kappa, self.K1, self.K2, self.K3 = setup_ewald_parameters(self.rc, self.ethresh, self._box0)
        self._engine = None

    def _solver_args(self):
        if not self.lpol:
            return {}
        if self.steps_pol is not None:
            return dict(solver="steps", steps=int(self.steps_pol), tol=0.0, maxiter=int(self.steps_pol))
        if self.solver == "reference":
            return dict(solver="reference", tol=POL_CONV if self.tol is None else self.tol,
                        maxiter=MAX_N_POL if self.maxiter is None else self.maxiter)
        return dict(solver="pcg", tol=1e-6 if self.tol is None else self.tol,
                    maxiter=500 if self.maxiter is None else self.maxiter)

    def _energy(self, positions, box, pairs, Q_local, pol, tholes, mScales, pScales, dScales, U_init_A):
        eng = self.engine
        dev = eng.device
        if pairs is not None:
            eng.set_pairs(pairs)
        pos_t = positions if isinstance(positions, torch.Tensor) else torch.as_tensor(np.asarray(positions), dtype=F64)
        box_t = box if isinstance(box, torch.Tensor) else torch.as_tensor(np.asarray(box), dtype=F64)
        Q_t = Q_local if isinstance(Q_local, torch.Tensor) else torch.as_tensor(np.asarray(Q_local), dtype=F64)
        pol_t = thole_t = None
        if self.lpol:
            pol_t = pol if isinstance(pol, torch.Tensor) else torch.as_tensor(np.asarray(pol), dtype=F64)
            thole_t = tholes if isinstance(tholes, torch.Tensor) else torch.as_tensor(np.asarray(tholes), dtype=F64)
        sargs = self._solver_args()
        out = {}

        def run(want_grad):
            E, g, U, info = eng.compute(pos_t, box_t, Q_t, pol_t, thole_t, mScales, pScales, dScales,
                                        U_init=U_init_A, want_grad=want_grad, **sargs)
            out["U"], out["info"] = U, info
            return (E if self.n_frames > 1 else E.reshape(())), g

        E = _ADMPEnergy.apply(pos_t, box_t, Q_t, pol_t, thole_t, run)
        self.last_info = out["info"]
        if E.device != pos_t.device:
            E = E.to(pos_t.device)
        return E, out

    # -------------------------------------------------------------- reference API
    def get_energy(self, positions, box, pairs, Q_local, *args, U_init=None, aux=None):
        """Non-polarisable: get_energy(positions, box, pairs, Q_local, mScales) (pme.py:120-141).
        Polarisable: get_energy(positions, box, pairs, Q_local, pol, tholes, mScales, pScales,
        dScales, U_init=None, aux=None) (pme.py:186-232); U_init in e*nm."""
        if not self.lpol:
            (mScales,) = args[:1]
            E, _ = self._energy(positions, box, pairs, Q_local, None, None, mScales, None, None, None)
            return E
        pol, tholes, mScales, pScales, dScales = args[:5]
        if len(args) > 5:
            U_init = args[5]
        U0 = None
        if U_init is not None:
            U0 = _dev(U_init, self.engine.device) * 10.0
        E, out = self._energy(positions, box, pairs, Q_local, pol, tholes, mScales, pScales, dScales, U0)
        if aux is not None:
            info = out["info"]
            U = out["U"] * 0.1
            aux["U_ind"] = U[0] if self.n_frames == 1 else U
            aux["lconverg"] = all(info["converged"])
            aux["n_cycle"] = info["n_iter"][0] if self.n_frames == 1 else info["n_iter"]
            return E, aux
        return E

    def Q_local_to_global(self, positions, box, Q_local):
        """pme.py:298-311: per-type local harmonic multipoles (expanded with
        map_atomtype) rotated into global Cartesian [q, mu, Theta_xx, Theta_yy,
        Theta_zz, Theta_xy, Theta_xz, Theta_yz] (N, 10).  Values only (no autograd)."""
        Q = torch.as_tensor(np.asarray(Q_local.detach().cpu() if isinstance(Q_local, torch.Tensor) else Q_local),
                            dtype=F64)
        Qa = Q[torch.as_tensor(np.asarray(self.map_atomtype))]
        eng = self._single_engine()
        return eng.local_to_global(positions, box, Qa)[0]

    def _single_engine(self):
        if self.n_frames == 1:
            return self.engine
        if getattr(self, "_engine1", None) is None:
            e = Engine(self.n_atoms, self.lmax, self.lpol, self.lpme, self.kappa, (self.K1, self.K2, self.K3), self.rc,
                       n_frames=1, precision=self.precision, device=self.device)
            ai = self.axis_indices if self.axis_indices is not None else np.full((self.n_atoms, 3), -1)
            e.set_topology(self.axis_type, ai)
            self._engine1 = e
        return self._engine1

    def generate_esp(self):
        """pme.py:236-296: returns get_esp(positions, grids, Q_local[, U_ind]) giving
        the potential (kJ/mol/e) of the multipoles (per-type Q_local, frames from
        positions in a 1000-unit cubic box) and induced dipoles (e nm) at the probe
        points, positions and probes in nm, no periodic images.  Values only."""
        lib = L.load()

        def esp(positions, grids, Q_local, U_ind=None):
            eng = self._single_engine()
            dev = eng.device
            pos = _dev(positions, dev).reshape(-1, 3)
            gr = _dev(grids, dev).reshape(-1, 3)
            Qg = self.Q_local_to_global(pos, torch.eye(3, dtype=F64, device=dev) * 1000.0, Q_local).contiguous()
            U = _dev(U_ind, dev).reshape(-1, 3) if (U_ind is not None and self.lpol) else None
            out = torch.empty(gr.shape[0], dtype=F64, device=dev)
            stream = torch.cuda.current_stream(dev).cuda_stream if dev.type == "cuda" else 0
            st = lib.admp_esp(pos.shape[0], ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(Qg.data_ptr()),
                              ctypes.c_void_p(U.data_ptr() if U is not None else 0), gr.shape[0],
                              ctypes.c_void_p(gr.data_ptr()), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
            if st != L.ADMP_OK:
                raise L.ADMPError(f"admp_esp failed with status {st}")
            return out

        if self.lpol:
            return lambda positions, grids, Q_local, U_ind: esp(positions, grids, Q_local, U_ind)
        return lambda positions, grids, Q_local: esp(positions, grids, Q_local)

    def get_forces(self, *args, **kwargs):
        """value_and_grad of get_energy with respect to positions (pme.py:339)."""
        positions = args[0]
        pos = (positions if isinstance(positions, torch.Tensor) else torch.as_tensor(np.asarray(positions), dtype=F64))
        pos = pos.detach().clone().requires_grad_(True)
        E = self.get_energy(pos, *args[1:], **kwargs)
        e = E[0] if isinstance(E, tuple) else E
        (g,) = torch.autograd.grad(e.sum(), pos)
        return E, g
