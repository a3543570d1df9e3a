"""Drop-in counterpart of DMFF's ADMPQeqForce (dmff/admp/qeq.py:200-405), charge
equilibration on the lmax-0 caller of the ADMP PME path.

    force = ADMPQeqForce(init_q, r_cut, kappa, (K1, K2, K3), damp_mod=3,
                         const_list=[[0, 1, ...], ...], const_vals=[...],
                         neutral_flag=True, slab_flag=False, constQ=True, pbc_flag=True)
    get_energy = force.generate_get_energy()
    E = get_energy(positions, box, pairs, mscales, eta, chi, J)            # has_aux False
    E, aux = get_energy(positions, box, pairs, mscales, eta, chi, J, aux)  # aux {"q", "lagmt"}

Units as the reference: positions, box and r_cut in nm, kappa in 1/nm, energy in
kJ/mol.  The energy is the reference's E_full (qeq.py:323-337): constraint +
short-range damping (E_sr/E_sr2/E_sr3) + site (E_site/E_site2/E_site3) +
Coulomb + optional slab correction, at the charges and Lagrange multipliers
where its gradient with respect to (q, lagmt) vanishes.

E_full is quadratic in (q, lagmt), so that stationary point solves one
symmetric (in general indefinite) linear system.  The reference finds it with
scipy's hybr root finder (tol 1e-10); here MINRES (Paige and Saunders, 1975)
solves it on the device.  Each MINRES step applies the system's matrix once:
the Coulomb part is the library's dE/dq at the direction (PME for pbc_flag,
the plain pair sum otherwise, both through the HIP library's lmax-0 path), the
other terms are pair and site expressions evaluated on the device.  The
solution is held fixed (the reference's stop_gradient), so gradients with
respect to positions, box, chi, J and eta are the partial derivatives of E_full
at that point: the Coulomb ones come from the library, the others from torch
autograd of the short expressions.
"""
import math

import numpy as np
import torch

from ._autograd import EnergyWithGrads, as_t
from .engine import Engine

F64 = torch.float64
DIELECTRIC = 1389.35455846      # dmff/common/constants.py:3


def minres(apply_a, b, x0, tol, maxiter):
    """MINRES for a symmetric matrix given as apply_a(v): returns (x, iterations,
    estimated relative residual).  Stops when the recurrence's residual norm
    falls below tol * |b - A x0|."""
    x = x0.clone()
    r1 = b - apply_a(x)
    beta1 = float(torch.linalg.norm(r1))
    if beta1 == 0.0:
        return x, 0, 0.0
    y = r1.clone()
    r2 = r1.clone()
    oldb, beta = 0.0, beta1
    dbar, epsln, phibar = 0.0, 0.0, beta1
    cs, sn = -1.0, 0.0
    w = torch.zeros_like(b)
    w2 = torch.zeros_like(b)
    it = 0
    while it < maxiter:
        it += 1
        v = y / beta
        y = apply_a(v)
        if it >= 2:
            y = y - (beta / oldb) * r1
        alfa = float(torch.dot(v, y))
        y = y - (alfa / beta) * r2
        r1, r2 = r2, y
        oldb, beta = beta, float(torch.linalg.norm(y))
        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta
        dbar = -cs * beta
        gamma = max(math.hypot(gbar, beta), 1e-300)
        cs, sn = gbar / gamma, beta / gamma
        phi = cs * phibar
        phibar = sn * phibar
        w1, w2 = w2, w
        w = (v - oldeps * w1 - delta * w2) / gamma
        x = x + phi * w
        if phibar <= tol * beta1 or beta == 0.0:
            break
    return x, it, phibar / beta1


def _mask_pairs(pairs, n):
    """Rows the reference keeps (pair_buffer_scales, dmff/utils.py:77-84: i < j)."""
    p = pairs[:, :2].to(torch.int64)
    return p[(p[:, 0] < p[:, 1]) & (p[:, 1] < n)]


class ADMPQeqForce:
    """qeq.py:200-319 (constructor arguments as the reference).  Extra keywords:
    precision ("fp64", or "mixed": FP32 Coulomb products, which bound MINRES's
    residual near 1e-6, so tol is raised to at least that), device, tol (MINRES
    relative residual) and maxiter."""

    def __init__(self, init_q, r_cut, kappa, K, damp_mod=3, const_list=[], const_vals=[], neutral_flag=True,
                 slab_flag=False, constQ=True, pbc_flag=True, has_aux=False, *, precision="fp64", device="cuda",
                 tol=1e-12, maxiter=5000):
        if damp_mod not in (1, 2, 3):
            raise ValueError("damp_mod must be 1, 2 or 3")
        const_vals = np.asarray(const_vals, dtype=np.float64)
        if neutral_flag and len(const_vals):
            const_vals = const_vals - np.sum(const_vals) / len(const_vals)
        if len(const_list) != len(const_vals):
            raise AssertionError("const_list and const_vals must have the same length")
        self.device = torch.device(device)
        self.has_aux = has_aux
        self.init_q = as_t(init_q).reshape(-1)
        self.n_atoms = int(self.init_q.shape[0])
        n, m = self.n_atoms, len(const_list)
        # constraint matrix (m, n): row k sums group k (qeq.py:52-75)
        C = torch.zeros((m, n), dtype=F64)
        for k, g in enumerate(const_list):
            g = [int(i) for i in g if int(i) < n]
            C[k, g] = 1.0
        self.C = C.to(self.device)
        self.const_vals = torch.as_tensor(const_vals, dtype=F64, device=self.device)
        self.init_lagmt = torch.ones(m, dtype=F64)
        self.r_cut = float(r_cut)
        self.kappa = float(kappa) if pbc_flag else 0.0
        self.K = tuple(int(k) for k in K)
        self.damp_mod, self.neutral_flag, self.slab_flag = damp_mod, neutral_flag, slab_flag
        self.constQ, self.pbc_flag = constQ, pbc_flag
        self.precision, self.tol, self.maxiter = precision, float(tol), int(maxiter)
        if precision != "fp64":
            self.tol = max(self.tol, 1e-6)
        self._engine = None
        self.last_iterations = 0
        self.last_residual = 0.0

    # ---- Coulomb through the library ----------------------------------------
    def _eng(self):
        if self._engine is None:
            n = self.n_atoms
            e = Engine(n, 0, False, self.pbc_flag, self.kappa / 10.0, self.K if self.pbc_flag else (1, 1, 1),
                       self.r_cut * 10.0, precision=self.precision, device=self.device)
            e.set_topology(np.full(n, 5), np.full((n, 3), -1))
            self._engine = e
        return self._engine

    def _coul_box(self, pos_A, box_A):
        """The library's pair kernel takes a periodic box; without pbc_flag (plain
        pair sum, qeq.py:293-316) a cube wider than twice the system's extent
        makes every minimum image the direct difference."""
        if self.pbc_flag:
            return box_A
        ext = float((pos_A.max(dim=0).values - pos_A.min(dim=0).values).max())
        return torch.eye(3, dtype=F64, device=pos_A.device) * (2.0 * ext + 4.0 * self.r_cut * 10.0 + 10.0)

    def _coul(self, pos_A, box_A, q, mscales, want_grad):
        E, g, _, _ = self._eng().compute(pos_A, self._coul_box(pos_A, box_A), q.reshape(-1, 1), mScales=mscales,
                                         want_grad=want_grad)
        return E, g

    # ---- the other terms of E_full, on the device ---------------------------
    def _pair_geometry(self, pos_nm, box_nm, pp):
        """ds_pairs (qeq.py:161-174): distances in nm, minimum image when periodic."""
        d = pos_nm[pp[:, 0]] - pos_nm[pp[:, 1]]
        if self.pbc_flag:
            s = d @ torch.linalg.inv(box_nm)
            s = s - torch.floor(s + 0.5)
            d = s @ box_nm
        return torch.linalg.norm(d + 1e-64, dim=1)

    def _e_rest(self, q, lagmt, chi, J, pos_nm, box_nm, pp, eta):
        """E_full without the Coulomb term (qeq.py:323-337)."""
        sums = self.C @ q
        if self.constQ:
            e = torch.sum((sums - self.const_vals) * lagmt)          # E_constQ, :71-74
        else:
            e = torch.sum(sums * self.const_vals)                     # E_constP, :77-79
        dm = self.damp_mod
        if dm != 1:
            ds = self._pair_geometry(pos_nm, box_nm, pp) * 10.0       # Angstrom, as E_full passes it
            i, j = pp[:, 0], pp[:, 1]
            if dm == 2:                                               # E_sr2, :96-105
                es = torch.sqrt(2 * (eta[i] ** 2 + eta[j] ** 2))
            else:                                                     # E_sr3, :108-120
                es = torch.sqrt(eta[i] ** 2 + eta[j] ** 2 + 1e-64)
            live = es > 1e-4                                          # eta_piecewise, :177-185
            piece = torch.where(live, torch.special.erfc(ds / torch.where(live, es, torch.ones_like(es))),
                                torch.zeros_like(es))
            e = e + torch.sum(-piece * DIELECTRIC * q[i] * q[j] / ds)
            ok = eta > 1e-4                                           # etainv_piecewise, :189-197
            inv = torch.where(ok, 1.0 / torch.where(ok, eta, torch.ones_like(eta)), torch.zeros_like(eta))
            e = e + torch.sum(inv / math.sqrt(2 * math.pi) * DIELECTRIC * q * q)
            if dm == 2:                                               # E_site2, :128-131
                e = e + torch.sum((chi * q + 0.5 * J * q ** 2) * 96.4869)
            else:                                                     # E_site3, :134-137
                e = e + torch.sum(chi * q * 4.184 + J * q ** 2 * DIELECTRIC * 2 * math.pi)
        if self.slab_flag:                                            # E_corr, :140-158
            pos_A, box_A = pos_nm * 10.0, box_nm * 10.0
            V = torch.linalg.det(box_A)
            z = pos_A[:, 2]
            Mz, Qt = torch.sum(q * z), torch.sum(q)
            Lz = torch.linalg.norm(box_A[2])    # the reference's box[3] of a 3x3 box (JAX clamps to box[2])
            e = e + 2 * math.pi / V * DIELECTRIC * (Mz ** 2 - Qt * torch.sum(q * z ** 2) - Qt ** 2 * Lz ** 2 / 12)
            if self.neutral_flag:
                e = e + (-math.pi / (2 * V * (self.kappa / 10.0) ** 2) * DIELECTRIC) * Qt ** 2
        return e

    # ---- the stationary point -------------------------------------------------
    def _solve(self, pos_nm, box_nm, pp, mscales, eta, chi, J, x0):
        n, m = self.n_atoms, self.C.shape[0]
        nl = m if self.constQ else 0
        pos_A, box_A = pos_nm * 10.0, box_nm * 10.0
        lag_fixed = torch.ones(m, dtype=F64, device=self.device)

        def grad_rest(q):
            q = q.detach().requires_grad_(True)
            with torch.enable_grad():
                e = self._e_rest(q, lag_fixed, chi, J, pos_nm, box_nm, pp, eta)
                return torch.autograd.grad(e, q)[0]

        g0 = grad_rest(torch.zeros(n, dtype=F64, device=self.device))   # linear part (+ C^T 1 for constQ)
        lin = g0 - (self.C.t() @ lag_fixed if self.constQ else 0.0)

        def apply_a(x):
            q = x[:n]
            _, g = self._coul(pos_A, box_A, q, mscales, True)
            out_q = g["Q"][0][:, 0] + (grad_rest(q) - g0)
            if not self.constQ:
                return out_q
            lam = x[n:]
            return torch.cat((out_q + self.C.t() @ lam, self.C @ q))

        # constQ: [A C^T; C 0] [q; lagmt] = [-lin; const_vals]; constP: A q = -(lin + C^T const_vals) = -g0
        rhs = torch.cat((-lin, self.const_vals)) if self.constQ else -g0
        x, it, res = minres(apply_a, rhs, x0[: n + nl], self.tol, self.maxiter)
        self.last_iterations, self.last_residual = it, res
        q = x[:n]
        lag = x[n:] if self.constQ else x0[n:]
        return q, lag

    def generate_get_energy(self):
        """qeq.py:321-405."""

        def get_energy(positions, box, pairs, mscales, eta, chi, J, aux=None):
            dev = self.device
            pos_in = as_t(positions)
            pos = pos_in.to(dev)
            bx = as_t(box).to(dev)
            eta, chi, J = (as_t(t).to(dev).reshape(-1) for t in (eta, chi, J))
            pt = pairs if isinstance(pairs, torch.Tensor) else torch.as_tensor(np.asarray(pairs))
            pp = _mask_pairs(pt.to(dev), self.n_atoms)
            self._eng().set_pairs(pt)
            if self.has_aux and aux is not None:
                x0 = torch.cat((as_t(aux["q"]).reshape(-1), as_t(aux["lagmt"]).reshape(-1))).to(dev)
            else:
                x0 = torch.cat((self.init_q, self.init_lagmt)).to(dev)
            with torch.no_grad():
                q, lag = self._solve(pos.detach(), bx.detach(), pp, mscales, eta.detach(), chi.detach(),
                                     J.detach(), x0)

            def run(want_grad):
                E, g = self._coul(pos.detach() * 10.0, bx.detach() * 10.0, q, mscales, want_grad)
                if not want_grad:
                    return E.reshape(()), (None, None)
                gb = g["box"][0] * 10.0 if self.pbc_flag else torch.zeros_like(bx)
                return E.reshape(()), (g["pos"][0] * 10.0, gb)

            e_coul = EnergyWithGrads.apply(run, pos, bx)
            e = e_coul + self._e_rest(q, lag, chi, J, pos, bx, pp, eta)
            e = e.to(pos_in.device) if e.device != pos_in.device else e
            if self.has_aux:
                if aux is None:
                    aux = {}
                aux["q"], aux["lagmt"] = q, lag
                return e, aux
            return e

        return get_energy
