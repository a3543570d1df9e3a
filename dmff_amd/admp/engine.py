"""One native ADMP context (libadmp_hip.so) driven from PyTorch tensors.

Owns the C context (grids, FFT tables, neighbour lists, solver state) for a
fixed atom count, frame count and PME set-up.  All array arguments are CUDA
float64 tensors; the current torch stream is handed to the library.
"""
import ctypes

import numpy as np
import torch

from .. import _lib as L

F64 = torch.float64


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _dev(x, device, dtype=F64):
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device).contiguous()


class Engine:
    def __init__(self, n_atoms, lmax, lpol, lpme, kappa, K, rc, n_frames=1, precision="fp64",
                 disp_kappa=0.0, dK=(0, 0, 0), disp_pmax=10, max_neighbors=0, device="cuda"):
        self.lib = L.load()
        self.device = torch.device(device)
        cfg = L.Config()
        cfg.n_atoms, cfg.n_frames, cfg.lmax = int(n_atoms), int(n_frames), int(lmax)
        cfg.lpol, cfg.lpme = int(bool(lpol)), int(bool(lpme))
        cfg.precision = L.ADMP_MIXED if precision == "mixed" else L.ADMP_FP64
        cfg.kappa = float(kappa)
        cfg.K1, cfg.K2, cfg.K3 = (int(k) for k in K)
        cfg.pme_order = 6
        cfg.rc = float(rc)
        cfg.max_neighbors = int(max_neighbors)
        cfg.disp_kappa = float(disp_kappa)
        cfg.dK1, cfg.dK2, cfg.dK3 = (int(k) for k in dK)
        cfg.disp_pmax = int(disp_pmax)
        self.cfg = cfg
        self.n_atoms, self.n_frames, self.lmax = cfg.n_atoms, cfg.n_frames, cfg.lmax
        self.nh = (self.lmax + 1) ** 2
        with torch.cuda.device(self.device):
            torch.cuda.init()
            h = ctypes.c_void_p()
            st = self.lib.admp_ctx_create(ctypes.byref(cfg), ctypes.byref(h))
            if st != 0 or not h.value:
                raise L.ADMPError(f"admp_ctx_create failed with status {st}")
        self.h = h
        self._pairs_key = None

    def __del__(self):
        try:
            if getattr(self, "h", None) and self.h.value:
                self.lib.admp_ctx_destroy(self.h)
                self.h = ctypes.c_void_p()
        except Exception:
            pass

    def _stream(self):
        L.check(self.h, self.lib.admp_set_stream(
            self.h, ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)), "admp_set_stream")

    def set_stream_mode(self, caller_mode):
        """caller_mode True: the context works directly on torch's current stream of
        each call (no stream of its own), for many contexts driven from a few
        worker streams; False (default): its own stream, ordered against the caller's."""
        L.check(self.h, self.lib.admp_set_stream_mode(self.h, int(bool(caller_mode))), "admp_set_stream_mode")

    def memory_bytes(self):
        """Device bytes held by the native context."""
        n = ctypes.c_int64()
        L.check(self.h, self.lib.admp_memory_bytes(self.h, ctypes.byref(n)), "admp_memory_bytes")
        return int(n.value)

    def profile(self, enable=True, reset=True):
        L.check(self.h, self.lib.admp_profile(self.h, int(enable), int(reset)), "admp_profile")

    def profile_get(self):
        """{family: (total_ms, launches)} accumulated since the last reset."""
        out = {}
        for k in range(16):
            ms, n, name = ctypes.c_double(), ctypes.c_int64(), ctypes.c_char_p()
            L.check(self.h, self.lib.admp_profile_get(self.h, k, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(name)),
                    "admp_profile_get")
            if name.value:
                out[name.value.decode()] = (ms.value, n.value)
        return out

    def set_topology(self, axis_type=None, axis_idx=None, exclusions=None):
        n = self.n_atoms
        at = np.ascontiguousarray(np.asarray(axis_type if axis_type is not None else np.full(n, 5)), dtype=np.int32)
        ai = np.ascontiguousarray(np.asarray(axis_idx if axis_idx is not None else np.full((n, 3), -1)),
                                  dtype=np.int32).reshape(n, 3)
        if exclusions:
            ex = np.asarray([(i, j, b) for (i, j), b in exclusions.items() if i < j and 1 <= b <= 7], dtype=np.int32)
        else:
            ex = np.zeros((0, 3), dtype=np.int32)
        ex = np.ascontiguousarray(ex.reshape(-1, 3))
        ei, ej, eb = (np.ascontiguousarray(ex[:, k]) for k in range(3))
        st = self.lib.admp_set_topology(self.h, at.ctypes.data_as(ctypes.c_void_p), ai.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.c_int64(len(ex)), ei.ctypes.data_as(ctypes.c_void_p),
                                        ej.ctypes.data_as(ctypes.c_void_p), eb.ctypes.data_as(ctypes.c_void_p))
        L.check(self.h, st, "admp_set_topology")

    def set_pairs(self, pairs, frame=0):
        """Reference-format pairs (Np,3) [i, j, nbond]; rows with i >= j ignored."""
        p = pairs if isinstance(pairs, torch.Tensor) else torch.as_tensor(np.asarray(pairs))
        p = p.to(device=self.device)
        if p.dtype not in (torch.int32, torch.int64):
            p = p.to(torch.int64)
        p = p.contiguous()
        key = (p.data_ptr(), p._version, tuple(p.shape), frame, p.dtype)
        if key == self._pairs_key and isinstance(pairs, torch.Tensor):
            return
        self._stream()
        st = self.lib.admp_set_pairs(self.h, frame, _ptr(p), ctypes.c_int64(p.shape[0]),
                                     1 if p.dtype == torch.int64 else 0)
        L.check(self.h, st, "admp_set_pairs")
        self._pairs_key = key if isinstance(pairs, torch.Tensor) else None
        self._pairs_hold = p

    def build_neighbors(self, pos, box):
        pos = _dev(pos, self.device).reshape(self.n_frames, self.n_atoms, 3)
        box = _dev(box, self.device).reshape(self.n_frames, 3, 3)
        self._stream()
        L.check(self.h, self.lib.admp_build_neighbors(self.h, _ptr(pos), _ptr(box)), "admp_build_neighbors")
        self._pairs_key = None

    def get_pairs(self, frame=0):
        n = ctypes.c_int64()
        self._stream()
        L.check(self.h, self.lib.admp_get_pairs(self.h, frame, ctypes.c_void_p(0), 0, ctypes.byref(n)), "admp_get_pairs")
        out = torch.empty((max(n.value, 1), 3), dtype=torch.int32, device=self.device)
        L.check(self.h, self.lib.admp_get_pairs(self.h, frame, _ptr(out), n.value, ctypes.byref(n)), "admp_get_pairs")
        return out[: n.value]

    def compute(self, pos, box, Q_local, pol=None, thole=None, mScales=None, pScales=None, dScales=None,
                U_init=None, solver="pcg", tol=1e-6, maxiter=500, steps=0, want_grad=True, fp32_products=False,
                precond="clusters", stats=True):
        """Energies (F,) and, with want_grad, dE/d{pos, box, Q_local, pol, thole}.
        U_init / returned U are global Cartesian e*A (F,N,3).  precond (PCG):
        "clusters" (exact blocks of up to 4 sites: a pair of mutually nearest coupled
        sites within 2 A and the sites whose nearest coupled neighbour lies in it),
        "pairs" (the pairs alone) or "diagonal".  stats=False: no solver statistics
        (info is None), and the call need not wait for the device at its end."""
        F, N = self.n_frames, self.n_atoms
        dev = self.device
        pos = _dev(pos, dev).reshape(F, N, 3)
        box = _dev(box, dev).reshape(F, 3, 3)
        Q = _dev(Q_local, dev).reshape(N, -1)
        if Q.shape[1] < self.nh:
            raise ValueError(f"Q_local needs {self.nh} harmonic components per site")
        Q = Q[:, : self.nh].contiguous()
        lpol = bool(self.cfg.lpol)
        polt = _dev(pol, dev).reshape(N) if lpol else None
        tht = _dev(thole, dev).reshape(N) if lpol else None
        sc = L.check_scales(mScales, pScales, dScales)
        ns = sc[0].shape[0]
        sc = [s if s is not None else np.ones(ns) for s in sc]
        # a polarisable call writes all of U (the solve's result)
        U = (torch.empty if lpol else torch.zeros)((F, N, 3), dtype=F64, device=dev)
        warm = 0
        if U_init is not None and lpol:
            U.copy_(_dev(U_init, dev).reshape(F, N, 3))
            warm = 1
        opts = L.SolverOpts()
        opts.kind = {"reference": L.SOLVER_REFERENCE, "steps": L.SOLVER_STEPS, "pcg": L.SOLVER_PCG}[solver]
        opts.maxiter, opts.steps, opts.tol, opts.warm_start = int(maxiter), int(steps), float(tol), warm
        opts.fp32_products = int(bool(fp32_products))
        opts.precond = {"clusters": L.ADMP_PRECOND_CLUSTERS, "pairs": L.ADMP_PRECOND_PAIRS,
                        "diagonal": L.ADMP_PRECOND_DIAGONAL}[precond]
        E = torch.empty(F, dtype=F64, device=dev)
        g = {}
        if want_grad:
            g["pos"] = torch.empty((F, N, 3), dtype=F64, device=dev)
            g["box"] = torch.empty((F, 3, 3), dtype=F64, device=dev)
            g["Q"] = torch.empty((F, N, self.nh), dtype=F64, device=dev)
            if lpol:
                g["pol"] = torch.empty((F, N), dtype=F64, device=dev)
                g["thole"] = torch.empty((F, N), dtype=F64, device=dev)
        st_out = L.Stats() if stats else None
        self._stream()
        st = self.lib.admp_compute(
            self.h, _ptr(pos), _ptr(box), _ptr(Q), _ptr(polt), _ptr(tht),
            sc[0].ctypes.data_as(ctypes.c_void_p), sc[1].ctypes.data_as(ctypes.c_void_p),
            sc[2].ctypes.data_as(ctypes.c_void_p), ns, _ptr(U) if lpol else ctypes.c_void_p(0),
            ctypes.byref(opts), int(want_grad), _ptr(E), _ptr(g.get("pos")), _ptr(g.get("box")), _ptr(g.get("Q")),
            _ptr(g.get("pol")), _ptr(g.get("thole")), ctypes.byref(st_out) if stats else None)
        L.check(self.h, st, "admp_compute")
        if not stats:
            return E, g, U, None
        stats = st_out
        info = {
            "n_iter": list(stats.n_iter[:F]), "converged": [bool(x) for x in stats.converged[:F]],
            "residual": list(stats.residual[:F]), "e_real": stats.e_real, "e_recip": stats.e_recip,
            "e_self": stats.e_self, "e_pol": stats.e_pol,
        }
        return E, g, U, info

    def disp_compute(self, pos, box, c_list, mScales, want_grad=True):
        F, N = self.n_frames, self.n_atoms
        dev = self.device
        pos = _dev(pos, dev).reshape(F, N, 3)
        box = _dev(box, dev).reshape(F, 3, 3)
        c = _dev(c_list, dev).reshape(N, 3)
        (ms,) = L.check_scales(mScales)[:1]
        E = torch.empty(F, dtype=F64, device=dev)
        g = {}
        if want_grad:
            g = {"pos": torch.empty((F, N, 3), dtype=F64, device=dev),
                 "box": torch.empty((F, 3, 3), dtype=F64, device=dev),
                 "c": torch.empty((F, N, 3), dtype=F64, device=dev)}
        self._stream()
        st = self.lib.admp_disp_compute(self.h, _ptr(pos), _ptr(box), _ptr(c), ms.ctypes.data_as(ctypes.c_void_p),
                                        len(ms), int(want_grad), _ptr(E), _ptr(g.get("pos")), _ptr(g.get("box")),
                                        _ptr(g.get("c")))
        L.check(self.h, st, "admp_disp_compute")
        return E, g

    def slater_compute(self, pos, box, a, b, mScales, want_grad=True):
        F, N = self.n_frames, self.n_atoms
        dev = self.device
        pos = _dev(pos, dev).reshape(F, N, 3)
        box = _dev(box, dev).reshape(F, 3, 3)
        a = _dev(a, dev).reshape(N)
        b = _dev(b, dev).reshape(N)
        (ms,) = L.check_scales(mScales)[:1]
        E = torch.empty(F, dtype=F64, device=dev)
        g = {}
        if want_grad:
            g = {"pos": torch.empty((F, N, 3), dtype=F64, device=dev),
                 "box": torch.empty((F, 3, 3), dtype=F64, device=dev),
                 "a": torch.empty((F, N), dtype=F64, device=dev), "b": torch.empty((F, N), dtype=F64, device=dev)}
        self._stream()
        st = self.lib.admp_slater_compute(self.h, _ptr(pos), _ptr(box), _ptr(a), _ptr(b),
                                          ms.ctypes.data_as(ctypes.c_void_p), len(ms), int(want_grad), _ptr(E),
                                          _ptr(g.get("pos")), _ptr(g.get("box")), _ptr(g.get("a")), _ptr(g.get("b")))
        L.check(self.h, st, "admp_slater_compute")
        return E, g

    def pairwise_compute(self, kind, pos, box, prm, mScales, want_grad=True):
        """generate_pairwise_interaction kernels (dmff/admp/pairwise.py:98-169) over the
        context's pair lists.  kind: _lib.SR_*; prm (N, n_prm) per-site parameters in
        the kernel's order.  Returns E (F,) and {pos, box, prm} gradients."""
        F, N = self.n_frames, self.n_atoms
        dev = self.device
        npr = L.SR_NPRM[int(kind)]
        pos = _dev(pos, dev).reshape(F, N, 3)
        box = _dev(box, dev).reshape(F, 3, 3)
        prm = _dev(prm, dev).reshape(N, npr)
        (ms,) = L.check_scales(mScales)[:1]
        E = torch.empty(F, dtype=F64, device=dev)
        g = {}
        if want_grad:
            g = {"pos": torch.empty((F, N, 3), dtype=F64, device=dev),
                 "box": torch.empty((F, 3, 3), dtype=F64, device=dev),
                 "prm": torch.empty((F, N, npr), dtype=F64, device=dev)}
        self._stream()
        st = self.lib.admp_pairwise_compute(self.h, int(kind), _ptr(pos), _ptr(box), _ptr(prm), npr,
                                            ms.ctypes.data_as(ctypes.c_void_p), len(ms), int(want_grad), _ptr(E),
                                            _ptr(g.get("pos")), _ptr(g.get("box")), _ptr(g.get("prm")))
        L.check(self.h, st, "admp_pairwise_compute")
        return E, g

    def local_to_global(self, pos, box, Q_local):
        """Q_local_to_global (dmff/admp/pme.py:298-311): per-site local harmonic
        multipoles (N, nh) -> global Cartesian (F, N, 10)."""
        F, N = self.n_frames, self.n_atoms
        dev = self.device
        pos = _dev(pos, dev).reshape(F, N, 3)
        box = _dev(box, dev).reshape(F, 3, 3)
        Q = _dev(Q_local, dev).reshape(N, -1)[:, : self.nh].contiguous()
        out = torch.empty((F, N, 10), dtype=F64, device=dev)
        self._stream()
        L.check(self.h, self.lib.admp_local_to_global(self.h, _ptr(pos), _ptr(box), _ptr(Q), _ptr(out)),
                "admp_local_to_global")
        return out
