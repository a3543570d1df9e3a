"""Drop-in counterpart of dmff/admp/pairwise.py on the MI355X HIP library:
generate_pairwise_interaction (pairwise.py:47-95) and its short-range kernels
(pairwise.py:98-169).

    pair_int = generate_pairwise_interaction(slater_sr_kernel, static_args={})
    E = pair_int(positions, box, pairs, mScales, a_list, b_list)

positions (N,3) and box (3,3) in Angstrom, pairs (Np, >=3) [i, j, nbond] in the
reference format (rows with i >= j, the (N, N) padding included, contribute
nothing), one per-atom array per kernel parameter in the kernel's order.  The
energy is a torch scalar with gradients with respect to positions, box and every
per-atom parameter array.  Each call is four asynchronous kernel launches on the
current stream (admp_pairwise_compute_noalloc): no allocation by the library,
no host synchronisation.
"""
import ctypes

import numpy as np
import torch

from .. import _lib as L
from ._autograd import EnergyWithGrads, as_t

F64 = torch.float64


class PairKernel:
    """Names one of the reference's pair kernels; generate_pairwise_interaction
    maps it to the library's implementation."""

    def __init__(self, name, kind, params):
        self.name, self.kind, self.params = name, kind, tuple(params)

    def __repr__(self):
        return f"<pair kernel {self.name}({', '.join(self.params)})>"


# pairwise.py:156-169, :98-117, :120-128, :131-153
slater_sr_kernel = PairKernel("slater_sr_kernel", L.SR_SLATER, ("a", "b"))
TT_damping_qq_c6_kernel = PairKernel("TT_damping_qq_c6_kernel", L.SR_TT_QQ_C6, ("a", "b", "q", "c"))
TT_damping_qq_kernel = PairKernel("TT_damping_qq_kernel", L.SR_TT_QQ, ("b", "q"))
slater_disp_damping_kernel = PairKernel("slater_disp_damping_kernel", L.SR_SLATER_DAMP, ("b", "c6", "c8", "c10"))


def generate_pairwise_interaction(pair_int_kernel, static_args=None, *, precision="fp64", device="cuda"):
    """pairwise.py:47-95.  precision "mixed" evaluates the pair arithmetic in FP32
    (sums in FP64)."""
    if not isinstance(pair_int_kernel, PairKernel):
        raise TypeError(f"{pair_int_kernel!r} is not one of this module's pair kernels")
    kind = pair_int_kernel.kind
    npr = L.SR_NPRM[kind]
    prec = L.ADMP_MIXED if precision == "mixed" else L.ADMP_FP64
    lib = L.load()

    def pair_int(positions, box, pairs, mScales, *atomic_params):
        if len(atomic_params) != npr:
            raise ValueError(f"{pair_int_kernel.name} takes {npr} per-atom parameters, got {len(atomic_params)}")
        dev = torch.device(device)
        if dev.type != "cuda" and not L.HOST_MEMORY:
            raise ValueError("the pair kernels run on the GPU: device must be a CUDA device")
        pos_in = as_t(positions)
        pos = pos_in.to(device=dev, dtype=F64)
        bx = as_t(box).to(device=dev, dtype=F64)
        prm = torch.stack([as_t(p).to(device=dev, dtype=F64).reshape(-1) for p in atomic_params], dim=1)
        n = pos.shape[0]
        if prm.shape[0] != n:
            raise ValueError(f"per-atom parameters have {prm.shape[0]} rows, positions {n}")
        pr = pairs if isinstance(pairs, torch.Tensor) else torch.as_tensor(np.asarray(pairs))
        pr = pr.to(device=dev)
        if pr.dtype not in (torch.int32, torch.int64):
            pr = pr.to(torch.int64)
        if pr.ndim != 2 or pr.shape[1] < 3:
            raise ValueError("pairs must be (Np, 3) [i, j, nbond]")
        pr = pr[:, :3].contiguous()
        (ms,) = L.check_scales(mScales)[:1]

        def run(want_grad):
            P, B, Q = pos.detach().contiguous(), bx.detach().contiguous(), prm.detach().contiguous()
            E = torch.empty(1, dtype=F64, device=dev)
            gp = torch.empty((n, 3), dtype=F64, device=dev) if want_grad else None
            gb = torch.empty((3, 3), dtype=F64, device=dev) if want_grad else None
            gq = torch.empty((n, npr), dtype=F64, device=dev) if want_grad else None
            vp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)
            st = lib.admp_pairwise_compute_noalloc(
                kind, prec, n, vp(P), vp(B), vp(pr), pr.shape[0], 1 if pr.dtype == torch.int64 else 0, vp(Q), npr,
                ms.ctypes.data_as(ctypes.c_void_p), len(ms), int(want_grad), vp(E), vp(gp), vp(gb), vp(gq),
                ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream if dev.type == "cuda" else 0))
            if st != L.ADMP_OK:
                raise L.ADMPError(f"admp_pairwise_compute_noalloc failed with status {st}: "
                                  f"{lib.admp_last_error(None).decode()}")
            return E.reshape(()), (gp, gb, gq)

        E = EnergyWithGrads.apply(run, pos, bx, prm)
        return E.to(pos_in.device) if E.device != pos_in.device else E

    return pair_int
