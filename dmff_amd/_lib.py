"""ctypes binding of libadmp_hip.so (the C ABI in include/admp_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (hipcc, gfx950).
There is no CPU fallback: if the library is missing, importing the compute
path raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libadmp_hip.so")

# True only for a build of the library whose "device" memory is host memory (the
# CPU emulation of the HIP runtime used by the test suite); the HIP library
# itself needs device pointers, and entry points that take torch tensors
# refuse CPU tensors while this is False.
HOST_MEMORY = False

ADMP_OK = 0
ADMP_WARN_NOT_CONVERGED = 5
ADMP_FP64, ADMP_MIXED = 0, 1
ADMP_PRECOND_CLUSTERS, ADMP_PRECOND_DIAGONAL, ADMP_PRECOND_PAIRS = 0, 1, 2
SOLVER_REFERENCE, SOLVER_STEPS, SOLVER_PCG = 0, 1, 2

SYMBOLS = [
    "admp_version", "admp_ctx_create", "admp_ctx_destroy", "admp_last_error", "admp_set_stream",
    "admp_set_topology", "admp_set_pairs", "admp_build_neighbors", "admp_get_pairs", "admp_compute",
    "admp_disp_compute", "admp_slater_compute", "admp_profile", "admp_profile_get", "admp_memory_bytes",
    "admp_type_reduce", "admp_pairwise_compute", "admp_pairwise_compute_noalloc", "admp_local_to_global",
    "admp_esp", "admp_add_forces_fixed", "admp_set_stream_mode",
]

# admp_sr_kind and each kind's per-site parameter count
SR_SLATER, SR_TT_QQ_C6, SR_TT_QQ, SR_SLATER_DAMP = 1, 2, 3, 4
SR_NPRM = {SR_SLATER: 2, SR_TT_QQ_C6: 4, SR_TT_QQ: 2, SR_SLATER_DAMP: 4}


class Config(ctypes.Structure):
    _fields_ = [
        ("n_atoms", ctypes.c_int32), ("n_frames", ctypes.c_int32), ("lmax", ctypes.c_int32),
        ("lpol", ctypes.c_int32), ("lpme", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("kappa", ctypes.c_double), ("K1", ctypes.c_int32), ("K2", ctypes.c_int32), ("K3", ctypes.c_int32),
        ("pme_order", ctypes.c_int32), ("rc", ctypes.c_double), ("max_neighbors", ctypes.c_int32),
        ("disp_kappa", ctypes.c_double), ("dK1", ctypes.c_int32), ("dK2", ctypes.c_int32), ("dK3", ctypes.c_int32),
        ("disp_pmax", ctypes.c_int32),
    ]


class SolverOpts(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("maxiter", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("tol", ctypes.c_double), ("warm_start", ctypes.c_int32), ("fp32_products", ctypes.c_int32),
                ("precond", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("n_iter", ctypes.c_int32 * 64), ("converged", ctypes.c_int32 * 64),
                ("residual", ctypes.c_double * 64), ("e_real", ctypes.c_double), ("e_recip", ctypes.c_double),
                ("e_self", ctypes.c_double), ("e_pol", ctypes.c_double)]


_lib = None


class ADMPError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ADMPError(f"{LIB_PATH} not found: build it with __graft_entry__.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64, d = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    lib.admp_version.restype = ctypes.c_char_p
    lib.admp_last_error.restype = ctypes.c_char_p
    lib.admp_last_error.argtypes = [P]
    lib.admp_ctx_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(P)]
    lib.admp_ctx_destroy.argtypes = [P]
    lib.admp_set_stream.argtypes = [P, P]
    lib.admp_set_stream_mode.argtypes = [P, i32]
    lib.admp_set_topology.argtypes = [P, P, P, i64, P, P, P]
    lib.admp_set_pairs.argtypes = [P, i32, P, i64, i32]
    lib.admp_build_neighbors.argtypes = [P, P, P]
    lib.admp_get_pairs.argtypes = [P, i32, P, i64, ctypes.POINTER(i64)]
    lib.admp_compute.argtypes = [P, P, P, P, P, P, P, P, P, i32, P, ctypes.POINTER(SolverOpts), i32,
                                 P, P, P, P, P, P, ctypes.POINTER(Stats)]
    lib.admp_disp_compute.argtypes = [P, P, P, P, P, i32, i32, P, P, P, P]
    lib.admp_slater_compute.argtypes = [P, P, P, P, P, P, i32, i32, P, P, P, P, P]
    lib.admp_profile.argtypes = [P, i32, i32]
    lib.admp_profile_get.argtypes = [P, i32, ctypes.POINTER(d), ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_char_p)]
    lib.admp_memory_bytes.argtypes = [P, ctypes.POINTER(i64)]
    lib.admp_type_reduce.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P, ctypes.c_int32, P, P]
    lib.admp_pairwise_compute.argtypes = [P, i32, P, P, P, i32, P, i32, i32, P, P, P, P]
    lib.admp_local_to_global.argtypes = [P, P, P, P, P]
    lib.admp_esp.argtypes = [i32, P, P, P, i32, P, P, P]
    lib.admp_add_forces_fixed.argtypes = [P, d, P, P, i32, i32, P]
    lib.admp_pairwise_compute_noalloc.argtypes = [i32, i32, i32, P, P, P, i64, i32, P, i32, P, i32, i32, P, P, P,
                                                  P, P]
    for name in SYMBOLS:
        getattr(lib, name).restype = ctypes.c_int if name not in ("admp_version", "admp_last_error") else \
            getattr(lib, name).restype
    _lib = lib
    return lib


def check_scales(mScales, pScales=None, dScales=None):
    """Topological scale vectors as the library takes them: equal lengths in 1..8
    (index nbond-1, with nbond = 0 wrapping to the last entry as pme.py:1111-1125
    does); raises ValueError otherwise."""
    import numpy as np
    out = []
    n = None
    for name, s in (("mScales", mScales), ("pScales", pScales), ("dScales", dScales)):
        if s is None:
            out.append(None)
            continue
        a = np.ascontiguousarray(np.asarray(s.detach().cpu() if hasattr(s, "detach") else s, dtype=np.float64)).reshape(-1)
        if n is None:
            n = a.shape[0]
        if a.shape[0] != n:
            raise ValueError(f"{name} has {a.shape[0]} entries, mScales has {n}: scale vectors must have equal lengths")
        out.append(a)
    if n is None or not 1 <= n <= 8:
        raise ValueError(f"scale vectors need 1..8 entries, got {n}")
    return out


def check(ctx, status, what):
    if status not in (ADMP_OK, ADMP_WARN_NOT_CONVERGED):
        msg = _lib.admp_last_error(ctx).decode() if ctx else ""
        raise ADMPError(f"{what} failed with status {status}: {msg}")
    return status
