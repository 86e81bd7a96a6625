"""ctypes binding of liblsqed.so (the C-ABI declared in include/lsqed.h).

The library is built in-tree by ``lsquantumed_amd/build.py``.  Loading fails
loudly when it is missing: there is no CPU fallback for any operation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsqed.so")

LSQED_OK = 0
ERRORS = {-1: "LSQED_ERR_ARG", -2: "LSQED_ERR_HIP", -3: "LSQED_ERR_OOM", -4: "LSQED_ERR_NOCONV",
          -5: "LSQED_ERR_STATE"}
CHAIN1D, SQUARE2D, HONEYCOMB2D = 0, 1, 2
BUILD_DIAG, BUILD_OFFDIAG, BUILD_DISORDER = 1, 2, 4


class LsqedError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("nspins", C.c_int), ("nup", C.c_int), ("n_nn", C.c_int), ("n_nnn", C.c_int),
                ("nn_i", C.POINTER(C.c_int)), ("nn_j", C.POINTER(C.c_int)),
                ("nnn_i", C.POINTER(C.c_int)), ("nnn_j", C.POINTER(C.c_int)),
                ("J1", C.c_double), ("D1", C.c_double), ("J2", C.c_double), ("D2", C.c_double)]


class EpsOpts(C.Structure):
    _fields_ = [("nev", C.c_int), ("ncv", C.c_int), ("mpd", C.c_int), ("tol", C.c_double),
                ("maxit", C.c_int), ("seed", C.c_uint64), ("multiplicity_rounds", C.c_int)]


class BatchOpts(C.Structure):
    _fields_ = [("maxit", C.c_int), ("tol", C.c_double), ("seed", C.c_uint64), ("lanes_per_state", C.c_int),
                ("keep_vectors", C.c_int), ("hv_kernel", C.c_int)]


_vp, _i32, _i64, _u64, _dbl = C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_double
_P = C.POINTER

# name -> (restype, argtypes); every symbol declared in include/lsqed.h
SIGNATURES = {
    "lsqed_lattice_bonds": (_i32, [_i32, _i32, _i32, _i32, _i32, _vp, _vp, _P(_i32)]),
    "lsqed_sector_dim": (_i64, [_i32, _i32]),
    "lsqed_ham_create": (_i32, [_P(Params), _P(_vp)]),
    "lsqed_ham_destroy": (_i32, [_vp]),
    "lsqed_ham_dim": (_i64, [_vp]),
    "lsqed_ham_nspins": (_i32, [_vp]),
    "lsqed_ham_nup": (_i32, [_vp]),
    "lsqed_ham_set_build": (_i32, [_vp, _i32]),
    "lsqed_ham_get_params": (_i32, [_vp, _P(_dbl), _P(_dbl), _P(_dbl), _P(_dbl), _P(_i32), _P(_i32), _P(_i32)]),
    "lsqed_ham_set_fields": (_i32, [_vp, _vp]),
    "lsqed_hv": (_i32, [_vp, _vp, _vp, _vp]),
    "lsqed_hv_c128": (_i32, [_vp, _vp, _vp, _vp]),
    "lsqed_ham_hv_fingerprint": (_i32, [_vp, _P(_u64)]),
    "lsqed_diag_c128": (_i32, [_vp, _vp, _vp]),
    "lsqed_ham_set_hv_kernel": (_i32, [_vp, _i32]),
    "lsqed_ham_get_hv_kernel": (_i32, [_vp]),
    "lsqed_ham_set_couplings": (_i32, [_vp, _dbl, _dbl, _dbl, _dbl]),
    "lsqed_diag": (_i32, [_vp, _vp, _vp]),
    "lsqed_export_cols": (_i32, [_vp, _i32, _i64, _i64, _i32, _vp, _vp, _vp]),
    "lsqed_unrank": (_i32, [_vp, _i64, _i64, _vp, _vp]),
    "lsqed_rank": (_i32, [_vp, _vp, _i64, _vp, _vp]),
    "lsqed_eps_create": (_i32, [_vp, _P(EpsOpts), _P(_vp)]),
    "lsqed_eps_solve": (_i32, [_vp, _P(_i32), _vp]),
    "lsqed_eps_solution": (_i32, [_vp, _i32, _P(_dbl), _vp, _P(_dbl), _vp]),
    "lsqed_eps_iterations": (_i32, [_vp]),
    "lsqed_eps_workspace_bytes": (_i64, [_vp]),
    "lsqed_eps_destroy": (_i32, [_vp]),
    "lsqed_lanczos_ground": (_i32, [_vp, _i32, _dbl, _u64, _vp, _P(_dbl), _P(_i32), _vp]),
    "lsqed_lanczos_step": (_i32, [_vp, _vp, _dbl, _vp, _dbl, _dbl, _vp, _P(_dbl), _P(_dbl), _vp]),
    "lsqed_lanczos_step_noalloc": (_i32, [_vp, _vp, _dbl, _vp, _dbl, _dbl, _vp, _P(_dbl), _P(_dbl), _vp]),
    "lsqed_obs_szsz": (_i32, [_vp, _vp, _vp, _vp]),
    "lsqed_obs_spsm": (_i32, [_vp, _vp, _vp, _vp]),
    "lsqed_obs_sz": (_i32, [_vp, _vp, _vp, _vp]),
    "lsqed_obs_dsf": (_i32, [_vp, _vp, _vp, _vp, _i32, _vp, _vp]),
    "lsqed_batch_create": (_i32, [_i32, _P(_vp), _P(BatchOpts), _P(_vp)]),
    "lsqed_batch_solve": (_i32, [_vp, _vp, _vp]),
    "lsqed_batch_info": (_i32, [_vp, _i32, _P(_i32), _P(_dbl), _P(_i32)]),
    "lsqed_batch_schedule": (_i32, [_vp, _P(_i32), _P(_i32), _P(_i32), _P(_i32), _P(_i32), _P(_dbl)]),
    "lsqed_batch_ground_state": (_i32, [_vp, _i32, _vp, _vp]),
    "lsqed_batch_observables": (_i32, [_vp, _vp, _vp]),
    "lsqed_batch_destroy": (_i32, [_vp]),
    "lsqed_batch_workspace_bytes": (_i64, [_i32, _P(_vp), _P(BatchOpts)]),
    "lsqed_stream_probe": (_i32, [_i32, _vp, _vp, _vp, _i64, _vp, _vp]),
    "lsqed_rd_seeds": (_i32, [_i32, _vp]),
    "lsqed_rd_fields": (_i32, [_i64, _i32, _dbl, _dbl, _vp]),
    "lsqed_last_error": (C.c_char_p, []),
    "lsqed_version": (C.c_char_p, []),
}

_LIB = None


def lib():
    """Load liblsqed.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m lsquantumed_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int) -> None:
    if rc != LSQED_OK:
        raise LsqedError(rc, lib().lsqed_last_error().decode())


def stream_ptr(stream=None):
    """hipStream_t of a torch stream (None = torch's current stream)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


_EPOCH = [0]


def ptr(t) -> C.c_void_p:
    """Device address of a tensor handed to the library.  Each hand-over
    advances the write epoch: the library may write there, which torch's
    version counter does not see."""
    _EPOCH[0] += 1
    return C.c_void_p(t.data_ptr())


def write_epoch() -> int:
    """Number of tensor hand-overs so far (a cache of results computed from a
    tensor is valid while this and the tensor's version are unchanged)."""
    return _EPOCH[0]
