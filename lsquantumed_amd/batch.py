"""Batched ground states: the realisation loop of examples/main_disorder.cpp
:82-120 (build_disorder, SolverInit(nev=1), solve, solution per
realisation) as one batched Lanczos over many models.

    hams = disorder_chains([18, 20, 22], [64, 64, 128], J1=1.0, J2=0.2, W=2.0, seed=1)
    b = BatchSolver(hams, tol=1e-9)
    e0 = b.solve()                       # numpy array, one E0 per model

The models may differ in N (<= 32), lattice, couplings and fields.  With
keep_vectors=True a second pass builds each model's unit-norm ground state
(`ground_state(k)`) and `observables()` returns the Hellmann-Feynman
derivatives <S^z_i>, bond <S^z S^z> and bond <(S^+S^- + S^-S^+)/2> sums.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr, stream_ptr
from .model import Basis, Environment, Hamiltonian, chain1D
from .tools import fields_from_seed

OBS_STRIDE = 40
OBS_NORM, OBS_ZZ_NN, OBS_FLIP_NN, OBS_ZZ_NNN, OBS_FLIP_NNN = 32, 33, 34, 35, 36


class BatchSolver:
    """Handle of one batch (lsqed_batch_* in include/lsqed.h)."""

    HV_AUTO, HV_CHUNK_TABLE, HV_SPECIALISED = 0, 1, 2

    def __init__(self, hams, maxit: int = 300, tol: float = 1e-9, seed: int = 1, lanes_per_state: int = 0,
                 keep_vectors: bool = False, hv_kernel: int = 0):
        self.hams = list(hams)
        if not self.hams:
            raise ValueError("empty batch")
        n = len(self.hams)
        self._arr = (C.c_void_p * n)(*[h.hamilt.value for h in self.hams])
        o = _lib.BatchOpts(int(maxit), float(tol), int(seed), int(lanes_per_state), int(bool(keep_vectors)),
                           int(hv_kernel))
        hdl = C.c_void_p()
        check(lib().lsqed_batch_create(n, self._arr, C.byref(o), C.byref(hdl)))
        self.handle = hdl
        self.n = n
        self.keep_vectors = bool(keep_vectors)

    @staticmethod
    def workspace_bytes(hams, keep_vectors: bool = False) -> int:
        arr = (C.c_void_p * len(hams))(*[h.hamilt.value for h in hams])
        o = _lib.BatchOpts(0, 0.0, 0, 0, int(bool(keep_vectors)), 0)
        return int(lib().lsqed_batch_workspace_bytes(len(hams), arr, C.byref(o)))

    def destroy(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().lsqed_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def solve(self, stream=None) -> np.ndarray:
        e0 = np.zeros(self.n)
        check(lib().lsqed_batch_solve(self.handle, e0.ctypes.data_as(C.c_void_p), stream_ptr(stream)))
        return e0

    def info(self, k: int):
        """(Lanczos steps, relative residual, converged) of model k."""
        steps, conv = C.c_int(), C.c_int()
        res = C.c_double()
        check(lib().lsqed_batch_info(self.handle, int(k), C.byref(steps), C.byref(res), C.byref(conv)))
        return steps.value, res.value, bool(conv.value)

    def schedule(self) -> dict:
        """Launch shape in use (the autotuner's choice)."""
        k, r, l, g, gr = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        check(lib().lsqed_batch_schedule(self.handle, C.byref(k), C.byref(r), C.byref(l), C.byref(g), C.byref(gr),
                                         C.byref(ms)))
        return {"hv_kernel": {1: "chunk table", 2: "specialised row split"}.get(k.value, "?"),
                "rows_per_task": r.value, "lanes_per_state": l.value, "blocks_per_cu": g.value,
                "graph": bool(gr.value), "step_ms": round(ms.value, 4)}

    def ground_state(self, k: int, out=None, stream=None):
        import torch

        D = self.hams[k].hamilt_size()
        if out is None:
            out = torch.empty(D, dtype=torch.float64, device="cuda")
        if not (out.is_cuda and out.dtype == torch.float64 and out.is_contiguous() and out.numel() == D):
            raise ValueError(f"expected a contiguous float64 CUDA tensor of {D} elements")
        check(lib().lsqed_batch_ground_state(self.handle, int(k), ptr(out), stream_ptr(stream)))
        return out

    def observables(self, stream=None) -> np.ndarray:
        """[n, OBS_STRIDE]: <S^z_i> (i < N), <psi|psi> before normalisation,
        and the nn / nnn bond sums of <S^z S^z> and <(S^+S^- + S^-S^+)/2>."""
        out = np.zeros((self.n, OBS_STRIDE))
        check(lib().lsqed_batch_observables(self.handle, out.ctypes.data_as(C.c_void_p), stream_ptr(stream)))
        return out


def disorder_chains(sizes, counts, J1: float, J2: float, W: float, seed: int = 1, D1=None, D2=None):
    """Periodic chains with random longitudinal fields, as main_disorder.cpp
    builds them (build_off_diag + build_disorder): for the k-th model overall,
    h_i = uniform_real(-W, W) from mt19937_64(seed + k)
    (src/tools/random_disorder.cpp:53-59).  Couplings default to isotropic
    (D1 = J1, D2 = J2)."""
    D1 = J1 if D1 is None else D1
    D2 = J2 if D2 is None else D2
    hams = []
    k = 0
    for N, cnt in zip(sizes, counts):
        env = Environment(N)
        basis = Basis(env, 0)
        lat = chain1D(env)
        for _ in range(cnt):
            h = Hamiltonian(env, basis, lat, J1, D1, J2, D2)
            h.build_diag(env)
            h.build_off_diag(env)
            h.build_disorder(env, fields_from_seed(seed + k, N, -W, W))
            hams.append(h)
            k += 1
    return hams
