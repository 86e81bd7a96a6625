"""Solver namespace of LSQuantumED (include/solver.hpp, src/solver.cpp) over
the device-resident Krylov-Schur of liblsqed.so.

    eps = Solver.SolverInit(h, nev, ncv, mpd)      # src/solver.cpp:7-24
    nconv = Solver.solve(env, eps)                 # :38-52
    lam, err = Solver.solution(eps, i, vec)        # :58-68 (vec: CUDA tensor or None)
    Solver.SolverClean(eps)                        # :28-32

Extra knobs that the reference takes from the PETSc options database
(``-eps_tol``, ``-eps_max_it``; EPSSetFromOptions, src/solver.cpp:21) are
keyword arguments here.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import check, lib, ptr, stream_ptr

PETSC_DEFAULT = -2
PETSC_DECIDE = -1


class EPS:
    """Handle of one eigensolve (the reference's `EPS` object)."""

    def __init__(self, h, nev, ncv=PETSC_DEFAULT, mpd=PETSC_DEFAULT, tol=None, maxit=None, seed=1,
                 multiplicity_rounds=0):
        self.h = h
        o = _lib.EpsOpts(int(nev), int(ncv) if ncv and ncv > 0 else 0, int(mpd) if mpd and mpd > 0 else 0,
                         float(tol) if tol else 0.0, int(maxit) if maxit else 0, int(seed), int(multiplicity_rounds))
        e = C.c_void_p()
        check(lib().lsqed_eps_create(h.hamilt, C.byref(o), C.byref(e)))
        self.handle = e
        self.nconv = 0

    def destroy(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().lsqed_eps_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def iterations(self) -> int:
        return lib().lsqed_eps_iterations(self.handle)

    def workspace_bytes(self) -> int:
        """Device bytes the solver holds (basis columns and scratch)."""
        return int(lib().lsqed_eps_workspace_bytes(self.handle))


class Solver:
    @staticmethod
    def SolverInit(h, nev, ncv=PETSC_DEFAULT, mpd=PETSC_DEFAULT, tol=None, maxit=None, seed=1,
                   multiplicity_rounds=0) -> EPS:
        """multiplicity_rounds > 0: after convergence, look for copies of
        degenerate levels that the start vector missed (at most that many
        rounds; include/lsqed.h); 0 = SLEPc's stopping rule."""
        return EPS(h, nev, ncv, mpd, tol, maxit, seed, multiplicity_rounds)

    @staticmethod
    def solve(env, eps: EPS, stream=None) -> int:
        n = C.c_int()
        check(lib().lsqed_eps_solve(eps.handle, C.byref(n), stream_ptr(stream)))
        eps.nconv = n.value
        return n.value

    @staticmethod
    def solution(eps: EPS, index: int, vec=None, stream=None):
        lam = C.c_double()
        err = C.c_double()
        vp = ptr(vec) if vec is not None else None
        check(lib().lsqed_eps_solution(eps.handle, int(index), C.byref(lam), vp, C.byref(err), stream_ptr(stream)))
        return lam.value, err.value

    @staticmethod
    def SolverClean(eps: EPS) -> int:
        eps.destroy()
        return 0


def lanczos_ground(h, maxit=2000, tol=1e-13, seed=1, work=None, stream=None):
    """Ground energy from the three-vector Lanczos (no eigenvector).  `work`:
    optional float64 CUDA tensor of 3*D elements."""
    import torch

    D = h.hamilt_size()
    if work is None:
        work = torch.empty(3 * D, dtype=torch.float64, device="cuda")
    e0 = C.c_double()
    steps = C.c_int()
    check(lib().lsqed_lanczos_ground(h.hamilt, int(maxit), float(tol), int(seed), ptr(work), C.byref(e0),
                                     C.byref(steps), stream_ptr(stream)))
    return e0.value, steps.value


def lanczos_step(h, q, sq, p, sp, beta, y, stream=None):
    """One three-term Lanczos step (see include/lsqed.h); returns (alpha, beta')."""
    a = C.c_double()
    b = C.c_double()
    check(lib().lsqed_lanczos_step(h.hamilt, ptr(q), float(sq), ptr(p), float(sp), float(beta), ptr(y), C.byref(a),
                                   C.byref(b), stream_ptr(stream)))
    return a.value, b.value


def lanczos_ground_residual(h, tol=1e-9, maxit=600, seed=7, check_every=5, progress=None, stream=None):
    """Ground energy by the three-vector Lanczos (q, p, y on the device, the
    tridiagonal matrix on the host; no reorthogonalisation, no eigenvector),
    stopped when the lowest Ritz pair (theta, s) has the relative residual
    |beta_m s_m| / |theta| < tol -- the EPS_CONV_REL test of the Krylov-Schur
    solver (src/solver.cpp:38-52) -- checked every `check_every` steps.  For
    sectors whose Krylov-Schur basis does not fit (N = 36: 72.6 GB per
    vector).  Start vector uniform in [-0.5, 0.5) from torch's device
    generator seeded with `seed`.  progress(step, theta, resid, seconds) is
    called at every check.  Returns a dict with E0, steps, the residual and
    the wall time of the iteration."""
    import time

    import numpy as np
    import scipy.linalg as sl
    import torch

    D = h.hamilt_size()
    g = torch.Generator(device="cuda").manual_seed(int(seed))
    q = torch.rand(D, dtype=torch.float64, device="cuda", generator=g)
    q -= 0.5
    p = torch.zeros(D, dtype=torch.float64, device="cuda")
    y = torch.empty(D, dtype=torch.float64, device="cuda")
    ss = sum(float(torch.dot(q[i:i + 2**30], q[i:i + 2**30]).item()) for i in range(0, D, 2**30))
    sq, sp, beta = 1.0 / ss ** 0.5, 0.0, 0.0
    al, be = [], []
    theta, res = None, None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for m in range(int(maxit)):
        a, b = lanczos_step(h, q, sq, p, sp, beta, y, stream)
        al.append(a)
        be.append(b)
        exhausted = len(al) >= D or not b > 1e-14 * abs(a)  # invariant subspace reached
        if not exhausted:
            q, p = p, q
            sp, sq, beta = sq, 1.0 / b, b
        if (m + 1) % check_every == 0 or m == maxit - 1 or exhausted:
            w, v = sl.eigh_tridiagonal(np.array(al), np.array(be[:-1]), select="i", select_range=(0, 0))
            theta, res = float(w[0]), abs(be[-1] * v[-1, 0]) / max(abs(float(w[0])), 1e-300)
            if progress:
                progress(m + 1, theta, res, time.perf_counter() - t0)
            if res < tol or exhausted:
                break
    wall = time.perf_counter() - t0
    del q, p, y
    return {"E0": theta, "steps": len(al), "relative_residual": res, "wall_s": wall}
