"""Observables of LSQuantumED over the GPU state vector.

Mirrors src/spinOperators/correlation.cpp (SzSz, SpSm, SmSp, SiSj),
include/spinOperators/spinOPft.tmpl_func.hpp (Sqz, Sqp) and
include/phys.tmpl_func.hpp (Phys::SquaredSubLattMagAF / STR,
Phys::DynStructFactor).  Each class keeps the reference's
``ExpectVal(env, i, j, state)`` call; underneath, one device pass computes
every pair at once (``correlations``), so an all-pairs sum costs one kernel
instead of N^2 operator builds.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import check, lib, ptr, stream_ptr, write_epoch
from .solver import Solver


def correlations(h, psi, stream=None) -> dict:
    """All-pairs <S^z_i S^z_j>, <S^+_i S^-_j>, <S^-_i S^+_j>, <S_i.S_j> and
    <S^z_i> of the real state psi (CUDA float64, length D)."""
    N = h.basis.nspins
    zz = np.zeros(N * N)
    pm = np.zeros(N * N)
    check(lib().lsqed_obs_szsz(h.hamilt, ptr(psi), zz.ctypes.data_as(C.c_void_p), stream_ptr(stream)))
    check(lib().lsqed_obs_spsm(h.hamilt, ptr(psi), pm.ctypes.data_as(C.c_void_p), stream_ptr(stream)))
    zz = zz.reshape(N, N)
    pm = pm.reshape(N, N)
    # <S^-_i S^+_j>: off-diagonal equal to <S^+_j S^-_i> (= <S^+_i S^-_j> for a
    # real state); diagonal <n_down_i> = P - <n_up_i> (SmSp, correlation.cpp:237-261).
    norm = 4.0 * zz[0, 0]
    mp = pm.T.copy()
    np.fill_diagonal(mp, norm - np.diag(pm))
    sisj = (pm + mp) / 2.0 + zz  # SiSj::ExpectVal, correlation.cpp:276-299
    sz = np.diag(pm) - 0.5 * norm
    return {"SzSz": zz, "SpSm": pm, "SmSp": mp, "SiSj": sisj, "Sz": sz}


class _Corr:
    key = ""

    def __init__(self, env=None, basis=None, h=None):
        self.h = h
        self._cache = None

    def ExpectVal(self, env, spin_i, spin_j, state, h=None):
        """One all-pairs device pass, reused while the state is unchanged:
        same tensor, same torch version and no library call with a tensor
        since (the library may have written into it, e.g. Solver.solution)."""
        hh = h or self.h
        key = (state.data_ptr(), state._version, write_epoch())
        if self._cache is None or self._cache[0] != key or self._cache[1] is not state:
            corr = correlations(hh, state)
            self._cache = ((state.data_ptr(), state._version, write_epoch()), state, corr)
        return float(self._cache[2][self.key][spin_i, spin_j])


class SzSz(_Corr):
    key = "SzSz"


class SpSm(_Corr):
    key = "SpSm"


class SmSp(_Corr):
    key = "SmSp"


class SiSj(_Corr):
    key = "SiSj"


class Phys:
    @staticmethod
    def SquaredSubLattMagAF(env, h, sublat, state) -> float:
        """m^2_AF = 8 sum_{sublattice pairs} <S_i.S_j> / (N (N+4)) (phys.tmpl_func.hpp:5-41)."""
        c = correlations(h, state)["SiSj"]
        acc = 0.0
        for lat in range(sublat.get_size()):
            sl = sublat.get_sl(lat)
            acc += float(c[np.ix_(sl, sl)].sum())
        N = h.basis.nspins
        return acc * 8.0 / float(N) / (float(N) + 4.0)

    @staticmethod
    def SquaredSubLattMagSTR(env, h, sublat, state) -> float:
        """m^2_STR = 4 sum / (N (N+4)) over the four striped sublattices (:45-81)."""
        c = correlations(h, state)["SiSj"]
        acc = 0.0
        for k in range(4):
            sl = sublat.get_sl(k)
            acc += float(c[np.ix_(sl, sl)].sum())
        N = h.basis.nspins
        return acc * 4.0 / float(N) / (float(N) + 4.0)

    @staticmethod
    def dsf_amplitudes(h0, h1, psi0, evecs, stream=None) -> np.ndarray:
        """a[n, i] = <n| S_i |0>: Sqz when h1 is h0's sector, Sqp when h1 has one
        more up spin (phys.tmpl_func.hpp:166-186).  evecs: (nev, D1) CUDA tensor."""
        nev = evecs.shape[0]
        N = h0.basis.nspins
        out = np.zeros(nev * N)
        check(lib().lsqed_obs_dsf(h0.hamilt, h1.hamilt, ptr(psi0), ptr(evecs), nev, out.ctypes.data_as(C.c_void_p),
                                  stream_ptr(stream)))
        return out.reshape(nev, N)

    @staticmethod
    def DynStructFactor(env, data: dict, gs: float, state, dis_iter: int, op: str = "Sqz"):
        """Phys::DynStructFactor<L, SQ> (phys.tmpl_func.hpp:87-215): solves the
        nev lowest states of the target sector, writes ``path/dsf_d<k>.dat``
        in the reference's format and returns (energies, amplitudes)."""
        import torch

        h0, h1 = data["h0"], data["h1"]
        nev = data.get("nev", 10)
        eps = Solver.SolverInit(h1, nev, data.get("ncv", -2), data.get("mpd", -2), tol=data.get("tol"))
        nconv = Solver.solve(env, eps)
        D1 = h1.hamilt_size()
        ev = torch.empty((nev, D1), dtype=torch.float64, device="cuda")
        eigs = []
        for n in range(nev):
            lam, _ = Solver.solution(eps, n, ev[n])
            eigs.append(lam)
        Solver.SolverClean(eps)
        amp = Phys.dsf_amplitudes(h0, h1, state, ev)
        path = data.get("path", ".")
        os.makedirs(path, exist_ok=True)
        fn = os.path.join(path, f"dsf_d{dis_iter}.dat")
        with open(fn, "w") as f:
            f.write("#------------------------------- Parameters -------------------------------------\n")
            f.write("#%7s:   %d  \n" % (" nspins", h0.basis.nspins))
            f.write("#%7s:   %d  \n" % (" nev", nev))
            f.write("#%7s:   %.16f  \n" % (" GS", gs))
            f.write("#%7s:   %.5f  %4s:   %.5f  %4s:   %.5f  %4s:   %.5f\n" % (
                "J1", h0.hamilt_J1(), "D1", h0.hamilt_D1(), "J2", h0.hamilt_J2(), "D2", h0.hamilt_D2()))
            if data.get("disorder"):
                for i, v in enumerate(data["hi"]):
                    f.write("# hi[%4d]: %.16f\n" % (i, v))
            f.write("#--------------------------------------------------------------------------------\n#\n#\n")
            f.write("#----- Energy -----  ------ Factor -----\n")
            for i in range(h0.basis.nspins):
                f.write("# spin: %d\n" % i)
                for n in range(nev):
                    f.write("%.13e  %.13e\n" % (eigs[n] - gs, amp[n, i]))
        return np.array(eigs), amp, nconv
