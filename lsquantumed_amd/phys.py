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


def magnetisation(h, psi, stream=None) -> np.ndarray:
    """<S^z_i> for every site i of the real state psi (CUDA float64, length D)
    in one device pass: sum_s |psi_s|^2 (bit_i(s) - 1/2), the diagonal of
    SzSz/SpSm reduced to one site (correlation.cpp:86-139)."""
    out = np.zeros(h.basis.nspins)
    check(lib().lsqed_obs_sz(h.hamilt, ptr(psi), out.ctypes.data_as(C.c_void_p), stream_ptr(stream)))
    return out


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


# --- post-processing of the DSF files (postprocessing/dynamicStep0_preprocess.ipynb) ---
def read_dsf(path: str):
    """Reads a dsf_d<k>.dat file written by Phys.DynStructFactor: returns
    (delta_e[nev], amp[nev, N], gs).  The file holds one block of nev rows
    (E_n - E_0, <n|S_i|0>) per spin i (phys.tmpl_func.hpp:190-200); the
    energies are those of the first block (notebook cell 9)."""
    gs = None
    blocks = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#") and line[1:].lstrip().startswith("GS:"):
                gs = float(line.split(":")[1])
            elif line.startswith("# spin:"):
                blocks.append([])
            elif line.strip() and not line.startswith("#"):
                e, a = line.split()
                blocks[-1].append((float(e), float(a)))
    if not blocks or any(len(b) != len(blocks[0]) for b in blocks):
        raise ValueError(f"{path}: not a DynStructFactor file")
    arr = np.array(blocks)  # [N, nev, 2]
    return arr[0, :, 0].copy(), arr[:, :, 1].T.copy(), gs


def site_coords(lx: int, ly: int = 1) -> np.ndarray:
    """r_i = (i // ly, i % ly): the site numbering of square2D/honeycomb2D
    (src/lattice.cpp:94) and of the notebook (cell 3); a chain is (N, 1)."""
    i = np.arange(lx * ly)
    return np.stack([i // ly, i % ly], 1).astype(float)


def q_grid(lx: int, ly: int = 1) -> np.ndarray:
    """The N reciprocal vectors of the notebook (cell 4): q_k = ((k // ly)
    2 pi / lx, (k % ly) 2 pi / ly), k = 0 .. N-1."""
    k = np.arange(lx * ly)
    return np.stack([(k // ly) * 2.0 * np.pi / lx, (k % ly) * 2.0 * np.pi / ly], 1)


def dsf_weights(amp, lx: int, ly: int = 1, q=None) -> np.ndarray:
    """W[q, n] = |sum_i exp(i q.r_i) a[n, i]|^2 / N (notebook cells 12-13)
    for amplitudes amp[n, i] = <n|S_i|0>; q defaults to q_grid(lx, ly)."""
    a = np.asarray(amp, dtype=float)
    N = lx * ly
    if a.ndim != 2 or a.shape[1] != N:
        raise ValueError("amp must be [nev, lx*ly]")
    qs = q_grid(lx, ly) if q is None else np.atleast_2d(np.asarray(q, dtype=float))
    phase = np.exp(1j * (qs @ site_coords(lx, ly).T))  # [nq, N]
    return np.abs(phase @ a.T) ** 2 / N


def multiplet_weights(delta_e, W, tol: float = 1e-8):
    """Sums W[:, n] over degenerate levels (|E_n - E_m| <= tol max(1, |E|)):
    returns (level energies, summed weights [nq, levels]).  Per-state
    amplitudes inside a degenerate level depend on the basis the solver
    picked; the sums do not."""
    e = np.asarray(delta_e, dtype=float)
    W = np.atleast_2d(np.asarray(W, dtype=float))
    levels, cols = [], []
    start = 0
    for n in range(1, len(e) + 1):
        if n == len(e) or abs(e[n] - e[start]) > tol * max(1.0, abs(e[n])):
            levels.append(e[start])
            cols.append(W[:, start:n].sum(axis=1))
            start = n
    return np.array(levels), np.stack(cols, 1)
