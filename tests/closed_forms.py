"""Spectra of sectors known in closed form (checkers for the oracle and the
GPU product alike)."""
import numpy as np


def polarised_energy(lat, N, J1, D1, J2, D2, hi=None, up=True):
    """The one state of the fully polarised sector (all up, or all down):
    every bond parallel, so H = D1/4 n_nn + D2/4 n_nnn, plus 1/2 sum h (all
    up) or -1/2 sum h (all down); n = N bonds of each kind on the periodic
    chain, 2 N on the square lattice."""
    nb = N if lat == "chain" else 2 * N
    e = D1 * 0.25 * nb + D2 * 0.25 * nb
    if hi is not None:
        e += (0.5 if up else -0.5) * float(np.sum(hi))
    return e


def magnon_levels(lat, N, lx, ly, J1, D1, J2, D2):
    """Single-magnon spectrum (one up spin, or one down spin, h = 0) of the
    periodic chain (N >= 5) or Lx x Ly square lattice (Lx, Ly >= 3): the
    flipped spin sees antiparallel bonds only on its own 2 (chain) or 4
    (square) bonds of each kind and hops with J/2 along each, so
    E(k) = D1/4 (n_nn - 2 z) + D2/4 (n_nnn - 2 z) + J1 g1(k) + J2 g2(k),
    z = 2 or 4 bonds per site, g1 = cos k, g2 = cos 2k (chain);
    g1 = cos kx + cos ky, g2 = cos(kx + ky) + cos(kx - ky) (square)."""
    if lat == "chain":
        k = 2 * np.pi * np.arange(N) / N
        return np.sort((D1 + D2) / 4 * (N - 4) + J1 * np.cos(k) + J2 * np.cos(2 * k))
    kx, ky = np.meshgrid(2 * np.pi * np.arange(lx) / lx, 2 * np.pi * np.arange(ly) / ly, indexing="ij")
    e = ((D1 + D2) / 4 * (2 * N - 8) + J1 * (np.cos(kx) + np.cos(ky))
         + J2 * (np.cos(kx + ky) + np.cos(kx - ky)))
    return np.sort(e.ravel())


# (lattice, N, Lx, Ly): periodic chains and square lattices up to N = 40
CLOSED = [("chain", 16, 0, 0), ("chain", 40, 0, 0), ("square", 24, 4, 6), ("square", 36, 6, 6),
          ("square", 40, 4, 10)]
