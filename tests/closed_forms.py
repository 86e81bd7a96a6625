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


# --- sublattice magnetisation of states whose correlators are known exactly --
# For a product state |s> (one basis state): <S_i.S_j> = 3/4 for i = j
# (S(S+1) with S = 1/2) and s_i s_j with s = +-1/2 otherwise (the S^+S^-
# terms move the state off itself).  Summed over the pairs of one sublattice
# with u up and d down spins: 3n/4 + ((u-d)^2 - n)/4 = n/2 + (u-d)^2/4.
# For a product of two-site singlets: <S_a.S_b> = -3/4 inside a singlet,
# 0 between different singlets (each has <S^x,y,z> = 0), 3/4 on the diagonal.
#
# Normalisation: m^2_AF is normalised so that the Neel state gives exactly 1
# (two sublattices of N/2 aligned spins: 2 (N/4 + N^2/16) = N (N+4)/8, whence
# the 8 / (N (N+4)) of Phys::SquaredSubLattMagAF); m^2_STR is the average of
# the two stripe orientations' order parameters, each normalised like the AF
# one, whence 8/2 = 4 over the four striped sublattices.


def product_block_sum(bits: int, sl) -> float:
    """sum_{i,j in sl} <s|S_i.S_j|s> for the basis state s (bit i = spin i up)."""
    u = sum((bits >> i) & 1 for i in sl)
    d = len(sl) - u
    return len(sl) / 2.0 + (u - d) ** 2 / 4.0


def singlet_block_sum(pairs, sl) -> float:
    """sum_{i,j in sl} <S_i.S_j> for the product of singlets on `pairs`."""
    s = set(sl)
    acc = 0.75 * len(sl)
    for a, b in pairs:
        if a in s and b in s:
            acc += 2 * (-0.75)
    return acc


def m2_closed(blocks, N: int, orientations: int = 1) -> float:
    """Sublattice sums normalised by the Neel state's sum N (N+4)/8, averaged
    over `orientations` (1: AF; 2: the x and y stripes of Striped)."""
    return sum(blocks) / orientations / (N * (N + 4) / 8.0)


def neel_bits(sublattice_a) -> int:
    return sum(1 << i for i in sublattice_a)


def singlet_product(pairs, rank):
    """Amplitudes of prod_{(a,b)} (|up_a dn_b> - |dn_a up_b>)/sqrt(2) as a
    dict rank -> amplitude (rank: state -> basis index)."""
    import itertools

    amp = 2.0 ** (-len(pairs) / 2.0)
    out = {}
    for choice in itertools.product((0, 1), repeat=len(pairs)):
        s, sign = 0, 1.0
        for (a, b), c in zip(pairs, choice):
            s |= 1 << (b if c else a)
            if c:
                sign = -sign
        out[rank(s)] = sign * amp
    return out
