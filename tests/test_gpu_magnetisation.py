"""GPU m^2_AF / m^2_STR (Phys::SquaredSubLattMagAF/STR,
include/phys.tmpl_func.hpp:5-81) on states whose <S_i.S_j> are known exactly:
the Neel and fully polarised states (m^2_AF = 1 by the normalisation),
stripes, random basis states, and products of singlets inside one AF
sublattice (m^2_AF = 0, which needs the S^+S^- half of SiSj) or across the
two (6/(N+4)).  Expected values come from tests/closed_forms.py, not from the
oracle's m^2 functions."""
import sys

import numpy as np
import pytest

from closed_forms import m2_closed, neel_bits, product_block_sum, singlet_block_sum, singlet_product
from test_gpu_hamiltonian import make

pytestmark = pytest.mark.gpu

CASES = [("chain", 12, 0, 0), ("square", 16, 4, 4), ("square", 24, 4, 6), ("honeycomb", 16, 4, 4),
         ("square", 12, 2, 6)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _state(torch, D, amps):
    psi = np.zeros(D)
    for r, v in amps.items():
        psi[r] = v
    return torch.tensor(psi, dtype=torch.float64, device="cuda")


def _lattice(Q, env, lat, lx, ly):
    return {"chain": lambda: Q.chain1D(env), "square": lambda: Q.square2D(env, lx, ly),
            "honeycomb": lambda: Q.honeycomb2D(env, lx, ly)}[lat]()


@pytest.mark.parametrize("lat,N,lx,ly", CASES)
def test_m2_closed_forms(O, lat, N, lx, ly):
    import torch

    import lsquantumed_amd as Q

    say("m2 closed forms", lat, N, lx, ly)
    env = Q.Environment(N)
    L = _lattice(Q, env, lat, lx, ly)
    af = Q.AF(L)
    st = Q.Striped(L) if lat == "square" else None
    sl_af = [list(af.get_sl(k)) for k in range(2)]
    sl_st = [list(st.get_sl(k)) for k in range(4)] if st else None
    h = make(lat, N, 0, 1.0, 1.0, 0.5, 0.5, lx, ly)
    D = h.hamilt_size()
    rank = lambda s: O.rank_ref(s, N)  # noqa: E731

    def m2(psi):
        a = Q.Phys.SquaredSubLattMagAF(env, h, af, psi)
        s = Q.Phys.SquaredSubLattMagSTR(env, h, st, psi) if st else None
        return a, s

    # Neel states: exactly 1
    for bits in (neel_bits(sl_af[0]), neel_bits(sl_af[1])):
        a, _ = m2(_state(torch, D, {rank(bits): 1.0}))
        assert abs(a - 1.0) <= 1e-14, (bits, a)
    # random basis states (and normalisation-independent sign: -psi)
    rng = np.random.default_rng(N)
    for t in range(5):
        bits = int(sum(1 << int(i) for i in rng.choice(N, size=N // 2, replace=False)))
        a, s = m2(_state(torch, D, {rank(bits): -1.0 if t % 2 else 1.0}))
        assert abs(a - m2_closed([product_block_sum(bits, sl) for sl in sl_af], N)) <= 1e-14
        if st:
            assert abs(s - m2_closed([product_block_sum(bits, sl) for sl in sl_st], N, orientations=2)) <= 1e-14
    # stripes (square lattices with an even number of rows)
    if st and ly % 2 == 0:
        bits = sum(1 << (ix * ly + iy) for ix in range(lx) for iy in range(0, ly, 2))
        a, s = m2(_state(torch, D, {rank(bits): 1.0}))
        assert abs(s - (N + 8) / (2 * (N + 4))) <= 1e-14 and abs(a - 4 / (N + 4)) <= 1e-14, (a, s)
    # singlet products
    A, B = sl_af
    same = [(A[k], A[k + 1]) for k in range(0, len(A), 2)] + [(B[k], B[k + 1]) for k in range(0, len(B), 2)]
    cross = list(zip(A, B))
    for pairs, want in ((same, 0.0), (cross, 6.0 / (N + 4))):
        psi = _state(torch, D, singlet_product(pairs, rank))
        a, s = m2(psi)
        assert abs(a - want) <= 1e-13, (pairs, a, want)
        if st:
            assert abs(s - m2_closed([singlet_block_sum(pairs, sl) for sl in sl_st], N, orientations=2)) <= 1e-13


def test_m2_polarised_sector():
    """The one state of the fully polarised sector (Basis total_mag = N/2):
    every pair in a sublattice is aligned, m^2_AF = 1."""
    import torch

    import lsquantumed_amd as Q

    for lat, N, lx, ly in CASES[:3]:
        say("m2 polarised", lat, N)
        env = Q.Environment(N)
        L = _lattice(Q, env, lat, lx, ly)
        h = make(lat, N, N // 2, 1.0, 1.0, 0.5, 0.5, lx, ly)
        assert h.hamilt_size() == 1
        psi = torch.ones(1, dtype=torch.float64, device="cuda")
        assert abs(Q.Phys.SquaredSubLattMagAF(env, h, Q.AF(L), psi) - 1.0) <= 1e-14
