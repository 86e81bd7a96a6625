"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/lsqed.h declares, and its host-only routines (bond lists, sector
size, RandomDisorder) agree with the oracle / the reference's literals.
No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LT = {"chain": 0, "square": 1, "honeycomb": 2}


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "lsqed.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(lsqed_\w+)\s*\(", txt, flags=re.M)))


@pytest.fixture(scope="module")
def L():
    from lsquantumed_amd import _lib
    return _lib.lib()


def test_exports_every_declared_symbol(L):
    syms = declared_symbols()
    assert len(syms) >= 29
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_bindings_cover_header():
    from lsquantumed_amd import _lib
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_version(L):
    assert b"gfx950" in L.lsqed_version()


@pytest.mark.parametrize("lat,N,lx,ly", [("chain", 6, 0, 0), ("chain", 16, 0, 0), ("chain", 2, 0, 0),
                                         ("square", 8, 2, 4), ("square", 12, 3, 4), ("square", 24, 4, 6),
                                         ("square", 32, 4, 8), ("square", 6, 6, 1), ("honeycomb", 12, 6, 2),
                                         ("honeycomb", 8, 2, 4), ("square", 36, 6, 6)])
def test_bonds_match_oracle(O, lat, N, lx, ly):
    from lsquantumed_amd import model
    cls = {"chain": model.chain1D, "square": model.square2D, "honeycomb": model.honeycomb2D}[lat]
    L = cls(N) if lat == "chain" else cls(N, lx, ly)
    for kind, got in ((0, L.nn_bonds), (1, L.nnn_bonds)):
        bi, bj = O.bonds(LT[lat], N, lx, ly, kind)
        assert list(got[0]) == list(bi) and list(got[1]) == list(bj)


def test_lattice_errors(L):
    """The reference's exceptions and messages (src/environment.cpp:31-35,
    src/basis.cpp:66-67, src/lattice.cpp:67-68, :150-154); the C-ABI itself
    rejects the same inputs with LSQED_ERR_ARG."""
    from lsquantumed_amd import model
    with pytest.raises(ValueError, match="^The product between the number of spins"):
        model.square2D(12, 3, 5)
    with pytest.raises(ValueError, match="^The number of spins in each direction has to be even.$"):
        model.honeycomb2D(12, 3, 4)
    with pytest.raises(ValueError, match="^The total number of spins has to be EVEN.$"):
        model.Environment(7)
    with pytest.raises(ValueError, match="^The total number of spins has to be EVEN.$"):
        model.Environment(-3)
    with pytest.raises(ValueError, match="^The total number of spins has to be POSITIVE.$"):
        model.Environment(0)
    with pytest.raises(ValueError, match="^The required constant magnetization has to be <= nspins/2.$"):
        model.Basis(8, 5)
    n = C.c_int()
    assert L.lsqed_lattice_bonds(LT["square"], 12, 3, 5, 0, None, None, C.byref(n)) == -1
    assert L.lsqed_lattice_bonds(LT["honeycomb"], 12, 3, 4, 0, None, None, C.byref(n)) == -1


def test_sector_dim(L):
    import math
    for N in (6, 16, 24, 32, 36, 62):
        for nup in (0, 1, N // 2, N):
            assert L.lsqed_sector_dim(N, nup) == math.comb(N, nup)


def test_random_disorder_literals(literals):
    from lsquantumed_amd import RandomDisorder
    r = literals["random"]
    rd = RandomDisorder(r["N"], True, r["min"], r["max"], r["rep"])
    assert list(rd.srand_seq) == r["seeds"]
    hi = rd.hi_init(r["hi_index"])
    assert np.max(np.abs(hi - np.array(r["hi"]))) <= r["tol"]


def test_random_fields_match_oracle(O):
    from lsquantumed_amd import fields_from_seed
    for seed in (1, 2, 77, 1000):
        assert np.array_equal(fields_from_seed(seed, 22, -2.0, 2.0), O.rd_hi(seed, 22, -2.0, 2.0))


def test_petsc_format_helpers(O):
    """The product's PETSc %g and row ownership equal the oracle's restatement
    of the reference (petsc_g; Basis split, src/basis.cpp:78-80)."""
    from lsquantumed_amd import petsc_io
    vals = [0.0, 0.5, -1.5, 1.0, 2.0, -3.0, 9.83333333333, 1e-7, 123456789.0, 1.0 / 3.0, -0.25, 1e20]
    for v in vals:
        assert petsc_io.petsc_g(v) == O.petsc_g(v), v
    for size in (1, 6, 7, 20, 495, 1001):
        for P in (1, 2, 3, 5, 8):
            for r in range(P):
                assert petsc_io.ownership(size, P, r) == O.split(size, P, r), (size, P, r)
    txt = petsc_io.vec_view([1.5], 1)
    assert txt == "Vec Object: 1 MPI processes\n  type: seq\n1.5\n"
