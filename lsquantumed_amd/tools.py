"""Tools::RandomDisorder (include/tools/random_disorder.hpp,
src/tools/random_disorder.cpp) over the library's libstdc++ engines."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


class RandomDisorder:
    """RandomDisorder(nspins, reprod, min, max, rep).  Only the reproducible
    seed list (mt19937_64(1) -> uniform_int(0, 1000)) is supported; the
    reference's non-reproducible branch draws from std::random_device."""

    def __init__(self, nspins: int, reprod: bool, lo: float, hi: float, rep: int):
        if not reprod:
            raise ValueError("only reproducible seeds are supported")
        self.nspins, self.min, self.max, self.rep = int(nspins), float(lo), float(hi), int(rep)
        self.srand_seq = np.zeros(self.rep, np.int64)
        check(lib().lsqed_rd_seeds(self.rep, self.srand_seq.ctypes.data_as(C.c_void_p)))
        self.hi = np.zeros(self.nspins)

    def hi_init(self, dis_index: int) -> np.ndarray:
        check(lib().lsqed_rd_fields(int(self.srand_seq[dis_index]), self.nspins, self.min, self.max,
                                    self.hi.ctypes.data_as(C.c_void_p)))
        return self.hi


def fields_from_seed(seed: int, nspins: int, lo: float, hi: float) -> np.ndarray:
    """h_i = uniform_real(lo, hi) from mt19937_64(seed) (hi_init with an explicit seed)."""
    out = np.zeros(nspins)
    check(lib().lsqed_rd_fields(int(seed), int(nspins), float(lo), float(hi), out.ctypes.data_as(C.c_void_p)))
    return out
