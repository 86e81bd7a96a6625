"""lsquantumed_amd -- MI355X-native matrix-free Sz-sector Hamiltonian and
Krylov eigensolver with LSQuantumED's host API.

All compute runs in liblsqed.so (hand-written HIP for gfx950) through the C-ABI
of include/lsqed.h; PyTorch only provides device memory and streams.
"""
from ._lib import BUILD_DIAG, BUILD_DISORDER, BUILD_OFFDIAG, LsqedError, lib  # noqa: F401
from .model import AF, Basis, Environment, Hamiltonian, Striped, chain1D, honeycomb2D, square2D  # noqa: F401
from .phys import (Phys, SiSj, SmSp, SpSm, SzSz, correlations, dsf_weights, magnetisation,  # noqa: F401
                   multiplet_weights, q_grid, read_dsf, site_coords)
from .solver import EPS, PETSC_DECIDE, PETSC_DEFAULT, Solver, lanczos_ground, lanczos_step  # noqa: F401
from .tools import RandomDisorder, fields_from_seed  # noqa: F401
from .batch import BatchSolver, disorder_chains  # noqa: F401
from .autograd import ground_energy  # noqa: F401
from .disorder import disorder_average  # noqa: F401
from . import petsc_io  # noqa: F401

__version__ = "0.1.0"
