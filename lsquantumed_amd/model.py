"""Host-side mirror of LSQuantumED's model classes over the HIP C-ABI.

Same names, argument meaning and error behaviour as the reference:
``Environment`` (include/environment.hpp), ``Basis`` (include/basis.hpp),
``chain1D`` / ``square2D`` / ``honeycomb2D`` (include/lattice.hpp),
``AF`` / ``Striped`` (include/sublattice.hpp), ``Hamiltonian``
(include/hamiltonian.hpp).  What differs by design: ``Basis`` stores no
``int_basis`` vector (states are unranked on the device on demand) and
``Hamiltonian`` holds no assembled matrix -- ``hamilt`` is the matrix-free
operator handle whose product runs on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr, stream_ptr


class Environment:
    """Environment(nspins): validates N like src/environment.cpp:31-35.  MPI
    and SLEPc initialisation have no counterpart on one GPU."""

    def __init__(self, nspins: int):
        # odd counts first, then non-positive ones, with the reference's messages
        if nspins % 2 != 0:
            raise ValueError("The total number of spins has to be EVEN.")
        if nspins <= 0:
            raise ValueError("The total number of spins has to be POSITIVE.")
        self.nspins = int(nspins)
        self.mpi_rank = 0
        self.mpi_size = 1
        self.disorder_flg = False


def _nspins(env) -> int:
    return env.nspins if isinstance(env, Environment) else int(env)


class Basis:
    """Basis(env, total_mag) -- src/basis.cpp:56-94 without the stored list."""

    def __init__(self, env, total_mag: int):
        N = _nspins(env)
        if total_mag > N // 2 or total_mag < -(N // 2):
            raise ValueError("The required constant magnetization has to be <= nspins/2.")
        self.nspins = N
        self.total_mag = int(total_mag)
        self.nspins_up = self.total_mag + N // 2
        self.size = int(lib().lsqed_sector_dim(N, self.nspins_up))
        self.local_size = self.size
        self.global_start_index = 0
        self.mpi_size, self.mpi_rank = 1, 0


class _Lattice:
    type_id = -1

    def __init__(self, env, lx: int = 0, ly: int = 0):
        self.nspins = _nspins(env)
        self.nspins_x, self.nspins_y = lx, ly
        self.nn_bonds = self._bonds(0)
        self.nnn_bonds = self._bonds(1)

    def _bonds(self, kind):
        n = C.c_int()
        check(lib().lsqed_lattice_bonds(self.type_id, self.nspins, self.nspins_x, self.nspins_y, kind, None, None,
                                        C.byref(n)))
        bi = np.zeros(n.value, np.int32)
        bj = np.zeros(n.value, np.int32)
        check(lib().lsqed_lattice_bonds(self.type_id, self.nspins, self.nspins_x, self.nspins_y, kind,
                                        bi.ctypes.data_as(C.c_void_p), bj.ctypes.data_as(C.c_void_p), C.byref(n)))
        return bi, bj

    def num_spins(self) -> int:
        return self.nspins

    def coords(self) -> np.ndarray:
        """Site positions r_i = (i // Ly, i % Ly) (src/lattice.cpp:94)."""
        ly = self.nspins_y if self.nspins_y else 1
        i = np.arange(self.nspins)
        if not self.nspins_y:
            return np.stack([i, np.zeros_like(i)], 1).astype(float)
        return np.stack([i // ly, i % ly], 1).astype(float)


class chain1D(_Lattice):
    """chain1D(env): periodic chain, nn (i, i+1), nnn (i, i+2) (src/lattice.cpp:7-53)."""
    type_id = _lib.CHAIN1D

    def __init__(self, env):
        super().__init__(env, 0, 0)


_PRODUCT_MSG = ("The product between the number of spins in x direction and the numer of spins in the y direction "
                "has to be equal to the total number of spins.")


class square2D(_Lattice):
    """square2D(env, lx, ly) (src/lattice.cpp:59-136)."""
    type_id = _lib.SQUARE2D

    def __init__(self, env, lx: int, ly: int):
        if lx * ly != _nspins(env):
            raise ValueError(_PRODUCT_MSG)
        super().__init__(env, lx, ly)


class honeycomb2D(_Lattice):
    """honeycomb2D(env, lx, ly) (src/lattice.cpp:142-264)."""
    type_id = _lib.HONEYCOMB2D

    def __init__(self, env, lx: int, ly: int):
        if lx * ly != _nspins(env):
            raise ValueError(_PRODUCT_MSG)
        if lx % 2 != 0 or ly % 2 != 0:
            raise ValueError("The number of spins in each direction has to be even.")
        super().__init__(env, lx, ly)


class AF:
    """Checkerboard sublattices (src/sublattice.cpp:5-57, :95-119)."""

    def __init__(self, lattice, size: int = 2):
        self.lattice = lattice
        N = lattice.nspins
        if isinstance(lattice, chain1D):
            self.sublat = [[i for i in range(N) if i % 2 == 0], [i for i in range(N) if i % 2 == 1]]
        else:
            lx, ly = lattice.nspins_x, lattice.nspins_y
            sl = [[], []]
            tagx = 0
            for ix in range(lx):
                tagy = tagx
                for iy in range(ly):
                    sl[1 if tagy % 2 else 0].append(ix * ly + iy)
                    tagy ^= 1
                tagx ^= 1
            self.sublat = sl

    def get_size(self) -> int:
        return len(self.sublat)

    def get_sl(self, k: int):
        return self.sublat[k]


class Striped:
    """Striped<square2D> (src/sublattice.cpp:61-91): sublattices 0/1 by the
    parity of iy, 2/3 by the parity of ix."""

    def __init__(self, lattice, size: int = 4):
        self.lattice = lattice
        lx, ly = lattice.nspins_x, lattice.nspins_y
        sl = [[], [], [], []]
        for ix in range(lx):
            for iy in range(ly):
                cur = ix * ly + iy
                sl[1 if iy % 2 else 0].append(cur)
                sl[3 if ix % 2 else 2].append(cur)
        self.sublat = sl

    def get_size(self) -> int:
        return len(self.sublat)

    def get_sl(self, k: int):
        return self.sublat[k]


class Hamiltonian:
    """Hamiltonian(env, basis, lattice, J1, D1, J2, D2)
    (include/hamiltonian.tmpl_func.hpp:69-91).

    The operator starts empty, like the freshly preallocated MATMPIAIJ; each
    ``build_*`` call adds its part: build_diag (tmpl:129-225),
    build_off_diag (tmpl:229-344), build_disorder (tmpl:97-125).
    """

    def __init__(self, env, basis: Basis, lattice, J1: float, D1: float, J2: float, D2: float):
        if basis.nspins != lattice.nspins:
            raise ValueError("basis and lattice have different numbers of spins")
        self.basis, self.lattice = basis, lattice
        self.size = basis.size
        self.J1, self.Delta1, self.J2, self.Delta2 = float(J1), float(D1), float(J2), float(D2)
        (a, b), (c, d) = lattice.nn_bonds, lattice.nnn_bonds
        self._keep = (a, b, c, d)
        P = C.POINTER(C.c_int)
        pr = _lib.Params(basis.nspins, basis.nspins_up, len(a), len(c),
                         a.ctypes.data_as(P), b.ctypes.data_as(P), c.ctypes.data_as(P), d.ctypes.data_as(P),
                         self.J1, self.Delta1, self.J2, self.Delta2)
        h = C.c_void_p()
        check(lib().lsqed_ham_create(C.byref(pr), C.byref(h)))
        self.hamilt = h
        self._flags = 0
        check(lib().lsqed_ham_set_build(self.hamilt, 0))
        self.hi = None

    def __del__(self):
        h = getattr(self, "hamilt", None)
        if h is not None and h.value:
            try:
                lib().lsqed_ham_destroy(h)
            except Exception:
                pass
            self.hamilt = None

    # --- build calls ------------------------------------------------------
    def build_diag(self, env=None) -> int:
        self._flags |= _lib.BUILD_DIAG
        check(lib().lsqed_ham_set_build(self.hamilt, self._flags))
        return 0

    def build_off_diag(self, env=None) -> int:
        self._flags |= _lib.BUILD_OFFDIAG
        check(lib().lsqed_ham_set_build(self.hamilt, self._flags))
        return 0

    def build_disorder(self, env, hi) -> int:
        hi = np.ascontiguousarray(hi, np.float64)
        if hi.shape != (self.basis.nspins,):
            raise ValueError("hi must hold one field per spin")
        self.hi = hi.copy()
        check(lib().lsqed_ham_set_fields(self.hamilt, hi.ctypes.data_as(C.c_void_p)))
        self._flags |= _lib.BUILD_DISORDER
        check(lib().lsqed_ham_set_build(self.hamilt, self._flags))
        return 0

    def set_couplings(self, J1: float, D1: float, J2: float, D2: float) -> None:
        """New couplings for this model, as if constructed with them
        (tmpl:69-91); lattice, sector and build calls are kept."""
        J1, D1, J2, D2 = float(J1), float(D1), float(J2), float(D2)
        check(lib().lsqed_ham_set_couplings(self.hamilt, J1, D1, J2, D2))
        self.J1, self.Delta1, self.J2, self.Delta2 = J1, D1, J2, D2

    # --- accessors (include/hamiltonian.hpp:257-299) -----------------------
    def hamilt_size(self) -> int:
        return self.size

    def hamilt_dim(self) -> int:
        return self.size * self.size

    def hamilt_J1(self) -> float:
        return self.J1

    def hamilt_J2(self) -> float:
        return self.J2

    def hamilt_D1(self) -> float:
        return self.Delta1

    def hamilt_D2(self) -> float:
        return self.Delta2

    # --- device operations ------------------------------------------------
    HV_AUTO, HV_GENERIC, HV_SPLIT, HV_SPLIT_JIT = 0, 1, 2, 3

    def set_hv_kernel(self, variant: int) -> None:
        """Select the H.v kernel (0 auto, 1 generic chunk table, 2 row split,
        3 row split specialised to this lattice with hipRTC)."""
        check(lib().lsqed_ham_set_hv_kernel(self.hamilt, int(variant)))

    def hv_kernel(self) -> int:
        """The kernel matvec uses (1 generic, 2 row split, 3 specialised row split)."""
        return int(lib().lsqed_ham_get_hv_kernel(self.hamilt))

    def hv_fingerprint(self) -> str:
        """16 hex digits identifying the H.v launch (kernel, grid, model and,
        for the specialised kernel, its generated source); measurement
        records are keyed by it."""
        import ctypes as C

        v = C.c_uint64()
        check(lib().lsqed_ham_hv_fingerprint(self.hamilt, C.byref(v)))
        return f"{v.value:016x}"

    def matvec(self, x, y=None, stream=None):
        """y = H x on the GPU (x, y: float64 CUDA tensors of length D)."""
        import torch

        _check_vec(x, self.size)
        if y is None:
            y = torch.empty_like(x)
        _check_vec(y, self.size)
        check(lib().lsqed_hv(self.hamilt, ptr(x), ptr(y), stream_ptr(stream)))
        return y

    def matvec_c128(self, x, y=None, stream=None):
        """y = H x on complex vectors (complex128 CUDA tensors of length D,
        the layout of a complex-scalar PETSc Vec); both parts in one pass."""
        import torch

        _check_vec(x, self.size, torch.complex128)
        if y is None:
            y = torch.empty_like(x)
        _check_vec(y, self.size, torch.complex128)
        check(lib().lsqed_hv_c128(self.hamilt, ptr(x), ptr(y), stream_ptr(stream)))
        return y

    def diagonal(self, stream=None, complex128=False):
        import torch

        if complex128:
            d = torch.empty(self.size, dtype=torch.complex128, device="cuda")
            check(lib().lsqed_diag_c128(self.hamilt, ptr(d), stream_ptr(stream)))
            return d
        d = torch.empty(self.size, dtype=torch.float64, device="cuda")
        check(lib().lsqed_diag(self.hamilt, ptr(d), stream_ptr(stream)))
        return d

    def export_cols(self, kind: int, row0: int = 0, nrows: int | None = None, stream=None):
        """Column lists in get_coup_elems emission order: (cols[n, maxc], ncol[n])."""
        import torch

        if nrows is None:
            nrows = self.size - row0
        nb = len(self.lattice.nn_bonds[0] if kind == 0 else self.lattice.nnn_bonds[0])
        maxc = max(nb, 1)
        cols = torch.full((max(nrows, 1), maxc), -1, dtype=torch.int64, device="cuda")
        ncol = torch.zeros(max(nrows, 1), dtype=torch.int32, device="cuda")
        check(lib().lsqed_export_cols(self.hamilt, kind, row0, nrows, maxc, ptr(cols), ptr(ncol), stream_ptr(stream)))
        return cols[:nrows], ncol[:nrows]

    def unrank(self, row0: int, n: int, stream=None):
        import torch

        out = torch.empty(max(n, 1), dtype=torch.int64, device="cuda")
        check(lib().lsqed_unrank(self.hamilt, row0, n, ptr(out), stream_ptr(stream)))
        return out[:n]

    def rank(self, states, stream=None):
        import torch

        states = states.to(torch.int64).contiguous()
        out = torch.empty(max(states.numel(), 1), dtype=torch.int64, device="cuda")
        check(lib().lsqed_rank(self.hamilt, ptr(states), states.numel(), ptr(out), stream_ptr(stream)))
        return out[: states.numel()]


def _check_vec(t, n, dtype=None):
    import torch

    dtype = dtype or torch.float64
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous()
            and t.numel() == n):
        raise ValueError(f"expected a contiguous {dtype} CUDA tensor of {n} elements")
