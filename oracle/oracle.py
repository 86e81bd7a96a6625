"""CPU oracle for the Sz-sector Hamiltonian path of LSQuantumED.

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product package ``lsquantumed_amd`` never does.

It wraps ``oracle/liboracle.so`` (built by ``oracle/Makefile`` from
``oracle/oracle.cpp``, a routine-by-routine restatement of the reference's
arithmetic) and adds the numpy-side pieces of the reference that are plain
linear algebra:

* dense / sparse eigensolves standing in for SLEPc's EPS (src/solver.cpp:7-68),
* sublattices (src/sublattice.cpp:5-91),
* <S_i.S_j>, m^2_AF, m^2_STR (src/spinOperators/correlation.cpp:276-299,
  include/phys.tmpl_func.hpp:5-81),
* DSF amplitudes and the PETSc/DSF text formats
  (include/phys.tmpl_func.hpp:87-215).

Parity status of every function is listed in ``PARITY`` below; every entry
is pinned by a reference fixture, a reference test literal, or a closed
form / brute-force check in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

CHAIN, SQUARE, HONEYCOMB = 0, 1, 2

# function -> what pins it (tests/test_oracle_*.py)
PARITY = {
    "binom_boost": "exact integer binomials for all n<=64 (closed form)",
    "rank_ref/unrank_ref/int_basis": "int_basis.ref, bit_basis.ref, test_basis1-4 literals, brute-force sort",
    "bonds": "test_lattice1-11 literals, get_coup_elems emission order literals",
    "elem_diag": "get_elem_diag1-5 literals, build_diag*.ref",
    "disorder_elem": "get_disorder_elem1/3 literals, build_disorder*.ref",
    "coup_elems": "get_coup_elems1-5 literals, build_off_diag*.ref",
    "csr": "build_hamilt*.ref, build_off_diag*.ref, build_disorder125.ref",
    "eigs": "test_solver1-5 / test_dsf1 eigenvalue literals (1e-13)",
    "szsz/spsm": "test_szsz1/2, test_spsm1 literals",
    "rd_seeds/rd_hi": "test_tools_random1 literals",
    "sublattices": "test_sublattice1-3 literals",
    "dsf": "dsf_d0_1.ref / dsf_d0_2.ref energies and multiplet weights",
    "m2_af/m2_str": "closed forms: Neel and polarised states (m2_AF = 1), stripes, random basis states, "
                    "singlet products inside / across the AF sublattices",
    "hv_matfree": "equal to the assembled CSR (build_*.ref) times x on 6 models with and without fields",
}


def build() -> str:
    """Compile liboracle.so with oracle/Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        i32, i64, u64, dbl, vp = C.c_int, C.c_int64, C.c_uint64, C.c_double, C.c_void_p
        P = C.POINTER
        sig = {
            "oq_binom_boost": (dbl, [i32, i32]),
            "oq_binom_exact": (u64, [i32, i32]),
            "oq_rank_ref": (i64, [u64, i32]),
            "oq_rank_exact": (u64, [u64]),
            "oq_unrank_ref": (u64, [i64, i32, i32]),
            "oq_compute_size": (i64, [i32, i32]),
            "oq_split": (None, [i64, i32, i32, P(i64), P(i64)]),
            "oq_int_basis": (None, [i32, i32, i64, i64, vp]),
            "oq_lattice": (i32, [i32, i32, i32, i32, vp, vp, vp, vp, P(i32), P(i32)]),
            "oq_bonds": (i32, [i32, i32, i32, i32, i32, vp, vp]),
            "oq_elem_diag": (i64, [u64, i32, vp, vp]),
            "oq_disorder_elem": (dbl, [u64, i32, vp]),
            "oq_coup_elems": (i32, [u64, i32, i32, vp, vp, vp]),
            "oq_model_create": (vp, [i32, i32, i32, i32, i32, dbl, dbl, dbl, dbl]),
            "oq_model_free": (None, [vp]),
            "oq_diag": (None, [vp, vp, i32, vp, vp]),
            "oq_csr_build": (vp, [vp, vp, i32]),
            "oq_csr_dim": (i64, [vp]),
            "oq_csr_nnz": (i64, [vp]),
            "oq_csr_copy": (None, [vp, vp, vp, vp]),
            "oq_csr_free": (None, [vp]),
            "oq_csr_spmv": (None, [vp, vp, vp, i32]),
            "oq_csrz_from": (vp, [vp]),
            "oq_csrz_free": (None, [vp]),
            "oq_csrz_spmv": (None, [vp, vp, vp, i32]),
            "oq_hv_matfree": (None, [vp, vp, vp, vp, i64, i64, i32]),
            "oq_hv_rows_formula": (None, [vp, vp, vp, i64, i64, i32]),
            "oq_formula_x": (C.c_double, [i64]),
            "oq_szsz": (dbl, [i32, i32, vp, i32, i32]),
            "oq_spsm": (dbl, [i32, i32, vp, i32, i32, i32]),
            "oq_sqz_apply": (None, [i32, i32, vp, i32, vp]),
            "oq_sqp_apply": (None, [i32, i32, vp, i32, vp]),
            "oq_rd_seeds": (None, [i32, vp]),
            "oq_rd_hi": (None, [i64, i32, dbl, dbl, vp]),
            "oq_max_threads": (i32, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# --- basis (src/basis.cpp) ------------------------------------------------
def binom_boost(n: int, k: int) -> float:
    return lib().oq_binom_boost(n, k)


def compute_size(N: int, nup: int) -> int:
    return lib().oq_compute_size(N, nup)


def nup_of(N: int, mag: int) -> int:
    """Basis::Basis: nspins_up = total_mag + nspins/2, |mag| <= N/2 (src/basis.cpp:66-71)."""
    if mag > N // 2 or mag < -(N // 2):
        raise ValueError("The required constant magnetization has to be <= nspins/2.")
    return mag + N // 2


def rank_ref(s: int, N: int) -> int:
    return lib().oq_rank_ref(s, N)


def unrank_ref(idx: int, N: int, nup: int) -> int:
    return lib().oq_unrank_ref(idx, N, nup)


def split(size: int, P: int, rank: int) -> tuple[int, int]:
    a, b = C.c_int64(), C.c_int64()
    lib().oq_split(size, P, rank, C.byref(a), C.byref(b))
    return a.value, b.value


def int_basis(N: int, nup: int, start: int = 0, count: int | None = None) -> np.ndarray:
    D = compute_size(N, nup)
    if count is None:
        count = D - start
    out = np.zeros(count, dtype=np.uint64)
    lib().oq_int_basis(N, nup, start, count, _p(out))
    return out


# --- lattices (src/lattice.cpp) -------------------------------------------
def lattice_lists(ltype: int, N: int, lx: int = 0, ly: int = 0) -> dict:
    stride = {CHAIN: (2, 2), SQUARE: (3, 3), HONEYCOMB: (3, 4)}[ltype]
    nn = np.zeros(N * stride[0], np.int64)
    nnn = np.zeros(N * stride[1], np.int64)
    nnX = np.zeros(N, np.int64)
    nnnX = np.zeros(N, np.int64)
    a, b = C.c_int(), C.c_int()
    rc = lib().oq_lattice(ltype, N, lx, ly, _p(nn), _p(nnn), _p(nnX), _p(nnnX), C.byref(a), C.byref(b))
    if rc != 0:
        raise ValueError("invalid lattice dimensions")
    return dict(nn=nn, nnn=nnn, nnXsite=nnX, nnnXsite=nnnX, num_nn=a.value, num_nnn=b.value)


def bonds(ltype: int, N: int, lx: int = 0, ly: int = 0, kind: int = 0) -> tuple[np.ndarray, np.ndarray]:
    n = lib().oq_bonds(ltype, N, lx, ly, kind, None, None)
    if n < 0:
        raise ValueError("invalid lattice dimensions")
    bi = np.zeros(n, np.int32)
    bj = np.zeros(n, np.int32)
    lib().oq_bonds(ltype, N, lx, ly, kind, _p(bi), _p(bj))
    return bi, bj


def elem_diag(s: int, bi: np.ndarray, bj: np.ndarray) -> int:
    bi = np.ascontiguousarray(bi, np.int32)
    bj = np.ascontiguousarray(bj, np.int32)
    return lib().oq_elem_diag(s, len(bi), _p(bi), _p(bj))


def disorder_elem(s: int, N: int, hi: np.ndarray) -> float:
    hi = np.ascontiguousarray(hi, np.float64)
    return lib().oq_disorder_elem(s, N, _p(hi))


def coup_elems(s: int, N: int, bi: np.ndarray, bj: np.ndarray) -> list[int]:
    bi = np.ascontiguousarray(bi, np.int32)
    bj = np.ascontiguousarray(bj, np.int32)
    out = np.zeros(len(bi) + 1, np.int64)
    n = lib().oq_coup_elems(s, N, len(bi), _p(bi), _p(bj), _p(out))
    return [int(v) for v in out[:n]]


class Model:
    """Lattice + couplings + sector; mirrors Hamiltonian<L>'s inputs
    (include/hamiltonian.tmpl_func.hpp:69-91)."""

    BUILD_DIAG, BUILD_OFFDIAG, BUILD_DISORDER = 1, 2, 4

    def __init__(self, ltype, N, nup, J1=1.0, D1=1.0, J2=0.0, D2=0.0, lx=0, ly=0):
        self.ltype, self.N, self.nup, self.lx, self.ly = ltype, N, nup, lx, ly
        self.J1, self.D1, self.J2, self.D2 = J1, D1, J2, D2
        self.h = lib().oq_model_create(ltype, N, lx, ly, nup, J1, D1, J2, D2)
        if not self.h:
            raise ValueError("invalid lattice")
        self.D = compute_size(N, nup)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oq_model_free(self.h)
            self.h = None

    def diag(self, hi=None, flags=BUILD_DIAG):
        d = np.zeros(self.D)
        pres = np.zeros(self.D, np.uint8)
        hp = None
        if hi is not None:
            hi = np.ascontiguousarray(hi, np.float64)
            hp = _p(hi)
        lib().oq_diag(self.h, hp, flags, _p(d), _p(pres))
        return d, pres.astype(bool)

    def csr_arrays(self, hi=None, flags=BUILD_DIAG | BUILD_OFFDIAG):
        hp = None
        if hi is not None:
            hi = np.ascontiguousarray(hi, np.float64)
            hp = _p(hi)
        a = lib().oq_csr_build(self.h, hp, flags)
        D, nnz = lib().oq_csr_dim(a), lib().oq_csr_nnz(a)
        rp = np.zeros(D + 1, np.int64)
        cols = np.zeros(nnz, np.int64)
        vals = np.zeros(nnz)
        lib().oq_csr_copy(a, _p(rp), _p(cols), _p(vals))
        lib().oq_csr_free(a)
        return rp, cols, vals

    def csr(self, hi=None, flags=BUILD_DIAG | BUILD_OFFDIAG):
        import scipy.sparse as sp

        rp, cols, vals = self.csr_arrays(hi, flags)
        return sp.csr_matrix((vals, cols, rp), shape=(self.D, self.D))

    def hv_matfree(self, x, hi=None, row0=0, nrows=None, nthreads=0):
        x = np.ascontiguousarray(x, np.float64)
        if nrows is None:
            nrows = self.D - row0
        y = np.zeros(nrows)
        hp = None
        if hi is not None:
            hi = np.ascontiguousarray(hi, np.float64)
            hp = _p(hi)
        lib().oq_hv_matfree(self.h, hp, _p(x), _p(y), row0, nrows, nthreads)
        return y

    def hv_rows_formula(self, row0, nrows, hi=None, nthreads=0):
        """Rows [row0, row0+nrows) of H x with x(c) = formula_x(c)."""
        y = np.zeros(nrows)
        hp = None
        if hi is not None:
            hi = np.ascontiguousarray(hi, np.float64)
            hp = _p(hi)
        lib().oq_hv_rows_formula(self.h, hp, _p(y), row0, nrows, nthreads)
        return y


class ComplexCsr:
    """The model's assembled matrix in the reference's layout (complex128
    values, int64 columns), for the CPU baseline's SpMV."""

    def __init__(self, model, hi=None, flags=Model.BUILD_DIAG | Model.BUILD_OFFDIAG):
        hp = None
        if hi is not None:
            hi = np.ascontiguousarray(hi, np.float64)
            hp = _p(hi)
        a = lib().oq_csr_build(model.h, hp, flags)
        self.D, self.nnz = lib().oq_csr_dim(a), lib().oq_csr_nnz(a)
        self.z = lib().oq_csrz_from(a)
        lib().oq_csr_free(a)

    def __del__(self):
        if getattr(self, "z", None):
            lib().oq_csrz_free(self.z)
            self.z = None

    def matvec(self, x: np.ndarray, y: np.ndarray, nthreads: int = 0) -> None:
        lib().oq_csrz_spmv(self.z, _p(x), _p(y), nthreads)


def lanczos_ground_cpu(A: ComplexCsr, tol: float = 1e-9, maxit: int = 500, nthreads: int = 0, seed: int = 1):
    """Three-term Lanczos on complex vectors with the CSR SpMV, stopped when
    the lowest Ritz pair has |beta_m s_m| / |theta| < tol (checked every 5
    steps).  Returns (E0, steps)."""
    import scipy.linalg as sl

    rng = np.random.default_rng(seed)
    v = (rng.random(A.D) - 0.5).astype(np.complex128)
    v /= np.linalg.norm(v)
    vp = np.zeros_like(v)
    w = np.empty_like(v)
    al, be = [], []
    beta = 0.0
    theta = 0.0
    for m in range(maxit):
        A.matvec(v, w, nthreads)
        a = np.vdot(v, w).real
        w -= a * v
        w -= beta * vp
        beta = np.linalg.norm(w)
        al.append(a)
        be.append(beta)
        if m % 5 == 4:
            ev, s = sl.eigh_tridiagonal(np.array(al), np.array(be[:-1]), select="i", select_range=(0, 0))
            theta = float(ev[0])
            if abs(be[-1] * s[-1, 0]) < tol * abs(theta):
                return theta, m + 1
        vp, v = v, w / beta
        w = np.empty_like(v)
    return theta, maxit


FORMULA_P = ((1000003, 7919), (999983, 104729))


def formula_x(c: np.ndarray) -> np.ndarray:
    """x(c) = (((c mod P1) K1 mod P1) + ((c mod P2) K2 mod P2)) * 2^-20 - 1,
    the vector of oq_hv_rows_formula (exact in floating point)."""
    c = np.asarray(c, np.int64)
    (p1, k1), (p2, k2) = FORMULA_P
    n = (c % p1) * k1 % p1 + (c % p2) * k2 % p2
    return n.astype(np.float64) * (1.0 / 1048576.0) - 1.0


# --- eigensolve (stands in for SLEPc EPS, src/solver.cpp:7-68) -------------
def eigs(A, nev: int, dense_max: int = 6000):
    """Lowest ``nev`` eigenpairs (ascending) of the real symmetric sparse A."""
    D = A.shape[0]
    if D <= dense_max:
        w, v = np.linalg.eigh(A.toarray())
        return w[:nev], v[:, :nev]
    import scipy.sparse.linalg as sla

    w, v = sla.eigsh(A, k=nev, which="SA", tol=1e-13, ncv=max(2 * nev + 1, 40))
    o = np.argsort(w)
    return w[o], v[:, o]


# --- observables -------------------------------------------------------------
def szsz(N, nup, psi, i, j) -> float:
    psi = np.ascontiguousarray(psi, np.float64)
    return lib().oq_szsz(N, nup, _p(psi), i, j)


def spsm(N, nup, psi, i, j) -> float:
    psi = np.ascontiguousarray(psi, np.float64)
    return lib().oq_spsm(N, nup, _p(psi), i, j, 1)


def smsp(N, nup, psi, i, j) -> float:
    psi = np.ascontiguousarray(psi, np.float64)
    return lib().oq_spsm(N, nup, _p(psi), i, j, 0)


def sisj(N, nup, psi, i, j) -> float:
    """SiSj::ExpectVal, src/spinOperators/correlation.cpp:276-299."""
    v = spsm(N, nup, psi, i, j)
    v += smsp(N, nup, psi, i, j)
    v /= 2.0
    v += szsz(N, nup, psi, i, j)
    return v


def sqz_apply(N, nup, psi, spin):
    psi = np.ascontiguousarray(psi, np.float64)
    out = np.zeros_like(psi)
    lib().oq_sqz_apply(N, nup, _p(psi), spin, _p(out))
    return out


def sqp_apply(N, nup, psi, spin):
    psi = np.ascontiguousarray(psi, np.float64)
    out = np.zeros(compute_size(N, nup + 1))
    lib().oq_sqp_apply(N, nup, _p(psi), spin, _p(out))
    return out


# --- sublattices (src/sublattice.cpp:5-91) -----------------------------------
def af_sublattices(ltype, N, lx=0, ly=0):
    """AF<chain1D> (:5-24), AF<square2D> / AF<honeycomb2D> (:28-57, :95-119)."""
    if ltype == CHAIN:
        sl = [[], []]
        for i in range(N):
            sl[1 if i % 2 else 0].append(i)
        return sl
    sl = [[], []]
    tagx = 0
    for ix in range(lx):
        tagy = tagx
        for iy in range(ly):
            sl[1 if tagy % 2 else 0].append(ix * ly + iy)
            tagy ^= 1
        tagx ^= 1
    return sl


def striped_sublattices(lx, ly):
    """Striped<square2D>, src/sublattice.cpp:61-91."""
    sl = [[], [], [], []]
    tagx = 0
    for ix in range(lx):
        tagy = 0
        for iy in range(ly):
            cur = ix * ly + iy
            sl[1 if tagy % 2 else 0].append(cur)
            sl[3 if tagx % 2 else 2].append(cur)
            tagy ^= 1
        tagx ^= 1
    return sl


def m2_af(N, nup, psi, sublat) -> float:
    """Phys::SquaredSubLattMagAF, include/phys.tmpl_func.hpp:5-41."""
    acc = 0.0
    for sl in sublat:
        for i in sl:
            for j in sl:
                acc += sisj(N, nup, psi, i, j)
    return acc * 8.0 / float(N) / float(N + 4.0)


def m2_str(N, nup, psi, sublat) -> float:
    """Phys::SquaredSubLattMagSTR, include/phys.tmpl_func.hpp:45-81."""
    acc = 0.0
    for st in range(2):
        for lat in range(2):
            sl = sublat[st * 2 + lat]
            for i in sl:
                for j in sl:
                    acc += sisj(N, nup, psi, i, j)
    return acc * 4.0 / float(N) / float(N + 4.0)


def dsf_amplitudes(N, nup0, psi0, evecs, op="z"):
    """a[n, i] = <n| S_i |0> as DynStructFactor forms it
    (include/phys.tmpl_func.hpp:166-186): rhs = S_i|0>, then VecDot(rhs, e_n)."""
    nev = evecs.shape[1]
    a = np.zeros((nev, N))
    for i in range(N):
        rhs = sqz_apply(N, nup0, psi0, i) if op == "z" else sqp_apply(N, nup0, psi0, i)
        a[:, i] = evecs.T @ rhs
    return a


def dsf_weight(a_row: np.ndarray, qvec, coords) -> float:
    """W_n(q) = |sum_i exp(i q.r_i) a_{n,i}|^2 / N
    (postprocessing/dynamicStep0_preprocess.ipynb cells 12-13)."""
    ph = np.exp(1j * (coords @ np.asarray(qvec, float)))
    return float(abs(np.sum(ph * a_row)) ** 2 / len(a_row))


def multiplets(evals, tol=1e-8):
    """Group ascending eigenvalues into degenerate multiplets (index lists)."""
    groups, cur = [], [0]
    for k in range(1, len(evals)):
        if abs(evals[k] - evals[cur[0]]) <= tol * max(1.0, abs(evals[k])):
            cur.append(k)
        else:
            groups.append(cur)
            cur = [k]
    groups.append(cur)
    return groups


# --- Tools::RandomDisorder (src/tools/random_disorder.cpp:33-59) -------------
def rd_seeds(rep: int) -> np.ndarray:
    out = np.zeros(rep, np.int64)
    lib().oq_rd_seeds(rep, _p(out))
    return out


def rd_hi(seed: int, N: int, lo: float, hi: float) -> np.ndarray:
    out = np.zeros(N)
    lib().oq_rd_hi(seed, N, lo, hi, _p(out))
    return out


# --- PETSc ASCII formats ---------------------------------------------------------
def petsc_g(v: float) -> str:
    """PETSc's ASCII %g: like C's %g, with a '.' forced onto integral output."""
    s = "%g" % v
    if not any(c in s for c in ".eEn"):
        s += "."
    return s


def parse_petsc_vec(text: str) -> list[str]:
    return [ln.strip() for ln in text.splitlines()
            if ln.strip() and not ln.startswith(("Vec Object", "  type", "Process"))]


def parse_petsc_mat(text: str) -> list[list[tuple[int, str]]]:
    rows = []
    for ln in text.splitlines():
        if not ln.startswith("row "):
            continue
        body = ln.split(":", 1)[1]
        ents = []
        for part in body.split(")"):
            part = part.strip()
            if not part:
                continue
            c, v = part.lstrip("(").split(",")
            ents.append((int(c), v.strip()))
        rows.append(ents)
    return rows


def csr_to_petsc_rows(rp, cols, vals) -> list[list[tuple[int, str]]]:
    return [[(int(cols[q]), petsc_g(vals[q])) for q in range(rp[r], rp[r + 1])] for r in range(len(rp) - 1)]


def binom_int(n: int, k: int) -> int:
    return math.comb(n, k) if 0 <= k <= n else 0
