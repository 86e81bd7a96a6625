"""PETSc ASCII views of the operator, in the format the reference's tests
dump and diff (MatView / VecView with the default ASCII viewer:
tests/unit_tests/build_off_diag/chain_1d/test_build_off_diag60.cpp:27,
tests/unit_tests/build_diag/chain_1d/test_build_diag60.cpp:29-33, compared
by tests/unit_tests/compare.sh against reference_data/*.ref).

    text = mat_view(h, nprocs=2)        # what MatView(h.hamilt) prints
    text = diag_view(h, nprocs=2)       # VecView of MatGetDiagonal(h.hamilt)

The rows are those of the matrix the reference assembles
(include/hamiltonian.tmpl_func.hpp:97-344): off-diagonal J_k/2 per
antiparallel bond of kind k with |J_k| >= 1e-14, duplicates (L = 2 lattices)
summed; the diagonal is stored when build_disorder inserted it (every row) or
when build_diag added a nonzero bond count of a kind with |D_k| >= 1e-14.
Columns come from the device column export, values of the diagonal from the
device diagonal (bit-exact), so the text is produced from the library's own
operator.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def petsc_g(v: float) -> str:
    """PETSc's ASCII real format: C's %g, with a '.' appended to integral
    output (PetscViewerASCIIPrintf of a PetscScalar)."""
    s = "%g" % v
    if not any(c in s for c in ".eEn"):
        s += "."
    return s


def ownership(size: int, nprocs: int, rank: int) -> tuple[int, int]:
    """(local size, first index) of `rank`: the row split of the reference's
    Basis (src/basis.cpp:78-80), which PETSc's default layout reproduces."""
    rest = size % nprocs
    local = size // nprocs + (1 if rest > rank else 0)
    start = rank * local + (rest if rest <= rank else 0)
    return local, start


def _procs(n: int) -> str:
    # the PETSc release of the reference's dumps prints "processes" for any count
    return f"{n} MPI processes"


def vec_view(values, nprocs: int = 1) -> str:
    """VecView of the vector MatCreateVecs gives on `nprocs` ranks: a
    sequential vector on one rank, else an MPI vector with one block per
    rank."""
    v = np.asarray(values, dtype=np.float64).ravel()
    out = [f"Vec Object: {_procs(nprocs)}", "  type: " + ("seq" if nprocs == 1 else "mpi")]
    for r in range(nprocs):
        n, s = ownership(len(v), nprocs, r)
        if nprocs > 1:
            out.append(f"Process [{r}]")
        out.extend(petsc_g(x) for x in v[s:s + n])
    return "\n".join(out) + "\n"


def diag_view(h, nprocs: int = 1, stream=None) -> str:
    """VecView of MatGetDiagonal(hamilt): unset diagonal entries read 0."""
    return vec_view(h.diagonal(stream).cpu().numpy(), nprocs)


def mat_rows(h, row0: int = 0, nrows: int | None = None, stream=None) -> list[list[tuple[int, float]]]:
    """Stored entries (column ascending) of rows [row0, row0 + nrows) of the
    matrix the reference assembles."""
    D = h.hamilt_size()
    if nrows is None:
        nrows = D - row0
    if row0 < 0 or nrows < 0 or row0 + nrows > D:
        raise ValueError("row range out of bounds")
    if nrows == 0:
        return []
    flags = h._flags
    J = (h.J1, h.J2)
    Dz = (h.Delta1, h.Delta2)
    nb = (len(h.lattice.nn_bonds[0]), len(h.lattice.nnn_bonds[0]))
    cols, ncol = [], []
    for kind in (0, 1):
        c, n = h.export_cols(kind, row0, nrows, stream)
        cols.append(c.cpu().numpy())
        ncol.append(n.cpu().numpy())
    diag = h.diagonal(stream)[row0:row0 + nrows].cpu().numpy()
    off = bool(flags & _lib.BUILD_OFFDIAG)
    rows = []
    for r in range(nrows):
        ent: dict[int, float] = {}
        if off:
            for kind in (0, 1):
                if abs(J[kind]) < 1e-14:
                    continue
                for c in cols[kind][r, :ncol[kind][r]]:
                    ent[int(c)] = ent.get(int(c), 0.0) + J[kind] * 0.5
        if flags & _lib.BUILD_DISORDER:
            present = True
        elif flags & _lib.BUILD_DIAG:
            present = any(abs(Dz[k]) >= 1e-14 and nb[k] - 2 * int(ncol[k][r]) != 0 for k in (0, 1))
        else:
            present = False
        if present:
            ent[row0 + r] = float(diag[r])
        rows.append(sorted(ent.items()))
    return rows


def mat_view(h, nprocs: int = 1, stream=None) -> str:
    """MatView of the assembled MATMPIAIJ over `nprocs` ranks."""
    out = [f"Mat Object: {_procs(nprocs)}", "  type: mpiaij"]
    for r, ents in enumerate(mat_rows(h, stream=stream)):
        out.append(f"row {r}:" + "".join(f" ({c}, {petsc_g(v)}) " for c, v in ents))
    return "\n".join(out) + "\n"


# ------------------------------------------------------------------ import --
# The reverse direction: a VecView / MatView ASCII dump (the files the
# reference's tests write and compare.sh diffs) read back into the library's
# vectors.  Values are PETSc's %g text (6 significant digits), so a loaded
# dump reproduces its own text exactly and the vector it was printed from to
# %g precision.

def _text(src) -> str:
    if isinstance(src, str) and ("\n" in src or src.startswith(("Vec Object", "Mat Object"))):
        return src
    with open(src) as fh:
        return fh.read()


def parse_scalar(tok: str) -> complex:
    """One PETSc ASCII scalar: %g with an optional forced '.', or a complex
    value written as 're + im i' / 're - im i' (complex-scalar builds)."""
    t = tok.strip()
    if t.endswith("i") and (" + " in t or " - " in t):
        sep = " + " if " + " in t else " - "
        re_s, im_s = t[:-1].rsplit(sep, 1)
        im = float(im_s.strip())
        return complex(float(re_s), im if sep == " + " else -im)
    return complex(float(t), 0.0)


def read_vec(src) -> tuple[np.ndarray, int]:
    """(values, number of ranks) of a VecView dump: 'Vec Object: <n> MPI
    processes', the type line, then one value per line, in 'Process [k]'
    blocks for an MPI vector.  Values are float64, or complex128 when any has
    an imaginary part."""
    lines = _text(src).splitlines()
    if not lines or not lines[0].startswith("Vec Object:"):
        raise ValueError("not a PETSc VecView dump (no 'Vec Object:' header)")
    nprocs = int(lines[0].split(":", 1)[1].split()[0])
    vals = [parse_scalar(ln) for ln in lines[1:]
            if ln.strip() and not ln.startswith(("  type", "Process ["))]
    arr = np.array(vals, dtype=np.complex128)
    if not np.any(arr.imag):
        arr = arr.real.copy()
    return arr, nprocs


def read_mat(src) -> tuple[np.ndarray, np.ndarray, np.ndarray, int]:
    """(row pointer, columns, values, number of ranks) of a MatView dump of
    an AIJ matrix: 'row i: (c, v)  (c, v) ...' per row, entries as stored."""
    lines = _text(src).splitlines()
    if not lines or not lines[0].startswith("Mat Object:"):
        raise ValueError("not a PETSc MatView dump (no 'Mat Object:' header)")
    nprocs = int(lines[0].split(":", 1)[1].split()[0])
    rp, cols, vals = [0], [], []
    for ln in lines[1:]:
        if not ln.startswith("row "):
            continue
        head, body = ln.split(":", 1)
        if int(head.split()[1]) != len(rp) - 1:
            raise ValueError(f"rows out of order at '{head}'")
        for part in body.split(")"):
            part = part.strip()
            if not part:
                continue
            c, v = part.lstrip("(").split(",", 1)
            cols.append(int(c))
            vals.append(parse_scalar(v))
        rp.append(len(cols))
    v = np.array(vals, dtype=np.complex128)
    if not np.any(v.imag):
        v = v.real.copy()
    return np.array(rp, np.int64), np.array(cols, np.int64), v, nprocs


def vec_load(src, device="cuda"):
    """A VecView dump (text or path) as a vector on the device: float64 (or
    complex128) CUDA tensor in basis-rank order, ready for Hamiltonian.matvec,
    the solver's start or a correlator.  Returns (tensor, number of ranks of
    the dump)."""
    import torch

    v, nprocs = read_vec(src)
    return torch.from_numpy(v).to(device), nprocs


def mat_load(src, device="cuda"):
    """A MatView dump (text or path) as a CSR matrix on the device
    (torch.sparse_csr_tensor, float64 or complex128, int64 indices) and the
    number of ranks of the dump."""
    import torch

    rp, cols, vals, nprocs = read_mat(src)
    D = len(rp) - 1
    A = torch.sparse_csr_tensor(torch.from_numpy(rp), torch.from_numpy(cols), torch.from_numpy(vals), size=(D, D))
    return A.to(device), nprocs


def mat_view_csr(rp, cols, vals, nprocs: int = 1) -> str:
    """MatView text of CSR arrays (the inverse of read_mat for real values)."""
    out = [f"Mat Object: {_procs(nprocs)}", "  type: mpiaij"]
    for r in range(len(rp) - 1):
        out.append(f"row {r}:" + "".join(f" ({int(cols[q])}, {petsc_g(float(vals[q]))}) "
                                         for q in range(rp[r], rp[r + 1])))
    return "\n".join(out) + "\n"
