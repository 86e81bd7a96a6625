/*
 * hamiltonian_shell.hpp -- LSQuantumED's Hamiltonian<L>::hamilt as a PETSc
 * MATSHELL over liblsqed.so (include/lsqed.h), for the reference's own build
 * (PETSc configured with HIP; complex or real scalars).
 *
 * The reference assembles a MATMPIAIJ in Hamiltonian<L>::MatInit
 * (include/hamiltonian.tmpl_func.hpp:8-56) and fills it in build_diag /
 * build_off_diag / build_disorder (tmpl:97-344); SLEPc receives it in
 * Solver::SolverInit (src/solver.cpp:16).  Here the matrix is a shell whose
 *   MATOP_MULT (and MULT_TRANSPOSE / MULT_HERMITIAN_TRANSPOSE: H is real
 *   symmetric)  -> lsqed_hv_c128 (complex PetscScalar) or lsqed_hv (real)
 *   MATOP_GET_DIAGONAL                     -> lsqed_diag_c128 / lsqed_diag
 * on the device arrays of VECHIP vectors (VecHIPGetArrayRead/Write): no host
 * copy and no allocation per product.  MatCreateVecs on the shell gives
 * VECHIP vectors (MatShellSetVecType), so SLEPc's basis lives on the device.
 * The build_* calls become HamiltonianShellBuild / HamiltonianShellSetFields,
 * which only record flags and upload the N fields.
 *
 * One MPI rank drives one GPU (the handle holds the whole sector).  Kernels
 * run on the stream set with HamiltonianShellSetStream (default: the legacy
 * default stream, which orders against PETSc's blocking streams).
 * MatGetInfo (examples/main.cpp:114) is not defined for shell matrices.
 */
#ifndef LSQED_HAMILTONIAN_SHELL_HPP
#define LSQED_HAMILTONIAN_SHELL_HPP

#include <petscmat.h>

#include <vector>

#include "lsqed.h"

namespace lsqed_petsc {

struct ShellCtx {
  lsqed_ham* h = nullptr;
  lsqed_stream_t stream = nullptr;
  PetscInt n = 0;
};

inline PetscErrorCode ShellContext(Mat A, ShellCtx** c) {
  PetscFunctionBeginUser;
  PetscCall(MatShellGetContext(A, (void*)c));
  PetscCheck(*c && (*c)->h, PETSC_COMM_SELF, PETSC_ERR_ARG_WRONGSTATE, "not an lsqed Hamiltonian shell");
  PetscFunctionReturn(PETSC_SUCCESS);
}

inline PetscErrorCode CheckDeviceVec(Vec v, PetscInt n) {
  PetscBool hip = PETSC_FALSE;
  PetscInt len = 0;
  PetscFunctionBeginUser;
  PetscCall(PetscObjectTypeCompareAny((PetscObject)v, &hip, VECSEQHIP, VECMPIHIP, VECHIP, ""));
  PetscCheck(hip, PETSC_COMM_SELF, PETSC_ERR_ARG_WRONG,
             "the lsqed shell multiplies VECHIP vectors (create them with MatCreateVecs)");
  PetscCall(VecGetLocalSize(v, &len));
  PetscCheck(len == n, PETSC_COMM_SELF, PETSC_ERR_ARG_SIZ, "vector length %d, sector size %d", (int)len, (int)n);
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* y = H x (MatMult of EPSSolve, src/solver.cpp:45). */
inline PetscErrorCode ShellMult(Mat A, Vec x, Vec y) {
  ShellCtx* c = nullptr;
  const PetscScalar* xa = nullptr;
  PetscScalar* ya = nullptr;
  PetscFunctionBeginUser;
  PetscCall(ShellContext(A, &c));
  PetscCall(CheckDeviceVec(x, c->n));
  PetscCall(CheckDeviceVec(y, c->n));
  PetscCheck(x != y, PETSC_COMM_SELF, PETSC_ERR_ARG_IDN, "x and y must be different vectors");
  PetscCall(VecHIPGetArrayRead(x, &xa));
  PetscCall(VecHIPGetArrayWrite(y, &ya));
#if defined(PETSC_USE_COMPLEX)
  const int rc = lsqed_hv_c128(c->h, reinterpret_cast<const double*>(xa), reinterpret_cast<double*>(ya), c->stream);
#else
  const int rc = lsqed_hv(c->h, xa, ya, c->stream);
#endif
  PetscCall(VecHIPRestoreArrayWrite(y, &ya));
  PetscCall(VecHIPRestoreArrayRead(x, &xa));
  PetscCheck(rc == LSQED_OK, PETSC_COMM_SELF, PETSC_ERR_LIB, "lsqed H.v: %s", lsqed_last_error());
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* MatGetDiagonal (tests/unit_tests/build_diag/chain_1d/test_build_diag60.cpp:29). */
inline PetscErrorCode ShellGetDiagonal(Mat A, Vec d) {
  ShellCtx* c = nullptr;
  PetscScalar* da = nullptr;
  PetscFunctionBeginUser;
  PetscCall(ShellContext(A, &c));
  PetscCall(CheckDeviceVec(d, c->n));
  PetscCall(VecHIPGetArrayWrite(d, &da));
#if defined(PETSC_USE_COMPLEX)
  const int rc = lsqed_diag_c128(c->h, reinterpret_cast<double*>(da), c->stream);
#else
  const int rc = lsqed_diag(c->h, da, c->stream);
#endif
  PetscCall(VecHIPRestoreArrayWrite(d, &da));
  PetscCheck(rc == LSQED_OK, PETSC_COMM_SELF, PETSC_ERR_LIB, "lsqed diagonal: %s", lsqed_last_error());
  PetscFunctionReturn(PETSC_SUCCESS);
}

inline PetscErrorCode ShellDestroy(Mat A) {
  ShellCtx* c = nullptr;
  PetscFunctionBeginUser;
  PetscCall(MatShellGetContext(A, (void*)&c));
  if (c) {
    lsqed_ham_destroy(c->h);
    delete c;
  }
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* The shell for a model given as bond lists (lsqed_params).  Nothing is
 * built yet, as after MatInit: call HamiltonianShellBuild for build_diag /
 * build_off_diag and HamiltonianShellSetFields for build_disorder. */
inline PetscErrorCode HamiltonianShellCreate(MPI_Comm comm, const lsqed_params* p, Mat* A) {
  PetscMPIInt size = 0;
  PetscFunctionBeginUser;
  PetscCallMPI(MPI_Comm_size(comm, &size));
  PetscCheck(size == 1, comm, PETSC_ERR_SUP, "the lsqed shell runs on one MPI rank per GPU (sector not split)");
  ShellCtx* c = new ShellCtx;
  if (lsqed_ham_create(p, &c->h) != LSQED_OK) {
    delete c;
    SETERRQ(comm, PETSC_ERR_LIB, "lsqed_ham_create: %s", lsqed_last_error());
  }
  c->n = (PetscInt)lsqed_ham_dim(c->h);
  if (lsqed_ham_set_build(c->h, 0) != LSQED_OK) {
    lsqed_ham_destroy(c->h);
    delete c;
    SETERRQ(comm, PETSC_ERR_LIB, "lsqed_ham_set_build: %s", lsqed_last_error());
  }
  PetscCall(MatCreateShell(comm, c->n, c->n, c->n, c->n, c, A));
  PetscCall(MatShellSetOperation(*A, MATOP_MULT, (void (*)(void))ShellMult));
  PetscCall(MatShellSetOperation(*A, MATOP_MULT_TRANSPOSE, (void (*)(void))ShellMult));
  PetscCall(MatShellSetOperation(*A, MATOP_MULT_HERMITIAN_TRANSPOSE, (void (*)(void))ShellMult));
  PetscCall(MatShellSetOperation(*A, MATOP_GET_DIAGONAL, (void (*)(void))ShellGetDiagonal));
  PetscCall(MatShellSetOperation(*A, MATOP_DESTROY, (void (*)(void))ShellDestroy));
  PetscCall(MatShellSetVecType(*A, VECHIP));
  PetscCall(MatSetOption(*A, MAT_SYMMETRIC, PETSC_TRUE));
  PetscCall(MatSetOption(*A, MAT_HERMITIAN, PETSC_TRUE));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* The shell of Hamiltonian<L>(env, basis, lattice, J1, D1, J2, D2)
 * (tmpl:69-91): the bond lists in the order HamiltHelper::get_coup_elems
 * visits them (tmpl:350-387: site i ascending, then its neighbour slots
 * 1..num_nnXsite(i) of lattice.nn / lattice.nnn with stride num_nn() + 1). */
template <typename L>
PetscErrorCode HamiltonianShellFromLattice(MPI_Comm comm, PetscInt nspins, PetscInt nup, L& lattice, PetscReal J1,
                                           PetscReal D1, PetscReal J2, PetscReal D2, Mat* A) {
  std::vector<int> ni, nj, mi, mj;
  PetscFunctionBeginUser;
  for (PetscInt i = 0; i < nspins; ++i) {
    for (PetscInt k = 1; k <= lattice.num_nnXsite(i); ++k) {
      ni.push_back((int)i);
      nj.push_back((int)lattice.nn[(lattice.num_nn() + 1) * i + k]);
    }
    for (PetscInt k = 1; k <= lattice.num_nnnXsite(i); ++k) {
      mi.push_back((int)i);
      mj.push_back((int)lattice.nnn[(lattice.num_nnn() + 1) * i + k]);
    }
  }
  lsqed_params p;
  p.nspins = (int)nspins;
  p.nup = (int)nup;
  p.n_nn = (int)ni.size();
  p.n_nnn = (int)mi.size();
  p.nn_i = ni.data();
  p.nn_j = nj.data();
  p.nnn_i = mi.data();
  p.nnn_j = mj.data();
  p.J1 = J1;
  p.D1 = D1;
  p.J2 = J2;
  p.D2 = D2;
  PetscCall(HamiltonianShellCreate(comm, &p, A));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* build_diag (LSQED_BUILD_DIAG) / build_off_diag (LSQED_BUILD_OFFDIAG):
 * adds the part to the operator. */
inline PetscErrorCode HamiltonianShellBuild(Mat A, int part) {
  ShellCtx* c = nullptr;
  int flags = 0;
  PetscFunctionBeginUser;
  PetscCall(ShellContext(A, &c));
  PetscCheck(lsqed_ham_get_params(c->h, nullptr, nullptr, nullptr, nullptr, &flags, nullptr, nullptr) == LSQED_OK &&
                 lsqed_ham_set_build(c->h, flags | part) == LSQED_OK,
             PETSC_COMM_SELF, PETSC_ERR_ARG_WRONG, "lsqed build: %s", lsqed_last_error());
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* build_disorder(env, hi) (tmpl:97-125): the N fields (copied) and the
 * diagonal rewritten with them. */
inline PetscErrorCode HamiltonianShellSetFields(Mat A, const PetscReal* hi) {
  ShellCtx* c = nullptr;
  PetscFunctionBeginUser;
  PetscCall(ShellContext(A, &c));
  PetscCheck(lsqed_ham_set_fields(c->h, hi) == LSQED_OK, PETSC_COMM_SELF, PETSC_ERR_LIB, "lsqed fields: %s",
             lsqed_last_error());
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* The HIP stream the shell's kernels go to (NULL: the legacy default). */
inline PetscErrorCode HamiltonianShellSetStream(Mat A, lsqed_stream_t stream) {
  ShellCtx* c = nullptr;
  PetscFunctionBeginUser;
  PetscCall(ShellContext(A, &c));
  c->stream = stream;
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* The library handle (for lsqed_eps_* / lsqed_obs_* / lsqed_batch_*). */
inline PetscErrorCode HamiltonianShellHandle(Mat A, lsqed_ham** h) {
  ShellCtx* c = nullptr;
  PetscFunctionBeginUser;
  PetscCall(ShellContext(A, &c));
  *h = c->h;
  PetscFunctionReturn(PETSC_SUCCESS);
}

}  // namespace lsqed_petsc

#endif /* LSQED_HAMILTONIAN_SHELL_HPP */
