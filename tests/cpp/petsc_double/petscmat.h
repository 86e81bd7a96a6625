/*
 * TEST DOUBLE, not PETSc.  The handful of PETSc declarations and behaviours
 * that lsquantumed_amd/petsc/hamiltonian_shell.hpp uses (and that its test,
 * tests/cpp/test_shell.cpp, calls), written here so that the adapter compiles
 * and runs in an image without PETSc: a shell Mat holding a context and an
 * operation table, and Vecs whose HIP types keep their entries in device
 * memory (VecHIPGetArray* hand out the device pointer, as PETSc's VECHIP
 * does).  Names, argument order and the PetscCall / PetscCheck error
 * convention follow PETSc's public API; nothing else of PETSc is modelled.
 * -DPETSC_USE_COMPLEX selects complex scalars, as a complex PETSc build does.
 */
#ifndef LSQED_TEST_PETSC_DOUBLE_H
#define LSQED_TEST_PETSC_DOUBLE_H

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

typedef int PetscErrorCode;
typedef int64_t PetscInt;  /* --with-64-bit-indices, as the reference requires */
typedef int PetscMPIInt;
typedef double PetscReal;
#if defined(PETSC_USE_COMPLEX)
typedef std::complex<double> PetscScalar;
#else
typedef double PetscScalar;
#endif
typedef enum { PETSC_FALSE = 0, PETSC_TRUE = 1 } PetscBool;
typedef int MPI_Comm;
typedef const char* VecType;

#define PETSC_SUCCESS 0
#define PETSC_COMM_SELF 1
#define PETSC_COMM_WORLD 2
enum {
  PETSC_ERR_SUP = 56,
  PETSC_ERR_ARG_WRONGSTATE = 73,
  PETSC_ERR_ARG_WRONG = 62,
  PETSC_ERR_ARG_SIZ = 60,
  PETSC_ERR_ARG_IDN = 61,
  PETSC_ERR_LIB = 76
};
#define VECSEQ "seq"
#define VECSEQHIP "seqhip"
#define VECMPIHIP "mpihip"
#define VECHIP "hip"

typedef enum {
  MATOP_MULT,
  MATOP_MULT_TRANSPOSE,
  MATOP_MULT_HERMITIAN_TRANSPOSE,
  MATOP_GET_DIAGONAL,
  MATOP_DESTROY,
  MATOP_COUNT_
} MatOperation;
typedef enum { MAT_SYMMETRIC, MAT_HERMITIAN } MatOption;

struct _p_PetscObject {
  std::string type;
};
typedef _p_PetscObject* PetscObject;

struct _p_Vec : _p_PetscObject {
  PetscInt n = 0;
  PetscScalar* a = nullptr;  /* device memory for the HIP types, host memory for VECSEQ */
};
typedef _p_Vec* Vec;

struct _p_Mat : _p_PetscObject {
  PetscInt m = 0, n = 0;
  void* ctx = nullptr;
  void (*ops[MATOP_COUNT_])(void) = {};
  std::string vectype = VECSEQ;
  PetscBool sym = PETSC_FALSE;
};
typedef _p_Mat* Mat;

/* The last error message (PETSc prints its traceback instead). */
inline std::string& PetscDoubleLastError() {
  static std::string s;
  return s;
}
inline PetscErrorCode PetscDoubleError(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  PetscDoubleLastError() = buf;
  return code;
}

#define PetscFunctionBeginUser do { } while (0)
#define PetscFunctionReturn(x) return (x)
#define PetscCall(e)                  \
  do {                                \
    PetscErrorCode _ierr = (e);       \
    if (_ierr) return _ierr;          \
  } while (0)
#define PetscCallMPI(e) PetscCall(e)
#define SETERRQ(comm, code, ...) return PetscDoubleError((code), __VA_ARGS__)
#define PetscCheck(cond, comm, code, ...) \
  do {                                    \
    if (!(cond)) SETERRQ(comm, code, __VA_ARGS__); \
  } while (0)

inline PetscErrorCode MPI_Comm_size(MPI_Comm, PetscMPIInt* size) {
  *size = 1;
  return 0;
}

inline PetscErrorCode PetscObjectTypeCompareAny(PetscObject obj, PetscBool* same, const char* t, ...) {
  va_list ap;
  va_start(ap, t);
  *same = PETSC_FALSE;
  for (const char* s = t; s && s[0]; s = va_arg(ap, const char*))
    if (obj->type == s) *same = PETSC_TRUE;
  va_end(ap);
  return 0;
}

/* ---- Vec ---- */
inline PetscErrorCode VecCreateType(PetscInt n, VecType type, Vec* v) {
  Vec x = new _p_Vec;
  x->type = type;
  x->n = n;
  if (x->type == VECSEQ) {
    x->a = new PetscScalar[n > 0 ? n : 1]();
  } else if (hipMalloc((void**)&x->a, sizeof(PetscScalar) * (n > 0 ? n : 1)) != hipSuccess) {
    delete x;
    return PetscDoubleError(PETSC_ERR_LIB, "hipMalloc");
  }
  *v = x;
  return 0;
}
inline PetscErrorCode VecDestroy(Vec* v) {
  if (*v) {
    if ((*v)->type == VECSEQ)
      delete[] (*v)->a;
    else
      (void)hipFree((*v)->a);
    delete *v;
    *v = nullptr;
  }
  return 0;
}
inline PetscErrorCode VecGetLocalSize(Vec v, PetscInt* n) {
  *n = v->n;
  return 0;
}
inline PetscErrorCode VecHIPGetArrayRead(Vec v, const PetscScalar** a) {
  if (v->type == VECSEQ) return PetscDoubleError(PETSC_ERR_ARG_WRONG, "not a HIP vector");
  *a = v->a;
  return 0;
}
inline PetscErrorCode VecHIPRestoreArrayRead(Vec, const PetscScalar** a) {
  *a = nullptr;
  return 0;
}
inline PetscErrorCode VecHIPGetArrayWrite(Vec v, PetscScalar** a) {
  if (v->type == VECSEQ) return PetscDoubleError(PETSC_ERR_ARG_WRONG, "not a HIP vector");
  *a = v->a;
  return 0;
}
inline PetscErrorCode VecHIPRestoreArrayWrite(Vec, PetscScalar** a) {
  *a = nullptr;
  return 0;
}

/* ---- shell Mat ---- */
inline PetscErrorCode MatCreateShell(MPI_Comm, PetscInt m, PetscInt n, PetscInt, PetscInt, void* ctx, Mat* A) {
  Mat M = new _p_Mat;
  M->type = "shell";
  M->m = m;
  M->n = n;
  M->ctx = ctx;
  *A = M;
  return 0;
}
inline PetscErrorCode MatShellGetContext(Mat A, void* ctx) {
  *(void**)ctx = A->ctx;
  return 0;
}
inline PetscErrorCode MatShellSetOperation(Mat A, MatOperation op, void (*f)(void)) {
  A->ops[op] = f;
  return 0;
}
inline PetscErrorCode MatShellSetVecType(Mat A, VecType t) {
  A->vectype = t;
  return 0;
}
inline PetscErrorCode MatSetOption(Mat A, MatOption o, PetscBool flg) {
  if (o == MAT_SYMMETRIC) A->sym = flg;
  return 0;
}
inline PetscErrorCode MatCreateVecs(Mat A, Vec* right, Vec* left) {
  if (right) PetscCall(VecCreateType(A->n, A->vectype.c_str(), right));
  if (left) PetscCall(VecCreateType(A->m, A->vectype.c_str(), left));
  return 0;
}
inline PetscErrorCode MatMult(Mat A, Vec x, Vec y) {
  if (!A->ops[MATOP_MULT]) return PetscDoubleError(PETSC_ERR_SUP, "no MATOP_MULT");
  return ((PetscErrorCode(*)(Mat, Vec, Vec))A->ops[MATOP_MULT])(A, x, y);
}
inline PetscErrorCode MatMultTranspose(Mat A, Vec x, Vec y) {
  if (!A->ops[MATOP_MULT_TRANSPOSE]) return PetscDoubleError(PETSC_ERR_SUP, "no MATOP_MULT_TRANSPOSE");
  return ((PetscErrorCode(*)(Mat, Vec, Vec))A->ops[MATOP_MULT_TRANSPOSE])(A, x, y);
}
inline PetscErrorCode MatGetDiagonal(Mat A, Vec d) {
  if (!A->ops[MATOP_GET_DIAGONAL]) return PetscDoubleError(PETSC_ERR_SUP, "no MATOP_GET_DIAGONAL");
  return ((PetscErrorCode(*)(Mat, Vec))A->ops[MATOP_GET_DIAGONAL])(A, d);
}
inline PetscErrorCode MatDestroy(Mat* A) {
  if (*A) {
    if ((*A)->ops[MATOP_DESTROY]) PetscCall(((PetscErrorCode(*)(Mat))(*A)->ops[MATOP_DESTROY])(*A));
    delete *A;
    *A = nullptr;
  }
  return 0;
}

#endif
