/*
 * lsquantumed.hpp -- C++ host interface of LSQuantumED over liblsqed.so.
 *
 * Classes and free functions keep the reference's names, argument meaning
 * and error behaviour, so code written against LSQuantumED's headers reads
 * the same:
 *   Environment                 include/environment.hpp, src/environment.cpp:8-49
 *   Basis                       include/basis.hpp, src/basis.cpp:56-94
 *   chain1D/square2D/honeycomb2D include/lattice.hpp, src/lattice.cpp
 *   AF, Striped                 include/sublattice.hpp, src/sublattice.cpp
 *   Hamiltonian<L>              include/hamiltonian.hpp, include/hamiltonian.tmpl_func.hpp
 *   Solver::                    include/solver.hpp, src/solver.cpp
 *   SzSz, SpSm, SmSp, SiSj      include/spinOperators/correlation.hpp
 *   Phys::                      include/phys.hpp, include/phys.tmpl_func.hpp
 *   Tools::RandomDisorder       include/tools/random_disorder.hpp
 *
 * The PETSc/SLEPc objects become handles of the library: Mat is the
 * matrix-free operator (lsqed_ham*), Vec a vector in device memory, EPS the
 * device eigensolver.  The PETSc calls the reference's tests and drivers make
 * on them (MatCreateVecs, MatMult, VecDestroy, EPSSetDimensions,
 * EPSGetEigenpair, ...) keep PETSc's argument order.  What differs by design:
 * one process drives one GPU (mpi_size 1, one block of rows), Basis keeps no
 * list of states (they are unranked on the device), the Hamiltonian is never
 * assembled, and every call runs on the legacy default HIP stream.
 *
 * Error behaviour: constructors throw std::invalid_argument with the
 * reference's messages; functions that return PetscErrorCode return 0 or the
 * library's negative LSQED_ERR_* code (lsqed_last_error() has the message).
 *
 * Header-only.  Link with liblsqed.so and libamdhip64.
 */
#ifndef LSQUANTUMED_HPP
#define LSQUANTUMED_HPP

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "lsqed.h"

namespace lsqed {

using PetscInt = int64_t;
using PetscReal = double;
using PetscScalar = double;
using PetscErrorCode = int;
using PetscBool = bool;
constexpr PetscBool PETSC_TRUE = true;
constexpr PetscBool PETSC_FALSE = false;
constexpr PetscInt PETSC_DECIDE = -1;
constexpr PetscInt PETSC_DEFAULT = -2;
inline PetscReal PetscRealPart(PetscScalar x) { return x; }
inline PetscReal PetscImaginaryPart(PetscScalar) { return 0.0; }

namespace detail {
inline uint64_t next_state() {
  static std::atomic<uint64_t> counter{0};
  return ++counter;
}
inline void check(int rc) {
  if (rc != LSQED_OK) throw std::runtime_error(lsqed_last_error());
}
inline PetscErrorCode hip_status(hipError_t e) { return e == hipSuccess ? 0 : LSQED_ERR_HIP; }
}  // namespace detail

/* ------------------------------------------------------------ Environment */
/* Environment(argc, argv[, nspins][, help]): the spin count is validated as in
 * src/environment.cpp:31-35 (odd first, then non-positive).  SLEPc/MPI start-up
 * has no counterpart: one process, rank 0 of 1. */
class Environment {
 public:
  PetscInt nspins = 0;
  int mpi_size = 1;
  int mpi_rank = 0;
  PetscBool disorder_flg = PETSC_FALSE;

  Environment(int /*argc*/, char** /*argv*/, const char* /*help*/ = nullptr) {}
  Environment(int argc, char** argv, PetscInt m_nspins, const char* help = nullptr) : Environment(argc, argv, help) {
    if (m_nspins % 2 != 0) throw std::invalid_argument("The total number of spins has to be EVEN.");
    if (m_nspins <= 0) throw std::invalid_argument("The total number of spins has to be POSITIVE.");
    nspins = m_nspins;
  }
  explicit Environment(PetscInt m_nspins) : Environment(0, nullptr, m_nspins) {}
};

/* ------------------------------------------------------------------ Basis */
/* The fixed-magnetisation sector: nspins_up = total_mag + nspins/2 up spins,
 * size C(nspins, nspins_up) (src/basis.cpp:38-46), row block of this rank
 * (src/basis.cpp:78-80).  States are colex-ranked bit strings; the list
 * itself stays on the device side. */
class Basis {
 public:
  PetscInt nspins;
  int mpi_size;
  int mpi_rank;
  PetscInt max_mag = 0;
  PetscInt total_mag = 0;
  PetscInt nspins_up = 0;
  PetscInt size = 0;
  PetscInt local_size = 0;
  PetscInt global_start_index = 0;

  Basis(Environment& env, PetscInt m_total_mag) : nspins{env.nspins}, mpi_size{env.mpi_size}, mpi_rank{env.mpi_rank} {
    if (m_total_mag > nspins / 2 || m_total_mag < -nspins / 2)
      throw std::invalid_argument("The required constant magnetization has to be <= nspins/2.");
    max_mag = nspins / 2;
    total_mag = m_total_mag;
    nspins_up = total_mag + nspins / 2;
    size = compute_size();
    if (size < mpi_size)
      throw std::invalid_argument("The size of the basis has to be at least equal to the number of MPI processes used.");
    const PetscInt rest = size % mpi_size;
    local_size = size / mpi_size + (rest > mpi_rank ? 1 : 0);
    global_start_index = mpi_rank * local_size + (rest <= mpi_rank ? rest : 0);
  }

  PetscInt compute_size() const { return lsqed_sector_dim(int(nspins), int(nspins_up)); }

  /* Operator handle of this sector without bonds: what the correlation
   * functions need (the rank tables of the sector).  Created on first use and
   * shared by copies of this Basis. */
  lsqed_ham* sector_handle() {
    if (!sector_) {
      lsqed_params p{};
      p.nspins = int(nspins);
      p.nup = int(nspins_up);
      lsqed_ham* h = nullptr;
      detail::check(lsqed_ham_create(&p, &h));
      sector_.reset(h, [](lsqed_ham* x) { lsqed_ham_destroy(x); });
    }
    return sector_.get();
  }

 private:
  std::shared_ptr<lsqed_ham> sector_;
};

/* ---------------------------------------------------------------- Lattices */
enum class lattice_type { chain1D, square2D, honeycomb2D };

/* Lattice (include/lattice.hpp:95-126).  The forward bonds of each kind are
 * kept in the order the Hamiltonian visits them, site by site
 * (HamiltHelper::get_coup_elems, include/hamiltonian.tmpl_func.hpp:350-387);
 * num_nnXsite(i) is the number of them that start at site i. */
class Lattice {
 public:
  explicit Lattice(Environment& env) : nspins{env.nspins} {}
  virtual ~Lattice() {}
  virtual lattice_type get_type() const = 0;
  virtual PetscInt num_nn() = 0;
  virtual PetscInt num_nnn() = 0;
  virtual PetscInt num_nnXsite(PetscInt lat_site) { return nnXsite.at(size_t(lat_site)); }
  virtual PetscInt num_nnnXsite(PetscInt lat_site) { return nnnXsite.at(size_t(lat_site)); }
  virtual PetscInt num_spins() { return nspins; }

  PetscInt nspins;
  std::vector<int> nn_i, nn_j;    // nearest-neighbour bonds (i, j)
  std::vector<int> nnn_i, nnn_j;  // next-nearest-neighbour bonds

 protected:
  void load_bonds(int kind_id, PetscInt lx, PetscInt ly) {
    for (int kind = 0; kind < 2; ++kind) {
      std::vector<int>& bi = kind ? nnn_i : nn_i;
      std::vector<int>& bj = kind ? nnn_j : nn_j;
      std::vector<PetscInt>& per_site = kind ? nnnXsite : nnXsite;
      int n = 0;
      detail::check(lsqed_lattice_bonds(kind_id, int(nspins), int(lx), int(ly), kind, nullptr, nullptr, &n));
      bi.resize(size_t(n));
      bj.resize(size_t(n));
      detail::check(lsqed_lattice_bonds(kind_id, int(nspins), int(lx), int(ly), kind, bi.data(), bj.data(), &n));
      per_site.assign(size_t(nspins), 0);
      for (int b = 0; b < n; ++b) ++per_site[size_t(bi[size_t(b)])];
    }
  }
  std::vector<PetscInt> nnXsite, nnnXsite;
};

/* Periodic chain: nn (i, i+1), nnn (i, i+2) (src/lattice.cpp:7-53). */
class chain1D : public Lattice {
 public:
  explicit chain1D(Environment& env) : Lattice{env} { load_bonds(LSQED_CHAIN1D, 0, 0); }
  lattice_type get_type() const override { return lattice_type::chain1D; }
  PetscInt num_nn() override { return 1; }
  PetscInt num_nnn() override { return 1; }
};

/* Periodic lx x ly square lattice, site ix*ly + iy (src/lattice.cpp:59-136). */
class square2D : public Lattice {
 public:
  PetscInt nspins_x, nspins_y;
  square2D(Environment& env, PetscInt m_nspins_x, PetscInt m_nspins_y)
      : Lattice{env}, nspins_x{m_nspins_x}, nspins_y{m_nspins_y} {
    if (m_nspins_x * m_nspins_y != env.nspins)
      throw std::invalid_argument(
          "The product between the number of spins in x direction and the numer of spins in the y direction has to "
          "be equal to the total number of spins.");
    load_bonds(LSQED_SQUARE2D, nspins_x, nspins_y);
  }
  lattice_type get_type() const override { return lattice_type::square2D; }
  PetscInt num_nn() override { return 2; }
  PetscInt num_nnn() override { return 2; }
};

/* Periodic honeycomb lattice on an lx x ly brick (src/lattice.cpp:142-264). */
class honeycomb2D : public Lattice {
 public:
  PetscInt nspins_x, nspins_y;
  honeycomb2D(Environment& env, PetscInt m_nspins_x, PetscInt m_nspins_y)
      : Lattice{env}, nspins_x{m_nspins_x}, nspins_y{m_nspins_y} {
    if (m_nspins_x * m_nspins_y != env.nspins)
      throw std::invalid_argument(
          "The product between the number of spins in x direction and the numer of spins in the y direction has to "
          "be equal to the total number of spins.");
    if (m_nspins_x % 2 != 0 || m_nspins_y % 2 != 0)
      throw std::invalid_argument("The number of spins in each direction has to be even.");
    load_bonds(LSQED_HONEYCOMB2D, nspins_x, nspins_y);
  }
  lattice_type get_type() const override { return lattice_type::honeycomb2D; }
  PetscInt num_nn() override { return 2; }
  PetscInt num_nnn() override { return 3; }
};

/* ------------------------------------------------------------ Sublattices */
template <typename L>
class Sublattice {
 public:
  explicit Sublattice(L& m_lattice) : lattice{m_lattice} {}
  virtual ~Sublattice() {}
  virtual PetscInt get_size() const = 0;
  virtual const std::vector<PetscInt>& get_sl(PetscInt sublat_i) = 0;

 protected:
  L& lattice;
};

/* Two antiferromagnetic sublattices (src/sublattice.cpp:5-57, :95-125): the
 * parity of i on the chain, of ix + iy on the 2D lattices. */
template <typename L>
class AF : public Sublattice<L> {
 public:
  PetscInt nspins;
  PetscInt size = 2;
  AF(L& m_lattice, PetscInt /*m_size*/ = 2) : Sublattice<L>{m_lattice}, nspins{m_lattice.num_spins()} {
    sublat.assign(size_t(size), {});
    if constexpr (std::is_same<L, chain1D>::value) {
      for (PetscInt i = 0; i < nspins; ++i) sublat[size_t(i & 1)].push_back(i);
    } else {
      for (PetscInt ix = 0; ix < m_lattice.nspins_x; ++ix)
        for (PetscInt iy = 0; iy < m_lattice.nspins_y; ++iy)
          sublat[size_t((ix + iy) & 1)].push_back(ix * m_lattice.nspins_y + iy);
    }
  }
  PetscInt get_size() const override { return size; }
  const std::vector<PetscInt>& get_sl(PetscInt sublat_i) override { return sublat.at(size_t(sublat_i)); }

 private:
  std::vector<std::vector<PetscInt>> sublat;
};

/* Striped sublattices (src/sublattice.cpp:61-91, :128-162).  Square: 0/1 by
 * the parity of iy, 2/3 by the parity of ix.  Honeycomb: six sublattices of
 * which the reference fills only 0/1 (parity of iy); 2..5 stay empty there
 * too. */
template <typename L>
class Striped : public Sublattice<L> {
  static_assert(!std::is_same<L, chain1D>::value, "Striped sublattices are defined on 2D lattices");

 public:
  PetscInt nspins;
  PetscInt size = std::is_same<L, honeycomb2D>::value ? 6 : 4;
  Striped(L& m_lattice, PetscInt /*m_size*/ = 4) : Sublattice<L>{m_lattice}, nspins{m_lattice.num_spins()} {
    sublat.assign(size_t(size), {});
    for (PetscInt ix = 0; ix < m_lattice.nspins_x; ++ix)
      for (PetscInt iy = 0; iy < m_lattice.nspins_y; ++iy) {
        const PetscInt site = ix * m_lattice.nspins_y + iy;
        sublat[size_t(iy & 1)].push_back(site);
        if (std::is_same<L, square2D>::value) sublat[size_t(2 + (ix & 1))].push_back(site);
      }
  }
  PetscInt get_size() const override { return size; }
  const std::vector<PetscInt>& get_sl(PetscInt sublat_i) override { return sublat.at(size_t(sublat_i)); }

 private:
  std::vector<std::vector<PetscInt>> sublat;
};

/* ------------------------------------------------------------ Mat and Vec */
using Mat = lsqed_ham*;

/* A vector of the sector in device memory.  `state` marks modifications
 * (PETSc's object state): every call here that writes a Vec gives it a new
 * value, unique across all vectors of the process (so a vector created where
 * a destroyed one lived never repeats its marks), and code that writes
 * through VecGetArray must call VecRestoreArray. */
struct _p_Vec {
  double* array = nullptr;
  PetscInt n = 0;
  uint64_t state = 0;
};
using Vec = _p_Vec*;

inline PetscErrorCode VecCreateDevice(PetscInt n, Vec* v) {
  if (!v || n < 0) return LSQED_ERR_ARG;
  *v = nullptr;
  double* a = nullptr;
  if (n > 0 && hipMalloc(reinterpret_cast<void**>(&a), size_t(n) * sizeof(double)) != hipSuccess) return LSQED_ERR_OOM;
  Vec x = new _p_Vec;
  x->array = a;
  x->n = n;
  x->state = detail::next_state();
  *v = x;
  return 0;
}

/* MatCreateVecs(mat, &right, &left): vectors of the operator's size; either
 * pointer may be NULL. */
inline PetscErrorCode MatCreateVecs(Mat mat, Vec* right, Vec* left) {
  if (!mat) return LSQED_ERR_ARG;
  const PetscInt D = lsqed_ham_dim(mat);
  if (right) {
    if (PetscErrorCode e = VecCreateDevice(D, right)) return e;
  }
  if (left) {
    if (PetscErrorCode e = VecCreateDevice(D, left)) return e;
  }
  return 0;
}

inline PetscErrorCode VecDestroy(Vec* v) {
  if (!v || !*v) return 0;
  const hipError_t e = (*v)->array ? hipFree((*v)->array) : hipSuccess;
  delete *v;
  *v = nullptr;
  return detail::hip_status(e);
}

inline PetscErrorCode VecGetSize(Vec v, PetscInt* n) {
  if (!v || !n) return LSQED_ERR_ARG;
  *n = v->n;
  return 0;
}

/* Device pointer of the entries; VecRestoreArray marks the vector modified. */
inline PetscErrorCode VecGetArray(Vec v, PetscScalar** a) {
  if (!v || !a) return LSQED_ERR_ARG;
  *a = v->array;
  return 0;
}
inline PetscErrorCode VecRestoreArray(Vec v, PetscScalar** a) {
  if (!v) return LSQED_ERR_ARG;
  v->state = detail::next_state();
  if (a) *a = nullptr;
  return 0;
}

/* Host copies of the whole vector (n doubles). */
inline PetscErrorCode VecCopyToHost(Vec v, PetscScalar* host) {
  if (!v || !host) return LSQED_ERR_ARG;
  return detail::hip_status(hipMemcpy(host, v->array, size_t(v->n) * sizeof(double), hipMemcpyDeviceToHost));
}
inline PetscErrorCode VecCopyFromHost(Vec v, const PetscScalar* host) {
  if (!v || !host) return LSQED_ERR_ARG;
  v->state = detail::next_state();
  return detail::hip_status(hipMemcpy(v->array, host, size_t(v->n) * sizeof(double), hipMemcpyHostToDevice));
}

inline PetscErrorCode MatGetSize(Mat mat, PetscInt* m, PetscInt* n) {
  if (!mat) return LSQED_ERR_ARG;
  const PetscInt D = lsqed_ham_dim(mat);
  if (m) *m = D;
  if (n) *n = D;
  return 0;
}

/* y = A x on the GPU (the product EPSSolve applies, src/solver.cpp:45). */
inline PetscErrorCode MatMult(Mat mat, Vec x, Vec y) {
  if (!mat || !x || !y || x == y) return LSQED_ERR_ARG;
  const PetscInt D = lsqed_ham_dim(mat);
  if (x->n != D || y->n != D) return LSQED_ERR_ARG;
  y->state = detail::next_state();
  return lsqed_hv(mat, x->array, y->array, nullptr);
}

/* The stored diagonal; entries the reference never sets read 0. */
inline PetscErrorCode MatGetDiagonal(Mat mat, Vec d) {
  if (!mat || !d || d->n != lsqed_ham_dim(mat)) return LSQED_ERR_ARG;
  d->state = detail::next_state();
  return lsqed_diag(mat, d->array, nullptr);
}

/* ------------------------------------------------------------- ASCII views */
/* PETSc's default ASCII viewer for the dumps the reference's tests diff
 * (MatView / VecView, which tests/unit_tests/compare.sh diffs against the
 * .ref files of tests/unit_tests/reference_data).  `nprocs` is the number of ranks whose layout the
 * text shows ("N MPI processes", one "Process [k]" block per rank of a
 * vector, the row split of src/basis.cpp:78-80); the library itself runs on
 * one process. */
struct _p_PetscViewer {
  FILE* fp = nullptr;
  int nprocs = 1;
};
using PetscViewer = _p_PetscViewer*;

inline PetscErrorCode PetscViewerASCIICreate(FILE* fp, int nprocs, PetscViewer* v) {
  if (!fp || !v || nprocs < 1) return LSQED_ERR_ARG;
  *v = new _p_PetscViewer;
  (*v)->fp = fp;
  (*v)->nprocs = nprocs;
  return 0;
}
inline PetscErrorCode PetscViewerDestroy(PetscViewer* v) {
  if (v && *v) {
    delete *v;
    *v = nullptr;
  }
  return 0;
}
/* Standard output in the layout of `nprocs` ranks. */
inline PetscViewer PetscViewerStdout(int nprocs = 1) {
  static _p_PetscViewer v[64];
  const int p = nprocs < 1 ? 1 : (nprocs > 64 ? 64 : nprocs);
  v[p - 1].fp = stdout;
  v[p - 1].nprocs = p;
  return &v[p - 1];
}
#define PETSC_VIEWER_STDOUT_WORLD (::lsqed::PetscViewerStdout(1))

namespace detail {
/* PETSc's ASCII scalar format: %g, with a '.' appended to integral output. */
inline std::string petsc_g(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%g", v);
  std::string s(buf);
  if (s.find_first_of(".eEn") == std::string::npos) s += ".";
  return s;
}
inline void ownership(PetscInt size, int nprocs, int rank, PetscInt& local, PetscInt& start) {
  const PetscInt rest = size % nprocs;
  local = size / nprocs + (rest > rank ? 1 : 0);
  start = rank * local + (rest <= rank ? rest : 0);
}
}  // namespace detail

/* VecView of a vector of the sector: a sequential vector on one rank, an
 * MPI vector with one block per rank otherwise. */
inline PetscErrorCode VecView(Vec v, PetscViewer viewer) {
  if (!v || !viewer || !viewer->fp) return LSQED_ERR_ARG;
  std::vector<double> h(size_t(v->n));
  if (v->n > 0) {
    if (PetscErrorCode e = VecCopyToHost(v, h.data())) return e;
  }
  const int P = viewer->nprocs;
  std::fprintf(viewer->fp, "Vec Object: %d MPI processes\n  type: %s\n", P, P == 1 ? "seq" : "mpi");
  for (int r = 0; r < P; ++r) {
    PetscInt n = 0, s = 0;
    detail::ownership(v->n, P, r, n, s);
    if (P > 1) std::fprintf(viewer->fp, "Process [%d]\n", r);
    for (PetscInt i = s; i < s + n; ++i) std::fprintf(viewer->fp, "%s\n", detail::petsc_g(h[size_t(i)]).c_str());
  }
  return 0;
}

/* MatView of the operator as the reference assembles it
 * (include/hamiltonian.tmpl_func.hpp:97-344): per row, J_k/2 for every
 * antiparallel bond of kind k with |J_k| >= 1e-14 (duplicate columns
 * summed), and the diagonal where build_disorder set it (every row) or
 * build_diag added a nonzero bond count of a kind with |D_k| >= 1e-14.
 * Columns come from the device column export, diagonal values from the
 * device diagonal. */
inline PetscErrorCode MatView(Mat mat, PetscViewer viewer) {
  if (!mat || !viewer || !viewer->fp) return LSQED_ERR_ARG;
  double J[2], Dz[2];
  int flags = 0, nb[2] = {0, 0};
  if (int rc = lsqed_ham_get_params(mat, &J[0], &Dz[0], &J[1], &Dz[1], &flags, &nb[0], &nb[1])) return rc;
  const PetscInt D = lsqed_ham_dim(mat);
  std::vector<double> diag(static_cast<size_t>(D));
  {
    Vec d = nullptr;
    if (PetscErrorCode e = VecCreateDevice(D, &d)) return e;
    PetscErrorCode e = MatGetDiagonal(mat, d);
    if (!e) e = VecCopyToHost(d, diag.data());
    VecDestroy(&d);
    if (e) return e;
  }
  std::fprintf(viewer->fp, "Mat Object: %d MPI processes\n  type: mpiaij\n", viewer->nprocs);
  const PetscInt block = 65536;
  const int maxc[2] = {std::max(nb[0], 1), std::max(nb[1], 1)};
  int64_t* d_cols = nullptr;
  int32_t* d_ncol = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&d_cols), size_t(block) * size_t(std::max(maxc[0], maxc[1])) * 8) !=
          hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&d_ncol), size_t(block) * 4) != hipSuccess) {
    if (d_cols) (void)hipFree(d_cols);
    return LSQED_ERR_OOM;
  }
  std::vector<int64_t> cols[2];
  std::vector<int32_t> ncol[2];
  PetscErrorCode err = 0;
  for (PetscInt r0 = 0; r0 < D && !err; r0 += block) {
    const PetscInt nr = std::min(block, D - r0);
    for (int k = 0; k < 2 && !err; ++k) {
      cols[k].assign(size_t(nr) * size_t(maxc[k]), -1);
      ncol[k].assign(size_t(nr), 0);
      err = lsqed_export_cols(mat, k, r0, nr, maxc[k], d_cols, d_ncol, nullptr);
      if (!err && hipMemcpy(cols[k].data(), d_cols, cols[k].size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
        err = LSQED_ERR_HIP;
      if (!err && hipMemcpy(ncol[k].data(), d_ncol, ncol[k].size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
        err = LSQED_ERR_HIP;
    }
    for (PetscInt r = 0; r < nr && !err; ++r) {
      std::vector<std::pair<int64_t, double>> ent;
      auto add = [&](int64_t c, double v) {
        for (auto& e : ent)
          if (e.first == c) {
            e.second += v;
            return;
          }
        ent.emplace_back(c, v);
      };
      if (flags & LSQED_BUILD_OFFDIAG)
        for (int k = 0; k < 2; ++k) {
          if (std::fabs(J[k]) < 1e-14) continue;
          for (int c = 0; c < ncol[k][size_t(r)]; ++c) add(cols[k][size_t(r) * size_t(maxc[k]) + size_t(c)], J[k] * 0.5);
        }
      bool present = false;
      if (flags & LSQED_BUILD_DISORDER)
        present = true;
      else if (flags & LSQED_BUILD_DIAG)
        for (int k = 0; k < 2; ++k)
          present = present || (std::fabs(Dz[k]) >= 1e-14 && nb[k] - 2 * ncol[k][size_t(r)] != 0);
      if (present) add(r0 + r, diag[size_t(r0 + r)]);
      std::sort(ent.begin(), ent.end());
      std::fprintf(viewer->fp, "row %lld:", (long long)(r0 + r));
      for (const auto& e : ent) std::fprintf(viewer->fp, " (%lld, %s) ", (long long)e.first, detail::petsc_g(e.second).c_str());
      std::fprintf(viewer->fp, "\n");
    }
  }
  (void)hipFree(d_cols);
  (void)hipFree(d_ncol);
  return err;
}

/* ------------------------------------------------------------ Hamiltonian */
/* Hamiltonian<L>(env, basis, lattice, J1, Delta1, J2, Delta2)
 * (include/hamiltonian.tmpl_func.hpp:69-91).  The operator starts empty, as
 * the freshly preallocated matrix does; build_diag, build_off_diag and
 * build_disorder each add their part. */
template <typename L>
class Hamiltonian {
 public:
  Mat hamilt = nullptr;

  Hamiltonian(Environment& env, Basis& m_basis, L& m_lattice, PetscReal m_J1, PetscReal m_D1, PetscReal m_J2,
              PetscReal m_D2)
      : basis{m_basis}, lattice{m_lattice}, J1{m_J1}, J2{m_J2}, D1{m_D1}, D2{m_D2} {
    (void)env;
    lsqed_params p{};
    p.nspins = int(basis.nspins);
    p.nup = int(basis.nspins_up);
    p.n_nn = int(lattice.nn_i.size());
    p.n_nnn = int(lattice.nnn_i.size());
    p.nn_i = lattice.nn_i.data();
    p.nn_j = lattice.nn_j.data();
    p.nnn_i = lattice.nnn_i.data();
    p.nnn_j = lattice.nnn_j.data();
    p.J1 = J1;
    p.D1 = D1;
    p.J2 = J2;
    p.D2 = D2;
    if (basis.nspins != lattice.num_spins() || lsqed_ham_create(&p, &hamilt) != LSQED_OK ||
        lsqed_ham_set_build(hamilt, 0) != LSQED_OK) {
      if (hamilt) lsqed_ham_destroy(hamilt);
      hamilt = nullptr;
      throw std::runtime_error("The matrix has not been properly initialized.\n");
    }
  }
  ~Hamiltonian() { MatClean(); }
  Hamiltonian(const Hamiltonian&) = delete;
  Hamiltonian& operator=(const Hamiltonian&) = delete;

  PetscErrorCode MatClean() {
    const int rc = hamilt ? lsqed_ham_destroy(hamilt) : 0;
    hamilt = nullptr;
    return rc;
  }

  /* Diagonal D1/4 c_nn + D2/4 c_nnn + 1/2 sum_i (+-h_i) on every row
   * (tmpl:97-125); hi holds nspins fields. */
  PetscErrorCode build_disorder(Environment& env, const PetscReal* hi) {
    if (!hi) return LSQED_ERR_ARG;
    (void)env;
    if (int rc = lsqed_ham_set_fields(hamilt, hi)) return rc;
    flags |= LSQED_BUILD_DISORDER;
    return lsqed_ham_set_build(hamilt, flags);
  }
  PetscErrorCode build_disorder(Environment& env, const std::vector<PetscReal>& hi) {
    if (PetscInt(hi.size()) != basis.nspins) return LSQED_ERR_ARG;
    return build_disorder(env, hi.data());
  }
  /* Ising part D_k/4 (parallel - antiparallel bonds) (tmpl:129-225). */
  PetscErrorCode build_diag(Environment& env) {
    (void)env;
    flags |= LSQED_BUILD_DIAG;
    return lsqed_ham_set_build(hamilt, flags);
  }
  /* Spin flips J_k/2 on antiparallel bonds (tmpl:229-344). */
  PetscErrorCode build_off_diag(Environment& env) {
    (void)env;
    flags |= LSQED_BUILD_OFFDIAG;
    return lsqed_ham_set_build(hamilt, flags);
  }

  PetscInt hamilt_size() const { return basis.size; }
  PetscInt hamilt_dim() const { return basis.size * basis.size; }
  PetscReal hamilt_J1() const { return J1; }
  PetscReal hamilt_J2() const { return J2; }
  PetscReal hamilt_D1() const { return D1; }
  PetscReal hamilt_D2() const { return D2; }
  lattice_type hamilt_type() const { return lattice.get_type(); }
  /* Which build calls were made (LSQED_BUILD_* bits); no counterpart in the
   * reference, whose matrix holds the built parts. */
  int build_flags() const { return flags; }
  Basis& hamilt_basis() const { return basis; }
  L& hamilt_lattice() const { return lattice; }

 private:
  Basis& basis;
  L& lattice;
  PetscReal J1, J2, D1, D2;
  int flags = 0;
};

/* -------------------------------------------------------------------- EPS */
/* The eigensolver object: settings of Solver::SolverInit (HEP, Krylov-Schur,
 * smallest real; src/solver.cpp:7-24), solved on the device when
 * Solver::solve runs. */
struct _p_EPS {
  Mat A = nullptr;
  lsqed_eps_opts opts{};
  lsqed_eps* eps = nullptr;
  PetscInt nconv = 0;
};
using EPS = _p_EPS*;
enum EPSErrorType { EPS_ERROR_ABSOLUTE = 0, EPS_ERROR_RELATIVE = 1 };

namespace detail {
inline void eps_reset(EPS s) {
  if (s && s->eps) {
    lsqed_eps_destroy(s->eps);
    s->eps = nullptr;
    s->nconv = 0;
  }
}
inline int dim_arg(PetscInt v) { return v > 0 ? int(v) : 0; }
}  // namespace detail

/* EPSSetDimensions(eps, nev, ncv, mpd): PETSC_DECIDE / PETSC_DEFAULT select
 * SLEPc's defaults. */
inline PetscErrorCode EPSSetDimensions(EPS s, PetscInt nev, PetscInt ncv, PetscInt mpd) {
  if (!s || nev <= 0) return LSQED_ERR_ARG;
  detail::eps_reset(s);
  s->opts.nev = int(nev);
  s->opts.ncv = detail::dim_arg(ncv);
  s->opts.mpd = detail::dim_arg(mpd);
  return 0;
}
inline PetscErrorCode EPSSetTolerances(EPS s, PetscReal tol, PetscInt maxit) {
  if (!s) return LSQED_ERR_ARG;
  detail::eps_reset(s);
  s->opts.tol = tol > 0 ? tol : 0.0;
  s->opts.maxit = detail::dim_arg(maxit);
  return 0;
}
inline PetscErrorCode EPSGetConverged(EPS s, PetscInt* nconv) {
  if (!s || !nconv) return LSQED_ERR_ARG;
  *nconv = s->nconv;
  return 0;
}
inline PetscErrorCode EPSGetIterationNumber(EPS s, PetscInt* its) {
  if (!s || !its || !s->eps) return LSQED_ERR_STATE;
  *its = lsqed_eps_iterations(s->eps);
  return 0;
}
/* Eigenpair i, ascending; the imaginary parts are zero (symmetric operator).
 * Any output may be NULL. */
inline PetscErrorCode EPSGetEigenpair(EPS s, PetscInt i, PetscScalar* kr, PetscScalar* ki, Vec xr, Vec xi) {
  if (!s || !s->eps) return LSQED_ERR_STATE;
  const PetscInt D = lsqed_ham_dim(s->A);
  if ((xr && xr->n != D) || (xi && xi->n != D)) return LSQED_ERR_ARG;
  double lambda = 0.0, err = 0.0;
  if (int rc = lsqed_eps_solution(s->eps, int(i), &lambda, xr ? xr->array : nullptr, &err, nullptr)) return rc;
  if (xr) xr->state = detail::next_state();
  if (xi) {
    xi->state = detail::next_state();
    if (hipMemset(xi->array, 0, size_t(D) * sizeof(double)) != hipSuccess) return LSQED_ERR_HIP;
  }
  if (kr) *kr = lambda;
  if (ki) *ki = 0.0;
  return 0;
}
/* ||A x - lambda x|| / |lambda| (relative) or ||A x - lambda x|| (absolute). */
inline PetscErrorCode EPSComputeError(EPS s, PetscInt i, EPSErrorType type, PetscReal* error) {
  if (!s || !s->eps || !error) return LSQED_ERR_STATE;
  double lambda = 0.0, rel = 0.0;
  if (int rc = lsqed_eps_solution(s->eps, int(i), &lambda, nullptr, &rel, nullptr)) return rc;
  *error = type == EPS_ERROR_RELATIVE ? rel : rel * (lambda != 0.0 ? (lambda < 0 ? -lambda : lambda) : 1.0);
  return 0;
}

namespace Solver {

/* src/solver.cpp:7-24. */
inline PetscErrorCode SolverInit(EPS& solver, Mat matrix, PetscInt nev, PetscInt ncv = PETSC_DEFAULT,
                                 PetscInt mpd = PETSC_DEFAULT) {
  solver = nullptr;
  if (!matrix || nev <= 0) return LSQED_ERR_ARG;
  solver = new _p_EPS;
  solver->A = matrix;
  solver->opts.nev = int(nev);
  solver->opts.ncv = detail::dim_arg(ncv);
  solver->opts.mpd = detail::dim_arg(mpd);
  solver->opts.seed = 1;
  return 0;
}

/* src/solver.cpp:28-32. */
inline PetscErrorCode SolverClean(EPS& solver) {
  if (solver) {
    detail::eps_reset(solver);
    delete solver;
  }
  solver = nullptr;
  return 0;
}

/* EPSSolve + EPSGetConverged (src/solver.cpp:36-52). */
inline PetscErrorCode solve(Environment& env, EPS& solver, PetscInt& nconv) {
  (void)env;
  nconv = 0;
  if (!solver) return LSQED_ERR_STATE;
  detail::eps_reset(solver);
  if (int rc = lsqed_eps_create(solver->A, &solver->opts, &solver->eps)) return rc;
  int n = 0;
  if (int rc = lsqed_eps_solve(solver->eps, &n, nullptr)) return rc;
  solver->nconv = n;
  nconv = n;
  return 0;
}

/* EPSGetEigenpair + EPSComputeError(EPS_ERROR_RELATIVE) (src/solver.cpp:58-68). */
inline PetscErrorCode solution(EPS& solver, PetscInt index, PetscScalar& real_value, Vec& real_vector,
                               PetscReal& error) {
  if (PetscErrorCode e = EPSGetEigenpair(solver, index, &real_value, nullptr, real_vector, nullptr)) return e;
  return EPSComputeError(solver, index, EPS_ERROR_RELATIVE, &error);
}

}  // namespace Solver

/* ------------------------------------------------------------ Correlation */
/* <state| O_ij |state> for the two-spin operators of
 * src/spinOperators/correlation.cpp.  One device pass computes every pair
 * (lsqed_obs_szsz / lsqed_obs_spsm); the result is kept until the state
 * vector changes, so a loop over pairs costs one pass. */
class Correlation {
 public:
  Correlation(Environment& env, Basis& m_basis) : basis{m_basis} { (void)env; }
  virtual ~Correlation() {}
  virtual PetscScalar ExpectVal(Environment& env, PetscInt spin_i, PetscInt spin_j, Vec& state) {
    (void)env;
    const PetscInt N = basis.nspins;
    if (spin_i < 0 || spin_i >= N || spin_j < 0 || spin_j >= N) throw std::out_of_range("spin index out of range");
    refresh(state);
    return pick(size_t(spin_i), size_t(spin_j), size_t(N));
  }

 protected:
  virtual PetscScalar pick(size_t i, size_t j, size_t N) const = 0;
  Basis& basis;
  std::vector<double> zz, pm;  // <S^z_i S^z_j>, <S^+_i S^-_j>, row-major N x N

 private:
  void refresh(Vec state) {
    if (!state || state->n != basis.size) throw std::invalid_argument("state vector does not belong to this basis");
    if (state == last_vec && state->state == last_state && !zz.empty()) return;
    const size_t N = size_t(basis.nspins);
    zz.assign(N * N, 0.0);
    pm.assign(N * N, 0.0);
    lsqed_ham* h = basis.sector_handle();
    detail::check(lsqed_obs_szsz(h, state->array, zz.data(), nullptr));
    detail::check(lsqed_obs_spsm(h, state->array, pm.data(), nullptr));
    last_vec = state;
    last_state = state->state;
  }
  Vec last_vec = nullptr;
  uint64_t last_state = 0;
};

/* S^z_i S^z_j (correlation.cpp:86-139). */
class SzSz : public Correlation {
 public:
  using Correlation::Correlation;

 protected:
  PetscScalar pick(size_t i, size_t j, size_t N) const override { return zz[i * N + j]; }
};

/* S^+_i S^-_j (correlation.cpp:154-224). */
class SpSm : public Correlation {
 public:
  using Correlation::Correlation;

 protected:
  PetscScalar pick(size_t i, size_t j, size_t N) const override { return pm[i * N + j]; }
};

/* S^-_i S^+_j (correlation.cpp:237-261): for a real state the off-diagonal
 * entries equal <S^+_j S^-_i>; on the diagonal <n_down_i> = <psi|psi> -
 * <n_up_i>, with <psi|psi> = 4 <S^z_i S^z_i>. */
class SmSp : public SpSm {
 public:
  using SpSm::SpSm;

 protected:
  PetscScalar pick(size_t i, size_t j, size_t N) const override {
    if (i != j) return pm[j * N + i];
    return 4.0 * zz[i * N + i] - pm[i * N + i];
  }
};

/* S_i . S_j = (S^+_i S^-_j + S^-_i S^+_j)/2 + S^z_i S^z_j (correlation.cpp:276-299). */
class SiSj : public Correlation {
 public:
  using Correlation::Correlation;

 protected:
  PetscScalar pick(size_t i, size_t j, size_t N) const override {
    const double mp = i != j ? pm[j * N + i] : 4.0 * zz[i * N + i] - pm[i * N + i];
    return (pm[i * N + j] + mp) / 2.0 + zz[i * N + j];
  }
};

/* ---------------------------------------------------------- Sq operators */
/* Operators of the dynamical structure factor (include/spinOperators/
 * spinOPft.hpp): Sqz applies S^z_i within the sector, Sqp applies S^+_i into
 * the sector with one more up spin.  Used as the SQ argument of
 * Phys::DynStructFactor. */
struct Sq_data {
  Basis* b0;
  Basis* b1;
};
class Sqz {
 public:
  static constexpr PetscInt up_shift = 0;
  Sqz(Environment&, Sq_data& m_data) : data{m_data} {}
  Sq_data data;
};
class Sqp {
 public:
  static constexpr PetscInt up_shift = 1;
  Sqp(Environment&, Sq_data& m_data) : data{m_data} {}
  Sq_data data;
};

/* ------------------------------------------------------------------- Phys */
namespace Phys {

template <typename L>
struct DSF_data {
  DSF_data() {}
  Basis* b0 = nullptr;
  Basis* b1 = nullptr;
  L* lat = nullptr;
  Hamiltonian<L>* h0 = nullptr;
  Hamiltonian<L>* h1 = nullptr;
  PetscInt nev = 0;
  PetscInt ncv = PETSC_DECIDE;
  PetscInt mpd = PETSC_DECIDE;
  PetscBool disorder = PETSC_FALSE;
  PetscReal* hi = nullptr;
  const char* path = ".";
};

namespace detail_phys {
template <typename SL>
double sublattice_sum(Environment& env, Basis& basis, SL& sublat, const std::vector<PetscInt>& which, Vec& state) {
  SiSj cij{env, basis};
  double acc = 0.0;
  for (PetscInt k : which) {
    const std::vector<PetscInt>& sl = sublat.get_sl(k);
    for (PetscInt si : sl)
      for (PetscInt sj : sl) acc += cij.ExpectVal(env, si, sj, state);
  }
  return acc;
}
}  // namespace detail_phys

/* m^2_AF = 8 sum_{sublattices} sum_{i,j in it} <S_i.S_j> / (N (N+4))
 * (include/phys.tmpl_func.hpp:5-41). */
template <typename L>
PetscScalar SquaredSubLattMagAF(Environment& env, Basis& basis, AF<L>& sublat, Vec& state) {
  std::vector<PetscInt> which;
  for (PetscInt k = 0; k < sublat.get_size(); ++k) which.push_back(k);
  const double acc = detail_phys::sublattice_sum(env, basis, sublat, which, state);
  return acc * 8.0 / PetscScalar(env.nspins) / (PetscScalar(env.nspins) + 4.0);
}

/* m^2_STR = 4 sum over sublattices 0..3 / (N (N+4)) (phys.tmpl_func.hpp:45-81). */
template <typename L>
PetscScalar SquaredSubLattMagSTR(Environment& env, Basis& basis, Striped<L>& sublat, Vec& state) {
  const double acc = detail_phys::sublattice_sum(env, basis, sublat, {0, 1, 2, 3}, state);
  return acc * 4.0 / PetscScalar(env.nspins) / (PetscScalar(env.nspins) + 4.0);
}

/* Phys::DynStructFactor<L, SQ> (include/phys.tmpl_func.hpp:87-215): the nev
 * lowest states |n> of data.h1's sector, the amplitudes <n| SQ_i |state> of
 * every spin i, written to <path>/dsf_d<dis_iter>.dat in the reference's
 * format.  The amplitudes of all spins and states come from one device pass
 * (lsqed_obs_dsf). */
template <typename L, typename SQ>
PetscErrorCode DynStructFactor(Environment& env, DSF_data<L>& data, PetscScalar gs, Vec& state,
                               PetscInt dis_iter = 0) {
  static_assert(std::is_same<SQ, Sqz>::value || std::is_same<SQ, Sqp>::value, "SQ is Sqz or Sqp");
  if (!data.b0 || !data.b1 || !data.h0 || !data.h1 || !state || data.nev <= 0) return LSQED_ERR_ARG;
  if (data.b1->nspins_up != data.b0->nspins_up + SQ::up_shift || state->n != data.b0->size) return LSQED_ERR_ARG;
  if (data.disorder && !data.hi) return LSQED_ERR_ARG;

  EPS solver = nullptr;
  PetscInt nconv = 0;
  PetscErrorCode ierr = Solver::SolverInit(solver, data.h1->hamilt, data.nev, data.ncv, data.mpd);
  if (!ierr) ierr = Solver::solve(env, solver, nconv);
  const PetscInt nev = data.nev;
  const PetscInt N = env.nspins;
  const PetscInt D1 = data.h1->hamilt_size();
  std::vector<double> eig(size_t(nev), 0.0), amp(size_t(nev * N), 0.0);
  double* evecs = nullptr;
  if (!ierr && hipMalloc(reinterpret_cast<void**>(&evecs), size_t(nev * D1) * sizeof(double)) != hipSuccess)
    ierr = LSQED_ERR_OOM;
  for (PetscInt i = 0; i < nev && !ierr; ++i) {
    double err = 0.0;
    ierr = lsqed_eps_solution(solver->eps, int(i), &eig[size_t(i)], evecs + i * D1, &err, nullptr);
  }
  if (!ierr)
    ierr = lsqed_obs_dsf(data.h0->hamilt, data.h1->hamilt, state->array, evecs, int(nev), amp.data(), nullptr);
  if (evecs) (void)hipFree(evecs);
  Solver::SolverClean(solver);
  if (ierr) return ierr;

  const std::string fn = std::string(data.path) + "/dsf_d" + std::to_string(dis_iter) + ".dat";
  FILE* fp = std::fopen(fn.c_str(), "w");
  if (!fp) return LSQED_ERR_ARG;  // "Cannot open file" (phys.tmpl_func.hpp:133)
  std::fprintf(fp, "#------------------------------- Parameters -------------------------------------\n");
  std::fprintf(fp, "#%7s:   %lld  \n", " nspins", (long long)N);
  std::fprintf(fp, "#%7s:   %lld  \n", " nev", (long long)nev);
  std::fprintf(fp, "#%7s:   %.16f  \n", " GS", gs);
  std::fprintf(fp, "#%7s:   %.5f  %4s:   %.5f  %4s:   %.5f  %4s:   %.5f\n", "J1", data.h0->hamilt_J1(), "D1",
               data.h0->hamilt_D1(), "J2", data.h0->hamilt_J2(), "D2", data.h0->hamilt_D2());
  if (data.disorder)
    for (PetscInt i = 0; i < N; ++i) std::fprintf(fp, "# hi[%4lld]: %.16f\n", (long long)i, data.hi[i]);
  std::fprintf(fp, "#--------------------------------------------------------------------------------\n#\n#\n");
  std::fprintf(fp, "#----- Energy -----  ------ Factor -----\n");
  for (PetscInt spin = 0; spin < N; ++spin) {
    std::fprintf(fp, "# spin: %lld\n", (long long)spin);
    for (PetscInt n = 0; n < nev; ++n)
      std::fprintf(fp, "%.13e  %.13e\n", eig[size_t(n)] - gs, amp[size_t(n * N + spin)]);
  }
  return std::fclose(fp) == 0 ? 0 : LSQED_ERR_ARG;
}

}  // namespace Phys

/* -------------------------------------------------------- batched solver */
/* The realisation loop of examples/main_disorder.cpp:82-120 (per
 * realisation: build_disorder, SolverInit with nev = 1, solve, solution) as
 * one batched three-term Lanczos over many models (lsqed_batch_*).  The
 * models are referenced, not copied: build_disorder on them between solves
 * takes effect at the next solve.  No counterpart in the reference, which
 * solves the realisations one at a time. */
template <typename L>
class BatchSolver {
 public:
  struct Options {
    PetscInt maxit = 300;
    PetscReal tol = 1e-9;
    uint64_t seed = 1;
    PetscBool keep_vectors = PETSC_FALSE;  // a second pass builds the ground states
  };

  explicit BatchSolver(const std::vector<Hamiltonian<L>*>& models) : BatchSolver(models, Options{}) {}
  BatchSolver(const std::vector<Hamiltonian<L>*>& models, const Options& o) {
    if (models.empty()) throw std::invalid_argument("the batch needs at least one model");
    for (Hamiltonian<L>* m : models) {
      if (!m || !m->hamilt) throw std::invalid_argument("model without an operator");
      hams.push_back(m->hamilt);
    }
    lsqed_batch_opts bo{};
    bo.maxit = int(o.maxit);
    bo.tol = o.tol;
    bo.seed = o.seed;
    bo.keep_vectors = o.keep_vectors ? 1 : 0;
    detail::check(lsqed_batch_create(int(hams.size()), hams.data(), &bo, &batch));
  }
  ~BatchSolver() {
    if (batch) lsqed_batch_destroy(batch);
  }
  BatchSolver(const BatchSolver&) = delete;
  BatchSolver& operator=(const BatchSolver&) = delete;

  PetscInt size() const { return PetscInt(hams.size()); }

  /* Lowest eigenvalue of every model, in the order given. */
  PetscErrorCode solve(std::vector<PetscScalar>& e0) {
    e0.assign(hams.size(), 0.0);
    return lsqed_batch_solve(batch, e0.data(), nullptr);
  }
  /* Lanczos steps, relative residual |beta_m s_m| / |theta| and convergence of model k. */
  PetscErrorCode info(PetscInt k, PetscInt& steps, PetscReal& resid, PetscBool& converged) const {
    int s = 0, c = 0;
    double r = 0.0;
    const int rc = lsqed_batch_info(batch, int(k), &s, &r, &c);
    steps = s;
    resid = r;
    converged = c != 0;
    return rc;
  }
  /* Unit-norm ground state of model k (needs keep_vectors) into a vector of
   * its sector (MatCreateVecs on that model). */
  PetscErrorCode ground_state(PetscInt k, Vec& state) {
    if (!state || k < 0 || k >= size() || state->n != lsqed_ham_dim(hams[size_t(k)])) return LSQED_ERR_ARG;
    state->state = detail::next_state();
    return lsqed_batch_ground_state(batch, int(k), state->array, nullptr);
  }
  /* dE0/dparameter of every model from its ground state (needs
   * keep_vectors): out[k * LSQED_BATCH_OBS_STRIDE + j], layout of
   * lsqed_batch_observables. */
  PetscErrorCode observables(std::vector<double>& out) {
    out.assign(hams.size() * LSQED_BATCH_OBS_STRIDE, 0.0);
    return lsqed_batch_observables(batch, out.data(), nullptr);
  }

 private:
  std::vector<lsqed_ham*> hams;
  lsqed_batch* batch = nullptr;
};

/* ------------------------------------------------------------------ Tools */
namespace Tools {
class RandomDisorder;
}

namespace Phys {
/* The realisation loop of examples/main_disorder.cpp:82-120 with the
 * ground states solved in batches: for every realisation k of `rd` not yet
 * on disk (resume: <data.path>/dsf_d<k>.dat exists), a copy of `h` (same
 * lattice, couplings and build calls) gets the fields rd.hi_init(k); the
 * pending realisations' ground states come from BatchSolver (batch_size at
 * a time), then DynStructFactor<L, SQ> writes each realisation's file with
 * data.h1 the target-sector model (nullptr: the realisation's own model; Sqp
 * needs it) carrying the same fields.  Returns {k: E0} for the realisations
 * computed in this call. */
template <typename L, typename SQ>
PetscErrorCode DisorderAverage(Environment& env, Hamiltonian<L>& h, Tools::RandomDisorder& rd, DSF_data<L> data,
                               std::vector<std::pair<PetscInt, PetscScalar>>& done, PetscBool resume = PETSC_TRUE,
                               PetscInt batch_size = 256, PetscReal tol = 1e-12);
}  // namespace Phys

namespace Tools {

/* Tools::RandomDisorder(nspins, reprod, min, max, rep)
 * (src/tools/random_disorder.cpp:9-59): one seed in [0, 1000] per
 * realisation (mt19937_64(1) when reprod, a random_device seed otherwise);
 * hi_init(k) draws nspins fields uniform in [min, max) from mt19937_64(seed k). */
class RandomDisorder {
 public:
  PetscReal min = 0.0, max = 0.0;
  PetscInt rep = 0, nspins = 0;
  std::vector<PetscInt> srand_seq;
  std::vector<PetscReal> hi;

  RandomDisorder() {}
  RandomDisorder(PetscInt m_nspins, PetscBool m_reprod, PetscReal m_min, PetscReal m_max, PetscInt m_rep)
      : min{m_min}, max{m_max}, rep{m_rep}, nspins{m_nspins}, srand_seq(size_t(m_rep), 0), hi(size_t(m_nspins), 0.0) {
    srand_seq_init(m_reprod);
  }

  PetscInt* srand_seq_init(PetscBool reprod) {
    srand_seq.assign(size_t(rep), 0);
    if (reprod) {
      static_assert(std::is_same<PetscInt, int64_t>::value, "seed list is int64_t");
      detail::check(lsqed_rd_seeds(int(rep), srand_seq.data()));
    } else {
      std::random_device rd;
      std::mt19937_64 gen(rd());
      std::uniform_int_distribution<PetscInt> dist(0, 1000);
      for (auto& v : srand_seq) v = dist(gen);
    }
    return srand_seq.data();
  }

  void hi_init(PetscInt dis_index) {
    detail::check(lsqed_rd_fields(srand_seq.at(size_t(dis_index)), int(nspins), min, max, hi.data()));
  }
};

}  // namespace Tools

namespace Phys {
template <typename L, typename SQ>
PetscErrorCode DisorderAverage(Environment& env, Hamiltonian<L>& h, Tools::RandomDisorder& rd, DSF_data<L> data,
                               std::vector<std::pair<PetscInt, PetscScalar>>& done, PetscBool resume,
                               PetscInt batch_size, PetscReal tol) {
  done.clear();
  if (batch_size <= 0 || !data.path || (std::is_same<SQ, Sqp>::value && !data.h1)) return LSQED_ERR_ARG;
  Hamiltonian<L>* target = data.h1;
  std::vector<PetscInt> pending;
  for (PetscInt k = 0; k < rd.rep; ++k) {
    const std::string fn = std::string(data.path) + "/dsf_d" + std::to_string(k) + ".dat";
    FILE* fp = resume ? std::fopen(fn.c_str(), "r") : nullptr;
    if (fp)
      std::fclose(fp);
    else
      pending.push_back(k);
  }
  auto make = [&](Hamiltonian<L>& like, const std::vector<PetscReal>& hi) {
    auto m = std::unique_ptr<Hamiltonian<L>>(new Hamiltonian<L>{env, like.hamilt_basis(), like.hamilt_lattice(),
                                                                like.hamilt_J1(), like.hamilt_D1(), like.hamilt_J2(),
                                                                like.hamilt_D2()});
    PetscErrorCode e = 0;
    if (like.build_flags() & LSQED_BUILD_DIAG) e = m->build_diag(env);
    if (!e && (like.build_flags() & LSQED_BUILD_OFFDIAG)) e = m->build_off_diag(env);
    if (!e) e = m->build_disorder(env, hi);
    if (e) throw std::runtime_error(lsqed_last_error());
    return m;
  };
  for (size_t c0 = 0; c0 < pending.size(); c0 += size_t(batch_size)) {
    const size_t c1 = std::min(pending.size(), c0 + size_t(batch_size));
    std::vector<std::unique_ptr<Hamiltonian<L>>> models;
    std::vector<std::vector<PetscReal>> fields;
    std::vector<Hamiltonian<L>*> ptrs;
    for (size_t j = c0; j < c1; ++j) {
      rd.hi_init(pending[j]);
      fields.push_back(rd.hi);
      models.push_back(make(h, rd.hi));
      ptrs.push_back(models.back().get());
    }
    typename BatchSolver<L>::Options o;
    o.tol = tol;
    o.keep_vectors = PETSC_TRUE;
    BatchSolver<L> batch{ptrs, o};
    std::vector<PetscScalar> e0;
    if (PetscErrorCode e = batch.solve(e0)) return e;
    for (size_t j = 0; j < ptrs.size(); ++j) {
      Vec psi = nullptr;
      if (PetscErrorCode e = MatCreateVecs(ptrs[j]->hamilt, &psi, NULL)) return e;
      PetscErrorCode e = batch.ground_state(PetscInt(j), psi);
      std::unique_ptr<Hamiltonian<L>> tgt;
      if (!e) {
        DSF_data<L> d = data;
        d.h0 = ptrs[j];
        d.b0 = &ptrs[j]->hamilt_basis();
        if (target) {
          tgt = make(*target, fields[j]);
          d.h1 = tgt.get();
          d.b1 = &target->hamilt_basis();
        } else {
          d.h1 = ptrs[j];
          d.b1 = d.b0;
        }
        d.disorder = PETSC_TRUE;
        d.hi = fields[j].data();
        e = DynStructFactor<L, SQ>(env, d, e0[j], psi, pending[c0 + j]);
      }
      VecDestroy(&psi);
      if (e) return e;
      done.emplace_back(pending[c0 + j], e0[j]);
    }
  }
  return 0;
}
}  // namespace Phys
}  // namespace lsqed

#endif  // LSQUANTUMED_HPP
