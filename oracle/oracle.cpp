// oracle.cpp -- CPU restatement of LSQuantumED's Sz-sector Hamiltonian path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product is
// compared against; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  The product library never links or calls it.
//
// Every function restates one routine of the reference (file:line cited on it)
// in plain C++17, with the reference's own arithmetic: binomials in double
// rounded to the nearest integer (the Boost routine the reference calls),
// ranks summed in double, diagonal terms evaluated left to right without FMA
// contraction (the reference is built for baseline x86-64, which has no FMA).
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <random>
#include <vector>
#include <omp.h>

namespace {

// Boost.Math binomial_coefficient<double>(n, k) for n below its factorial
// table limit (170): n! / (n-k)! / k! from a table of correctly rounded
// factorials, then ceil(x - 0.5).  Called by src/basis.cpp:26 and
// src/hamiltonian.cpp:23.  Boost's version is unpinned by the reference
// (meson_options.txt:4 takes boostdir from the user); this is the published
// algorithm of boost/math/special_functions/binomial.hpp.
double factorial_d(int n) {
  // n! as a double (accumulated in long double).  The ratio below is rounded
  // to the nearest integer, so the last-bit rounding of the factorials cannot
  // change the result while C(n,k) < 2^50; tests check every n <= 64 against
  // exact integer binomials.
  long double f = 1.0L;
  for (int i = 2; i <= n; ++i) f *= (long double)i;
  return (double)f;
}

double binom_boost(int n, int k) {
  if (k > n) return 0.0;  // Boost raises a domain error; callers never pass it
  if (k == 0 || k == n) return 1.0;
  if (k == 1 || k == n - 1) return (double)n;
  double r = factorial_d(n);
  r /= factorial_d(n - k);
  r /= factorial_d(k);
  return std::ceil(r - 0.5);
}

struct Lattice {
  // Flat neighbour lists exactly as src/lattice.cpp builds them:
  // [site, nb_1 .. nb_K] per site with stride K+1, and per-site counts.
  std::vector<int64_t> nn, nnn, nnX, nnnX;
  int num_nn = 1, num_nnn = 1;
};

// chain1D::get_nn / get_nnn, src/lattice.cpp:27-53
void chain_lattice(int N, Lattice& L) {
  L.num_nn = 1; L.num_nnn = 1;
  for (int i = 0; i < N; ++i) { L.nn.push_back(i); L.nn.push_back((i + 1) % N); L.nnX.push_back(1); }
  for (int i = 0; i < N; ++i) { L.nnn.push_back(i); L.nnn.push_back((i + 2) % N); L.nnnX.push_back(1); }
}

// square2D::get_nn / get_nnn, src/lattice.cpp:86-136
void square_lattice(int lx, int ly, Lattice& L) {
  L.num_nn = 2; L.num_nnn = 2;
  for (int ix = 0; ix < lx; ++ix)
    for (int iy = 0; iy < ly; ++iy) {
      int cur = ix * ly + iy;
      L.nn.push_back(cur);
      L.nn.push_back(ix * ly + (iy + 1) % ly);
      L.nn.push_back(((ix + 1) % lx) * ly + iy);
      L.nnX.push_back(2);
    }
  for (int ix = 0; ix < lx; ++ix)
    for (int iy = 0; iy < ly; ++iy) {
      int cur = ix * ly + iy;
      L.nnn.push_back(cur);
      L.nnn.push_back(((ix + 1) % lx) * ly + (iy + 1) % ly);
      L.nnn.push_back(((ix + 1) % lx) * ly + (iy - 1 + ly) % ly);
      L.nnnX.push_back(2);
    }
}

// honeycomb2D::get_nn / get_nnn, src/lattice.cpp:172-264 (padding -1 kept).
void honeycomb_lattice(int lx, int ly, Lattice& L) {
  L.num_nn = 2; L.num_nnn = 3;
  for (int ix = 0; ix < lx; ++ix)
    for (int iy = 0; iy < ly; ++iy) {
      int cur = ix * ly + iy;
      L.nn.push_back(cur);
      bool two = (ix % 2 == 0) ? (cur % 2 == 0) : (cur % 2 != 0);
      L.nn.push_back(ix * ly + (iy + 1) % ly);
      if (two) { L.nn.push_back(((ix + 1) % lx) * ly + iy); L.nnX.push_back(2); }
      else { L.nn.push_back(-1); L.nnX.push_back(1); }
    }
  for (int ix = 0; ix < lx; ++ix)
    for (int iy = 0; iy < ly; ++iy) {
      int cur = ix * ly + iy;
      L.nnn.push_back(cur);
      L.nnnX.push_back(3);
      int ix1 = (ix + 1) % lx;
      L.nnn.push_back(ix * ly + (iy + 2) % ly);
      L.nnn.push_back(ix1 * ly + (iy + 1) % ly);
      L.nnn.push_back(ix1 * ly + (iy - 1 + ly) % ly);
    }
}

bool make_lattice(int type, int N, int lx, int ly, Lattice& L) {
  if (type == 0) { chain_lattice(N, L); return true; }
  if (lx * ly != N) return false;  // src/lattice.cpp:67-68, :150-151
  if (type == 1) { square_lattice(lx, ly, L); return true; }
  if (type == 2) {
    if (lx % 2 || ly % 2) return false;  // src/lattice.cpp:153-154
    honeycomb_lattice(lx, ly, L); return true;
  }
  return false;
}

// Ordered bond list of one neighbour kind, in the loop order of
// HamiltHelper::get_coup_elems / get_elem_diag
// (include/hamiltonian.tmpl_func.hpp:360-384, :399-423).
void ordered_bonds(const Lattice& L, int N, int kind, std::vector<int>& bi, std::vector<int>& bj) {
  const std::vector<int64_t>& lst = kind == 0 ? L.nn : L.nnn;
  const std::vector<int64_t>& cnt = kind == 0 ? L.nnX : L.nnnX;
  int stride = (kind == 0 ? L.num_nn : L.num_nnn) + 1;
  for (int i = 0; i < N; ++i)
    for (int nb = 1; nb <= cnt[i]; ++nb) {
      bi.push_back(i);
      bj.push_back((int)(short)lst[(size_t)stride * i + nb]);  // `short int spin` in the reference
    }
}

// HamiltHelper::search_index, src/hamiltonian.cpp:9-27: rank summed in double.
int64_t rank_ref(uint64_t s, int N) {
  int64_t up = 0;
  double idx = 0.0;
  for (int bit = 0; bit < N; ++bit)
    if ((s >> bit) & 1) {
      ++up;
      if (!(bit < up)) idx += binom_boost(bit, (int)up);
    }
  return (int64_t)std::floor(idx + 0.5);
}

// Exact uint64 binomial table, used by the matrix-free CPU H.v (the CPU
// baseline) so that it does not time Boost calls; checked against rank_ref.
struct BinomTable {
  uint64_t c[65][65];
  BinomTable() {
    std::memset(c, 0, sizeof(c));
    for (int n = 0; n <= 64; ++n) {
      c[n][0] = 1;
      for (int k = 1; k <= n; ++k) c[n][k] = c[n - 1][k - 1] + (k <= n - 1 ? c[n - 1][k] : 0);
    }
  }
};
const BinomTable& btab() { static BinomTable t; return t; }

inline uint64_t rank_exact(uint64_t s, const BinomTable& T) {
  uint64_t r = 0; int k = 0;
  while (s) {
    int p = __builtin_ctzll(s);
    ++k;
    r += T.c[p][k];  // c[p][k] == 0 for p < k
    s &= s - 1;
  }
  return r;
}

// BasisHelper::search_elem, src/basis.cpp:9-36 (greedy unrank, double binomials).
uint64_t unrank_ref(int64_t index, int N, int nup) {
  int64_t pi = N - 1, rest = index, sup = nup;
  uint64_t elem = 0;
  while (sup > 0) {
    double b = (pi < sup ? 0.0 : binom_boost((int)pi, (int)sup));
    if ((int64_t)std::floor(b + 0.5) <= rest) {
      rest -= (int64_t)b;  // `rest -= binomial` converts through double; exact here
      elem |= (1ull << pi);
      sup -= 1;
    }
    pi -= 1;
  }
  return elem;
}

// Gosper's hack as in Basis::generate_int_basis, src/basis.cpp:208-213.
// The reference never steps past the last state of a sector; the loops here
// do, once, and the empty state (nup = 0, a sector of one state) has no
// lowest set bit to divide by, so it maps to itself.
inline uint64_t gosper_next(uint64_t e) {
  if (e == 0) return 0;
  uint64_t u = e & (0 - e);
  uint64_t v = u + e;
  return (((v ^ e) / u) >> 2) + v;
}

// get_elem_diag_AUX summed over one kind, src/hamiltonian.cpp:53-68 and
// include/hamiltonian.tmpl_func.hpp:391-426.
int64_t elem_diag(uint64_t s, const std::vector<int>& bi, const std::vector<int>& bj) {
  int64_t d = 0;
  for (size_t b = 0; b < bi.size(); ++b) {
    uint64_t a = (s >> bi[b]) & 1, c = (s >> bj[b]) & 1;
    d += (a ^ c) ? -1 : 1;
  }
  return d;
}

// HamiltHelper::get_disorder_elem, src/hamiltonian.cpp:31-49.
double disorder_elem(uint64_t s, int N, const double* hi) {
  double d = 0.0;
  for (int i = 0; i < N; ++i) {
    if ((s >> i) & 1) d += hi[i];
    else d -= hi[i];
  }
  return d;
}

struct Model {
  int N = 0, nup = 0;
  std::vector<int> bi[2], bj[2];  // kind 0 = nn (J1, D1), kind 1 = nnn (J2, D2)
  double J[2] = {0, 0}, Dz[2] = {0, 0};
};

struct Csr {
  int64_t D = 0;
  std::vector<int64_t> rowptr, cols;
  std::vector<double> vals;
};

}  // namespace

extern "C" {

double oq_binom_boost(int n, int k) { return binom_boost(n, k); }
uint64_t oq_binom_exact(int n, int k) { return (n < 0 || k < 0 || n > 64 || k > n) ? 0 : btab().c[n][k]; }
int64_t oq_rank_ref(uint64_t s, int N) { return rank_ref(s, N); }
uint64_t oq_rank_exact(uint64_t s) { return rank_exact(s, btab()); }
uint64_t oq_unrank_ref(int64_t index, int N, int nup) { return unrank_ref(index, N, nup); }

// Basis::compute_size, src/basis.cpp:38-46.
int64_t oq_compute_size(int N, int nup) {
  double size = 1.0;
  for (int i = 1; i <= (N - nup); ++i) size *= ((double)(i + nup) / (double)i);
  return (int64_t)std::floor(size + 0.5);
}

// Row split of Basis::Basis, src/basis.cpp:78-80 (also the PETSc split used
// by Hamiltonian::prealloc, include/hamiltonian.tmpl_func.hpp:437-444).
void oq_split(int64_t size, int P, int rank, int64_t* local_size, int64_t* start) {
  int64_t rest = size % P;
  *local_size = size / P + (rest > rank);
  *start = (int64_t)rank * (*local_size) + (rest <= rank) * rest;
}

// Basis::generate_int_basis, src/basis.cpp:192-215: unrank the first state,
// then Gosper's hack.
void oq_int_basis(int N, int nup, int64_t start, int64_t count, uint64_t* out) {
  if (count <= 0) return;
  uint64_t e = unrank_ref(start, N, nup);
  out[0] = e;
  for (int64_t i = 1; i < count; ++i) { e = gosper_next(e); out[i] = e; }
}

// Flat lattice lists (src/lattice.cpp).  Returns 0 on success, -1 on the
// reference's invalid_argument cases.  Arrays must hold N*(K+1) / N entries.
int oq_lattice(int type, int N, int lx, int ly, int64_t* nn, int64_t* nnn, int64_t* nnX, int64_t* nnnX,
               int* num_nn, int* num_nnn) {
  Lattice L;
  if (!make_lattice(type, N, lx, ly, L)) return -1;
  if (nn) std::copy(L.nn.begin(), L.nn.end(), nn);
  if (nnn) std::copy(L.nnn.begin(), L.nnn.end(), nnn);
  if (nnX) std::copy(L.nnX.begin(), L.nnX.end(), nnX);
  if (nnnX) std::copy(L.nnnX.begin(), L.nnnX.end(), nnnX);
  *num_nn = L.num_nn; *num_nnn = L.num_nnn;
  return 0;
}

// Ordered bond list for one kind; returns the count (or -1).  bi/bj may be
// NULL to query the count.
int oq_bonds(int type, int N, int lx, int ly, int kind, int* bi, int* bj) {
  Lattice L;
  if (!make_lattice(type, N, lx, ly, L)) return -1;
  std::vector<int> a, b;
  ordered_bonds(L, N, kind, a, b);
  if (bi) std::copy(a.begin(), a.end(), bi);
  if (bj) std::copy(b.begin(), b.end(), bj);
  return (int)a.size();
}

int64_t oq_elem_diag(uint64_t s, int nb, const int* bi, const int* bj) {
  std::vector<int> a(bi, bi + nb), b(bj, bj + nb);
  return elem_diag(s, a, b);
}

double oq_disorder_elem(uint64_t s, int N, const double* hi) { return disorder_elem(s, N, hi); }

// HamiltHelper::get_coup_elems for one state, in emission order
// (include/hamiltonian.tmpl_func.hpp:350-387; src/hamiltonian.cpp:72-99).
int oq_coup_elems(uint64_t s, int N, int nb, const int* bi, const int* bj, int64_t* cols) {
  int ncol = 0;
  for (int b = 0; b < nb; ++b) {
    uint64_t a = (s >> bi[b]) & 1, c = (s >> bj[b]) & 1;
    if (a ^ c) {
      uint64_t t = s ^ (1ull << bi[b]);
      t ^= (1ull << bj[b]);
      cols[ncol++] = rank_ref(t, N);
    }
  }
  return ncol;
}

// --- model handle: lattice + couplings -------------------------------------
void* oq_model_create(int type, int N, int lx, int ly, int nup, double J1, double D1, double J2, double D2) {
  Lattice L;
  if (!make_lattice(type, N, lx, ly, L)) return nullptr;
  Model* m = new Model;
  m->N = N; m->nup = nup;
  ordered_bonds(L, N, 0, m->bi[0], m->bj[0]);
  ordered_bonds(L, N, 1, m->bi[1], m->bj[1]);
  m->J[0] = J1; m->Dz[0] = D1; m->J[1] = J2; m->Dz[1] = D2;
  return m;
}
void oq_model_free(void* h) { delete (Model*)h; }

// Diagonal of H after the reference's build sequence:
//  flags & 1: build_diag   (tmpl:129-225)  ADD D*0.25*c, skipped for |D|<1e-14 or c==0
//  flags & 4: build_disorder(tmpl:97-125)  INSERT D1*0.25*c1 + D2*0.25*c2 + 0.5*dis
// `present[r]` is set when the entry is stored in the PETSc matrix.
void oq_diag(void* h, const double* hi, int flags, double* d, uint8_t* present) {
  Model& m = *(Model*)h;
  int64_t D = oq_compute_size(m.N, m.nup);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < D; ++r) {
    uint64_t s = unrank_ref(r, m.N, m.nup);
    double v = 0.0; uint8_t p = 0;
    if (flags & 1) {
      if (!(std::fabs(m.Dz[0]) < 1e-14 && std::fabs(m.Dz[1]) < 1e-14)) {
        for (int k = 0; k < 2; ++k) {
          if (std::fabs(m.Dz[k]) < 1e-14) continue;
          int64_t c = elem_diag(s, m.bi[k], m.bj[k]);
          if (c == 0) continue;
          v += m.Dz[k] * 0.25 * (double)c;
          p = 1;
        }
      }
    }
    if ((flags & 4) && hi) {
      int64_t c1 = elem_diag(s, m.bi[0], m.bj[0]);
      int64_t c2 = elem_diag(s, m.bi[1], m.bj[1]);
      v = m.Dz[0] * 0.25 * (double)c1 + m.Dz[1] * 0.25 * (double)c2 + 0.5 * disorder_elem(s, m.N, hi);
      p = 1;
    }
    d[r] = v;
    if (present) present[r] = p;
  }
}

// Assembled matrix after the build sequence (flags: 1 diag, 2 off-diag,
// 4 disorder), as PETSc stores it: per row, columns sorted ascending,
// duplicate (row, col) ADDs summed (build_off_diag uses ADD_VALUES,
// tmpl:258-266).  Returns an opaque handle.
void* oq_csr_build(void* h, const double* hi, int flags) {
  Model& m = *(Model*)h;
  int64_t D = oq_compute_size(m.N, m.nup);
  std::vector<double> d(D);
  std::vector<uint8_t> pres(D);
  oq_diag(h, hi, flags, d.data(), pres.data());
  Csr* A = new Csr;
  A->D = D;
  int nthr = omp_get_max_threads();
  std::vector<std::vector<int64_t>> rc(nthr);
  std::vector<std::vector<double>> rv(nthr);
  std::vector<int64_t> rowlen(D);
  std::vector<int64_t> chunk_start(nthr + 1);
  for (int t = 0; t <= nthr; ++t) chunk_start[t] = D * t / nthr;
#pragma omp parallel num_threads(nthr)
  {
    int t = omp_get_thread_num();
    std::map<int64_t, double> row;
    std::vector<int64_t> cols(4 * 64 + 8);
    for (int64_t r = chunk_start[t]; r < chunk_start[t + 1]; ++r) {
      row.clear();
      uint64_t s = unrank_ref(r, m.N, m.nup);
      if (pres[r]) row[r] += d[r];
      if (flags & 2) {
        for (int k = 0; k < 2; ++k) {
          if (std::fabs(m.J[k]) < 1e-14) continue;
          int nc = oq_coup_elems(s, m.N, (int)m.bi[k].size(), m.bi[k].data(), m.bj[k].data(), cols.data());
          for (int c = 0; c < nc; ++c) row[cols[c]] += m.J[k] * 0.5;
        }
      }
      rowlen[r] = (int64_t)row.size();
      for (auto& kv : row) { rc[t].push_back(kv.first); rv[t].push_back(kv.second); }
    }
  }
  A->rowptr.resize(D + 1);
  A->rowptr[0] = 0;
  for (int64_t r = 0; r < D; ++r) A->rowptr[r + 1] = A->rowptr[r] + rowlen[r];
  A->cols.reserve(A->rowptr[D]);
  A->vals.reserve(A->rowptr[D]);
  for (int t = 0; t < nthr; ++t) {
    A->cols.insert(A->cols.end(), rc[t].begin(), rc[t].end());
    A->vals.insert(A->vals.end(), rv[t].begin(), rv[t].end());
  }
  return A;
}
int64_t oq_csr_dim(void* a) { return ((Csr*)a)->D; }
int64_t oq_csr_nnz(void* a) { return (int64_t)((Csr*)a)->cols.size(); }
void oq_csr_copy(void* a, int64_t* rowptr, int64_t* cols, double* vals) {
  Csr& A = *(Csr*)a;
  std::copy(A.rowptr.begin(), A.rowptr.end(), rowptr);
  std::copy(A.cols.begin(), A.cols.end(), cols);
  std::copy(A.vals.begin(), A.vals.end(), vals);
}
void oq_csr_free(void* a) { delete (Csr*)a; }

// y = A x over the assembled CSR (the reference's MatMult on MATMPIAIJ data).
void oq_csr_spmv(void* a, const double* x, double* y, int nthreads) {
  Csr& A = *(Csr*)a;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < A.D; ++r) {
    double acc = 0.0;
    for (int64_t q = A.rowptr[r]; q < A.rowptr[r + 1]; ++q) acc += A.vals[q] * x[A.cols[q]];
    y[r] = acc;
  }
}

// The reference's data layout for timing: PETSc with complex scalars
// (README.md:18) stores complex128 values and int64 columns (24 B per
// nonzero) and multiplies complex vectors.  Built from the assembled CSR.
struct CsrZ {
  int64_t D = 0;
  std::vector<int64_t> rowptr, cols;
  std::vector<std::complex<double>> vals;
};
void* oq_csrz_from(void* a) {
  Csr& A = *(Csr*)a;
  CsrZ* Z = new CsrZ;
  Z->D = A.D;
  Z->rowptr = A.rowptr;
  Z->cols = A.cols;
  Z->vals.assign(A.vals.begin(), A.vals.end());
  return Z;
}
void oq_csrz_free(void* z) { delete (CsrZ*)z; }
// y = A x; x, y: D complex numbers as interleaved (re, im) doubles.
void oq_csrz_spmv(void* z, const double* x, double* y, int nthreads) {
  CsrZ& A = *(CsrZ*)z;
  const std::complex<double>* xc = reinterpret_cast<const std::complex<double>*>(x);
  std::complex<double>* yc = reinterpret_cast<std::complex<double>*>(y);
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < A.D; ++r) {
    std::complex<double> acc = 0.0;
    for (int64_t q = A.rowptr[r]; q < A.rowptr[r + 1]; ++q) acc += A.vals[q] * xc[A.cols[q]];
    yc[r] = acc;
  }
}

// Matrix-free y = H x for rows [row0, row0+nrows): the same row semantics as
// the assembled H (diag + J/2 per antiparallel bond of every kind with
// |J| >= 1e-14), ranks from an exact binomial table.  `hi` may be NULL.  The
// diagonal follows build_diag (+ build_disorder when hi != NULL).  This is
// the CPU H.v baseline; it gathers from the full x.
void oq_hv_matfree(void* h, const double* hi, const double* x, double* y, int64_t row0, int64_t nrows,
                   int nthreads) {
  Model& m = *(Model*)h;
  const BinomTable& T = btab();
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double jh[2] = {std::fabs(m.J[0]) >= 1e-14 ? m.J[0] * 0.5 : 0.0, std::fabs(m.J[1]) >= 1e-14 ? m.J[1] * 0.5 : 0.0};
  double dq[2] = {std::fabs(m.Dz[0]) >= 1e-14 ? m.Dz[0] * 0.25 : 0.0,
                  std::fabs(m.Dz[1]) >= 1e-14 ? m.Dz[1] * 0.25 : 0.0};
  if (hi) { dq[0] = m.Dz[0] * 0.25; dq[1] = m.Dz[1] * 0.25; }
#pragma omp parallel
  {
    int nt = omp_get_num_threads(), t = omp_get_thread_num();
    int64_t a = row0 + nrows * t / nt, b = row0 + nrows * (t + 1) / nt;
    if (a < b) {
      uint64_t s = unrank_ref(a, m.N, m.nup);
      for (int64_t r = a; r < b; ++r) {
        double diag = 0.0, acc = 0.0;
        for (int k = 0; k < 2; ++k) {
          int64_t c = 0;
          const int* bi = m.bi[k].data();
          const int* bj = m.bj[k].data();
          for (size_t q = 0; q < m.bi[k].size(); ++q) {
            uint64_t mask = (1ull << bi[q]) | (1ull << bj[q]);
            uint64_t both = s & mask;
            if (both == 0 || both == mask) { ++c; continue; }
            --c;
            if (jh[k] != 0.0) acc += jh[k] * x[rank_exact(s ^ mask, T)];
          }
          diag += dq[k] * (double)c;
        }
        if (hi) diag += 0.5 * disorder_elem(s, m.N, hi);
        y[r - row0] = diag * x[r] + acc;
        s = gosper_next(s);
      }
    }
  }
}

// Rows of H x for a vector given by a formula of the rank (no D-sized array:
// for sectors whose vectors do not fit in host memory).  The formula has
// period 1000003 * 999983 (> 2^39), so a rank that is wrong by any amount
// up to that (a 32-bit truncation included) reads a different value:
//   x(c) = (((c mod P1) K1 mod P1) + ((c mod P2) K2 mod P2)) * 2^-20 - 1,
// exact in floating point (both terms < 2^20), so any implementation
// reproduces it bit for bit.
double oq_formula_x(int64_t c) {
  const int64_t P1 = 1000003, P2 = 999983, K1 = 7919, K2 = 104729;
  const int64_t n = ((c % P1) * K1) % P1 + ((c % P2) * K2) % P2;
  return (double)n * (1.0 / 1048576.0) - 1.0;
}

void oq_hv_rows_formula(void* h, const double* hi, double* y, int64_t row0, int64_t nrows, int nthreads) {
  Model& m = *(Model*)h;
  const BinomTable& T = btab();
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double jh[2] = {std::fabs(m.J[0]) >= 1e-14 ? m.J[0] * 0.5 : 0.0, std::fabs(m.J[1]) >= 1e-14 ? m.J[1] * 0.5 : 0.0};
  double dq[2] = {std::fabs(m.Dz[0]) >= 1e-14 ? m.Dz[0] * 0.25 : 0.0,
                  std::fabs(m.Dz[1]) >= 1e-14 ? m.Dz[1] * 0.25 : 0.0};
  if (hi) { dq[0] = m.Dz[0] * 0.25; dq[1] = m.Dz[1] * 0.25; }
#pragma omp parallel
  {
    int nt = omp_get_num_threads(), t = omp_get_thread_num();
    int64_t a = row0 + nrows * t / nt, b = row0 + nrows * (t + 1) / nt;
    if (a < b) {
      uint64_t s = unrank_ref(a, m.N, m.nup);
      for (int64_t r = a; r < b; ++r) {
        double diag = 0.0, acc = 0.0;
        for (int k = 0; k < 2; ++k) {
          int64_t c = 0;
          for (size_t q = 0; q < m.bi[k].size(); ++q) {
            uint64_t mask = (1ull << m.bi[k][q]) | (1ull << m.bj[k][q]);
            uint64_t both = s & mask;
            if (both == 0 || both == mask) { ++c; continue; }
            --c;
            if (jh[k] != 0.0) acc += jh[k] * oq_formula_x((int64_t)rank_exact(s ^ mask, T));
          }
          diag += dq[k] * (double)c;
        }
        if (hi) diag += 0.5 * disorder_elem(s, m.N, hi);
        y[r - row0] = diag * oq_formula_x(r) + acc;
        s = gosper_next(s);
      }
    }
  }
}

// --- Observables (src/spinOperators/correlation.cpp, spinOPft.tmpl_func.hpp) --

// SzSz::ExpectVal: <psi| 0.25*(+1 parallel / -1 antiparallel) |psi>
// (correlation.cpp:86-139 with Correlation::ExpectVal :38-70).
double oq_szsz(int N, int nup, const double* psi, int i, int j) {
  int64_t D = oq_compute_size(N, nup);
  double acc = 0.0;
  uint64_t s = unrank_ref(0, N, nup);
  for (int64_t r = 0; r < D; ++r) {
    double v = psi[r] * 0.25;
    uint64_t a = (s >> i) & 1, b = (s >> j) & 1;
    v *= (a ^ b) ? -1.0 : 1.0;
    acc += psi[r] * v;
    s = gosper_next(s);
  }
  return acc;
}

// SpSm (plus=1) / SmSp (plus=0) ::ExpectVal: the one-nonzero-per-row operator
// of correlation.cpp:154-224 / :237-261 applied to psi, then <psi|.>.
double oq_spsm(int N, int nup, const double* psi, int i, int j, int plus) {
  int64_t D = oq_compute_size(N, nup);
  double acc = 0.0;
  uint64_t s = unrank_ref(0, N, nup);
  for (int64_t r = 0; r < D; ++r) {
    uint64_t a = (s >> i) & 1, b = (s >> j) & 1;
    bool sw = false; int64_t col = 0;
    if (i == j) {
      if ((plus && a == 1) || (!plus && a == 0)) { col = rank_ref(s, N); sw = true; }
    } else if ((plus && a == 0 && b == 1) || (!plus && a == 1 && b == 0)) {
      uint64_t t = s ^ (1ull << i);
      t ^= (1ull << j);
      col = rank_ref(t, N); sw = true;
    }
    if (sw) acc += psi[r] * psi[col];
    s = gosper_next(s);
  }
  return acc;
}

// Sqz::OpOnStateVector (spinOPft.tmpl_func.hpp:28-60): rhs = 0.5 * (+-1) psi.
void oq_sqz_apply(int N, int nup, const double* psi, int spin, double* rhs) {
  int64_t D = oq_compute_size(N, nup);
  uint64_t s = unrank_ref(0, N, nup);
  for (int64_t r = 0; r < D; ++r) {
    double v = ((s >> spin) & 1) ? 1.0 : -1.0;
    rhs[r] = psi[r] * 0.5 * v;
    s = gosper_next(s);
  }
}

// Sqp::OpOnStateVector (spinOPft.tmpl_func.hpp:91-143): S^+_spin from the
// sector with nup up-spins into the sector with nup+1 (scatter by rank).
void oq_sqp_apply(int N, int nup, const double* psi, int spin, double* rhs) {
  int64_t D0 = oq_compute_size(N, nup), D1 = oq_compute_size(N, nup + 1);
  for (int64_t r = 0; r < D1; ++r) rhs[r] = 0.0;
  uint64_t s = unrank_ref(0, N, nup);
  for (int64_t r = 0; r < D0; ++r) {
    if (!((s >> spin) & 1)) rhs[rank_ref(s ^ (1ull << spin), N)] += psi[r];
    s = gosper_next(s);
  }
}

// --- Tools::RandomDisorder (src/tools/random_disorder.cpp:33-59) -----------
// Uses libstdc++'s <random> engines and distributions, as the reference does.
void oq_rd_seeds(int rep, int64_t* seeds) {
  std::uniform_int_distribution<long> dist(0, 1000);
  std::mt19937_64 gen(1);
  for (int i = 0; i < rep; ++i) seeds[i] = dist(gen);
}
void oq_rd_hi(int64_t seed, int N, double lo, double hi_, double* out) {
  std::mt19937_64 gen((uint64_t)seed);
  std::uniform_real_distribution<double> dist(lo, hi_);
  for (int i = 0; i < N; ++i) out[i] = dist(gen);
}

int oq_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
