// lsqed_internal.hpp -- shared host/device definitions of liblsqed.so.
//
// Sector layout: a basis state is an N-bit string (bit i = spin i up) with
// nup set bits; its index is the colex combinadic rank
//   r(s) = sum_{k=1..nup} C(p_k, k),  p_k = position of the k-th set bit,
// which is the numeric order of the states (src/basis.cpp:205-213,
// src/hamiltonian.cpp:9-27).  Nothing of size D is stored: every kernel
// unranks its own row and ranks flipped states from a chunk table:
//
//   the N bits are cut into chunks of W bits; for chunk c, chunk value v and
//   k = number of set bits below the chunk,
//     T[c][k][v] = sum_{l-th set bit of v, at chunk position q} C(c*W + q, k + l)
//   so that r(s) = sum_c T[c][k_c(s)][v_c(s)].
//
// Flipping an antiparallel bond (lo < hi) changes only the chunks clo..chi:
// the chunk values at clo and chi, and k_c by -1 (particle moved up) or +1
// for clo < c <= chi.  The new rank is r + sum_{c=clo..chi} (T'[c] - T[c]),
// i.e. (chi-clo+1) table lookups per flip, with the old terms kept in
// registers.  Only rows k in [kmin_c, kmax_c] reachable by sector states are
// stored: N=32, nup=16 needs 36 rows x 256 x 4 B = 36.9 KB (staged in LDS).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lsqed.h"

namespace lsqed {

constexpr int kMaxChunks = 11;   // W=6 chunks for N up to 64
constexpr int kMaxBonds = 512;   // per kind

// Host-built description of a sector's rank tables.
struct SectorTables {
  int N = 0, nup = 0, W = 8, NC = 0;
  bool wide = false;            // 64-bit states/ranks (D >= 2^32 or N > 32)
  int rowbase[kMaxChunks] = {}; // first table row of chunk c
  int kmin[kMaxChunks] = {};
  int kmax[kMaxChunks] = {};
  int nrows = 0;                // table rows (each 2^W entries)
  std::vector<uint64_t> entries;  // nrows << W
  std::vector<uint64_t> binom;    // C(p, k) for p < N, k <= nup: [p*(nup+1) + k]
};

void build_sector_tables(int N, int nup, SectorTables& t);
uint64_t binom_u64(int n, int k);
int64_t sector_dim(int N, int nup);

// FNV-1a over bytes (fingerprints of generated kernel sources and launch shapes).
inline uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull) {
  const unsigned char* p = (const unsigned char*)data;
  for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
  return h;
}
template <typename T>
inline uint64_t fnv1a_val(const T& v, uint64_t h) { return fnv1a(&v, sizeof(T), h); }

// Packed bond: lo | hi << 8 (lo < hi, positions of the two sites).
inline uint32_t pack_bond(int a, int b) {
  int lo = a < b ? a : b, hi = a < b ? b : a;
  return (uint32_t)lo | ((uint32_t)hi << 8);
}

}  // namespace lsqed

// The handle behind lsqed_ham* (one model: lattice bonds, couplings, sector).
struct lsqed_ham {
  int N = 0, nup = 0;
  int64_t D = 0;
  double J[2] = {0, 0}, Dz[2] = {0, 0};
  std::vector<int> bi[2], bj[2];     // reference emission order, per kind
  int build = LSQED_BUILD_DIAG | LSQED_BUILD_OFFDIAG;
  std::vector<double> fields;        // N, when LSQED_BUILD_DISORDER
  lsqed::SectorTables tab;
  // device copies
  void* d_table = nullptr;   // uint32 or uint64 entries
  void* d_binom = nullptr;
  uint32_t* d_bonds = nullptr;  // kind 0 then kind 1, packed
  uint32_t* d_bonds_ordered = nullptr;  // same, but (bi, bj) in emission order: bi | bj << 8
  double* d_fields = nullptr;
  double* d_scratch = nullptr;   // reduction partials
  size_t scratch_doubles = 0;
  int device = 0;
  int hv_variant = 0;          // 0 auto, 1 generic chunk-table, 2 row split, 3 row split specialised (hipRTC)
  void* split = nullptr;       // row-split plan of the single-model kernels (hv_split.hip), built on first use
  void* split_b = nullptr;     // row-split plan of the batched kernel (its own row split)
  void* jit = nullptr;         // specialised row-split module (hv_jit.cpp), cached per bond structure
  void* jit_c = nullptr;       // its complex128 form (lsqed_hv_c128)
  bool jit_failed = false;
};

namespace lsqed {

// Error handling (thread-local last error, returned as negative codes).
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define LSQED_HIP(call)                                                         \
  do {                                                                          \
    hipError_t _e = (call);                                                     \
    if (_e != hipSuccess)                                                       \
      return ::lsqed::fail(LSQED_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Kernel-side parameter block for the row kernels (H.v, diag, export, rank).
struct RowParams {
  const void* table;   // chunk table (uint32 or uint64)
  const void* binom;   // binomials [p*(nup+1)+k]
  const uint32_t* bonds;  // packed (lo|hi<<8), kind 0 then kind 1
  const double* fields;   // N doubles or null
  int64_t D;
  int N, nup;
  int nb[2];
  int table_entries, binom_entries;
  int rowbase[kMaxChunks];
  int kmin[kMaxChunks];
  double jh[2];        // J_k * 0.5 applied per flip (0: kind skipped)
  double dq[2];        // D_k * 0.25
  int flip[2];         // flips of kind k are part of H
  int diag_mode;       // 0: build_diag rules, 1: build_disorder expression, 2: no diagonal
  int dz_on[2];        // |D_k| >= 1e-14 (build_diag applies the kind)
};

RowParams make_row_params(const lsqed_ham* h);
int ensure_scratch(lsqed_ham* h, size_t doubles);

// Launch helpers implemented in hamiltonian.hip (used by krylov.hip).
// hv_launch_grid() is the number of blocks launch_hv() uses, i.e. the number
// of partial sums it writes when dot_partials != null.
int hv_launch_grid(lsqed_ham* h);
bool hv_plan_ready(const lsqed_ham* h);
int launch_hv(lsqed_ham* h, const double* x, double* y, double* dot_partials, hipStream_t s);

// Row-split kernel (hv_split.hip).  The unspecialised kernel handles
// N <= 32 (split_supported); the tables (split plan) and the specialised
// kernel go to N <= 40 with 64-bit row bases above 32.
struct SplitTables {
  const uint16_t *U, *RL0, *RL1;
  const uint32_t *binom, *tasks;
  const uint64_t* binom64;
  int ntasks;
  const uint32_t* tasks_ord;  // the items of one H.v in cache-reuse order (split_grid blocks)
  int ntasks_ord;
  const uint32_t* rows_ord;  // whole rows in the same order (one block per row)
  int nrows_ord;
  const uint32_t* rows;      // whole rows in h order, no padding
  int nrows;
  const int* offU;
  uint64_t order_hash;       // of the item and row orders
};
bool split_supported(const lsqed_ham* h);
int split_low_bits(const lsqed_ham* h, bool batch = false);
bool split_plan_supported(const lsqed_ham* h);
int split_tables(lsqed_ham* h, SplitTables& t, bool batch = false);
int split_plan_create(lsqed_ham* h, bool batch = false);
void split_plan_destroy(lsqed_ham* h);
int split_grid(const lsqed_ham* h);
int launch_hv_split(lsqed_ham* h, const double* x, double* y, double* dot_part, int grid, hipStream_t s);
bool split_jit_fits(const lsqed_ham* h, bool batch = false);

// Site renumbering for the device-resident eigensolver (relabel.hip):
// perm[old site] = new site when it puts more flipping bonds among the row
// split's low sites; the renumbered copy of a model, its synchronisation with
// the model, and the gather of a renumbered vector back into h's basis order.
bool relabel_plan(const lsqed_ham* h, std::vector<int>& perm);
int relabel_handle(const lsqed_ham* h, const std::vector<int>& perm, lsqed_ham** out);
int relabel_sync(const lsqed_ham* h, lsqed_ham* hk, const std::vector<int>& perm);
int relabel_gather(const lsqed_ham* h, const int* d_perm, const double* src, double* dst, hipStream_t s);

// Specialised row-split kernel (hv_jit.cpp).
bool jit_enabled();
int jit_prepare(lsqed_ham* h);
int launch_hv_jit(lsqed_ham* h, const double* x, double* y, double* dot_part, int grid, hipStream_t s);
int launch_hv_jit_c128(lsqed_ham* h, const double* x, double* y, hipStream_t s);
int launch_hv_c128(lsqed_ham* h, const double* x, double* y, hipStream_t s);
int jit_grid(lsqed_ham* h);
uint64_t jit_source_hash(const lsqed_ham* h);  // of the prepared module's generated source (0: none)
// Generated source of the specialised kernel (batch: the per-problem-coupling
// variant; cplx: complex128 entries).  With lowtab non-null and the source
// using low-bond records (LOWTAB), *lowtab receives them (hv_jit.cpp).
std::string jit_source(const lsqed_ham* h, bool batch, bool cplx, std::vector<uint32_t>* lowtab = nullptr);
int jit_threads(const lsqed_ham* h);

// Batch variant of the specialised kernel (hv_jit.cpp): problems sharing a
// bond structure and build mode form a group; one launch covers the group.
struct BatchJitLaunch {
  const double* x;
  double* y;
  double* dot;            // per (problem of the group, row-split task) partial of x.y, or null
  const int64_t* poff;    // per batch problem
  const double* pcoup;    // per batch problem: jh0, jh1, dq0, dq1
  const double* pfields;  // per batch problem: 32 doubles
  const int* active;      // per batch problem
  const int* gidx;        // batch indices of the group's problems
  int nprob;              // problems in the group
};
int batch_jit_prepare(lsqed_ham* h, void** module, std::string* key);
int batch_task_count(void* module, lsqed_ham* h, int* ntasks);
int launch_hv_batch_jit(void* module, lsqed_ham* h0, const BatchJitLaunch& L, int grid, hipStream_t s);

// Reductions (krylov.hip)
int device_dot(lsqed_ham* h, const double* x, const double* y, int64_t n, double* out_host, hipStream_t s);

}  // namespace lsqed
