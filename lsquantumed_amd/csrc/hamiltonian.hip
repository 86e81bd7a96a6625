// hamiltonian.hip -- matrix-free Sz-sector Hamiltonian on MI355X (gfx950).
//
// Replaces the assembled MATMPIAIJ of Hamiltonian<L> (include/hamiltonian.hpp,
// include/hamiltonian.tmpl_func.hpp, src/hamiltonian.cpp) with a gather
// kernel, one lane per basis state:
//   y[r] = diag(s_r) x[r] + sum_{antiparallel bonds b of kind k} J_k/2 x[rank(s_r ^ m_b)]
// The row's state s_r is unranked from r; flipped states are ranked with the
// chunk tables of lsqed_internal.hpp staged in LDS, so no state list and no
// matrix is stored.  The diagonal follows build_diag (tmpl:129-225) or, with
// fields, build_disorder (tmpl:97-125), evaluated in the reference's order
// without FMA contraction so that lsqed_diag is bit-exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "lsqed_internal.hpp"

namespace lsqed {

// ------------------------------------------------------------ host tables --
uint64_t binom_u64(int n, int k) {
  if (k < 0 || n < 0 || k > n) return 0;
  if (k > n - k) k = n - k;
  unsigned __int128 r = 1;
  for (int i = 1; i <= k; ++i) r = r * (unsigned)(n - k + i) / (unsigned)i;
  return (uint64_t)r;
}

int64_t sector_dim(int N, int nup) { return (int64_t)binom_u64(N, nup); }

void build_sector_tables(int N, int nup, SectorTables& t) {
  t.N = N;
  t.nup = nup;
  t.wide = N > 32;
  t.W = t.wide ? 6 : 8;
  t.NC = (N + t.W - 1) / t.W;
  t.nrows = 0;
  for (int c = 0; c < t.NC; ++c) {
    int below = c * t.W;
    t.kmin[c] = std::max(0, nup - (N - below));
    t.kmax[c] = std::min(nup, below);
    t.rowbase[c] = t.nrows;
    t.nrows += t.kmax[c] - t.kmin[c] + 1;
  }
  const int V = 1 << t.W;
  t.entries.assign((size_t)t.nrows * V, 0);
  for (int c = 0; c < t.NC; ++c) {
    int below = c * t.W;
    int width = std::min(t.W, N - below);
    for (int k = t.kmin[c]; k <= t.kmax[c]; ++k)
      for (int v = 0; v < V; ++v) {
        if (v >> width) continue;  // bits beyond N never occur
        uint64_t val = 0;
        int l = 0;
        for (int q = 0; q < width; ++q)
          if ((v >> q) & 1) {
            ++l;
            val += binom_u64(below + q, k + l);
          }
        t.entries[(size_t)(t.rowbase[c] + k - t.kmin[c]) * V + v] = val;
      }
  }
  t.binom.assign((size_t)N * (nup + 1), 0);
  for (int p = 0; p < N; ++p)
    for (int k = 0; k <= nup; ++k) t.binom[(size_t)p * (nup + 1) + k] = binom_u64(p, k);
}

RowParams make_row_params(const lsqed_ham* h) {
  RowParams p;
  std::memset(&p, 0, sizeof(p));
  p.table = h->d_table;
  p.binom = h->d_binom;
  p.bonds = h->d_bonds;
  p.fields = h->d_fields;
  p.D = h->D;
  p.N = h->N;
  p.nup = h->nup;
  p.nb[0] = (int)h->bi[0].size();
  p.nb[1] = (int)h->bi[1].size();
  p.table_entries = (int)h->tab.entries.size();
  p.binom_entries = (int)h->tab.binom.size();
  for (int c = 0; c < kMaxChunks; ++c) {
    p.rowbase[c] = h->tab.rowbase[c];
    p.kmin[c] = h->tab.kmin[c];
  }
  const bool off = (h->build & LSQED_BUILD_OFFDIAG) != 0;
  for (int k = 0; k < 2; ++k) {
    // build_off_diag skips a kind with |J| < 1e-14 (tmpl:243, :278).
    p.flip[k] = off && std::fabs(h->J[k]) >= 1e-14;
    p.jh[k] = p.flip[k] ? h->J[k] * 0.5 : 0.0;
    p.dq[k] = h->Dz[k] * 0.25;
    p.dz_on[k] = std::fabs(h->Dz[k]) >= 1e-14;
  }
  if ((h->build & LSQED_BUILD_DISORDER) && h->d_fields)
    p.diag_mode = 1;
  else if (h->build & LSQED_BUILD_DIAG)
    p.diag_mode = 0;
  else
    p.diag_mode = 2;
  return p;
}

}  // namespace lsqed

#include "sector_device.hpp"

namespace lsqed {

constexpr int kHvThreads = 256;
constexpr int kBatch = 8;

// --------------------------------------------------------------- H.v kernel --
template <typename S, typename R, int W, int NC, bool DOT, bool CPLX = false>
__global__ __launch_bounds__(kHvThreads) void hv_kernel(RowParams p, const R* __restrict__ gtab,
                                                        const R* __restrict__ gbinom,
                                                        const uint32_t* __restrict__ bonds,
                                                        const double* __restrict__ x, double* __restrict__ y,
                                                        double* __restrict__ dot_part, int64_t ntiles) {
  extern __shared__ __align__(16) unsigned char smem[];
  R* s_tab = reinterpret_cast<R*>(smem);
  R* s_binom = s_tab + p.table_entries;
  double* s_f = reinterpret_cast<double*>(smem + (((size_t)(p.table_entries + p.binom_entries) * sizeof(R) + 15) & ~(size_t)15));
  double* s_red = s_f + 64;
  for (int i = threadIdx.x; i < p.table_entries; i += blockDim.x) s_tab[i] = gtab[i];
  for (int i = threadIdx.x; i < p.binom_entries; i += blockDim.x) s_binom[i] = gbinom[i];
  if (p.diag_mode == 1)
    for (int i = threadIdx.x; i < p.N; i += blockDim.x) s_f[i] = p.fields[i];
  __syncthreads();

  const int64_t G = gridDim.x;
  const int64_t lb = xcd_block();
  const int64_t t0 = lb * ntiles / G, t1 = (lb + 1) * ntiles / G;
  double dotacc = 0.0;

  for (int64_t tile = t0; tile < t1; ++tile) {
    const int64_t row = tile * kHvThreads + threadIdx.x;
    if (row >= p.D) continue;
    const S s = unrank_row<S, R>((R)row, p.N, p.nup, s_binom);
    int kc[NC];
    R tc[NC];
    chunk_terms<S, R, W, NC>(s, kc, tc, s_tab, p);
    typedef Entry<CPLX> E;
    typename E::T acc = E::zero();
    int cnt[2] = {0, 0};
#pragma unroll
    for (int kind = 0; kind < 2; ++kind) {
      const uint32_t* __restrict__ bk = bonds + (kind ? p.nb[0] : 0);
      const int nb = p.nb[kind];
      const bool fl = p.flip[kind] != 0;
      const double jh = p.jh[kind];
      int b = 0;
      for (; b + kBatch <= nb; b += kBatch) {
        typename E::T vals[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          const uint32_t bd = __builtin_amdgcn_readfirstlane(bk[b + u]);
          const int lo = bd & 0xff, hi = (bd >> 8) & 0xff;
          const S m = ((S)1 << lo) | ((S)1 << hi);
          const S sm = s & m;
          const bool anti = (sm != 0) && (sm != m);
          cnt[kind] += anti;
          vals[u] = E::zero();
          if (fl && anti) {
            const R rn = flip_rank<S, R, W, NC>((R)row, s, lo, hi, kc, tc, s_tab, p);
            vals[u] = E::load(x, (int64_t)rn);
          }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) madd(acc, jh, vals[u]);
      }
      for (; b < nb; ++b) {
        const uint32_t bd = __builtin_amdgcn_readfirstlane(bk[b]);
        const int lo = bd & 0xff, hi = (bd >> 8) & 0xff;
        const S m = ((S)1 << lo) | ((S)1 << hi);
        const S sm = s & m;
        const bool anti = (sm != 0) && (sm != m);
        cnt[kind] += anti;
        if (fl && anti) {
          const R rn = flip_rank<S, R, W, NC>((R)row, s, lo, hi, kc, tc, s_tab, p);
          madd(acc, jh, E::load(x, (int64_t)rn));
        }
      }
    }
    const int c0 = p.nb[0] - 2 * cnt[0], c1 = p.nb[1] - 2 * cnt[1];
    const double d = diag_value(p, c0, c1, (uint64_t)s, s_f);
    const typename E::T xr = E::load(x, row);
    const typename E::T yr = diag_apply(d, xr, acc);
    E::store(y, row, yr);
    if constexpr (DOT && !CPLX) dotacc += xr * yr;
  }
  if (DOT) {
    double t = block_sum(dotacc, s_red);
    if (threadIdx.x == 0) dot_part[blockIdx.x] = t;
  }
}

// ------------------------------------------------------- auxiliary kernels --
template <typename S, typename R, int W, int NC>
__global__ void diag_kernel(RowParams p, const R* __restrict__ tab, const R* __restrict__ binom,
                            const uint32_t* __restrict__ bonds, double* __restrict__ d, int cplx) {
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < p.D;
       row += (int64_t)gridDim.x * blockDim.x) {
    const S s = unrank_row<S, R>((R)row, p.N, p.nup, binom);
    int cnt[2] = {0, 0};
    for (int kind = 0; kind < 2; ++kind) {
      const uint32_t* bk = bonds + (kind ? p.nb[0] : 0);
      for (int b = 0; b < p.nb[kind]; ++b) {
        const int lo = bk[b] & 0xff, hi = (bk[b] >> 8) & 0xff;
        cnt[kind] += (((s >> lo) ^ (s >> hi)) & 1) ? 1 : 0;
      }
    }
    const double v = diag_value(p, p.nb[0] - 2 * cnt[0], p.nb[1] - 2 * cnt[1], (uint64_t)s, p.fields);
    if (cplx)
      reinterpret_cast<double2*>(d)[row] = make_double2(v, 0.0);
    else
      d[row] = v;
  }
}

template <typename S, typename R, int W, int NC>
__global__ void export_kernel(RowParams p, const R* __restrict__ tab, const R* __restrict__ binom,
                              const uint32_t* __restrict__ ordered, int kind, int64_t row0, int64_t nrows,
                              int max_cols, int64_t* __restrict__ cols, int32_t* __restrict__ ncol) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = row0 + i;
    const S s = unrank_row<S, R>((R)row, p.N, p.nup, binom);
    const uint32_t* bk = ordered + (kind ? p.nb[0] : 0);
    int n = 0;
    for (int b = 0; b < p.nb[kind] && n < max_cols; ++b) {
      const int a = bk[b] & 0xff, c = (bk[b] >> 8) & 0xff;
      if (((s >> a) ^ (s >> c)) & 1) {
        const S sp = s ^ (((S)1 << a) | ((S)1 << c));
        cols[i * max_cols + n] = (int64_t)rank_state<S, R, W, NC>(sp, tab, p);
        ++n;
      }
    }
    ncol[i] = n;
  }
}

template <typename S, typename R, int W, int NC>
__global__ void unrank_kernel(RowParams p, const R* __restrict__ binom, int64_t row0, int64_t n,
                              uint64_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (uint64_t)unrank_row<S, R>((R)(row0 + i), p.N, p.nup, binom);
}

template <typename S, typename R, int W, int NC>
__global__ void rank_kernel(RowParams p, const R* __restrict__ tab, const uint64_t* __restrict__ st, int64_t n,
                            int64_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t s = st[i];
    // A state outside the sector (a bit at or above N, or not nup set bits)
    // has no rank; it would index table rows outside [kmin, kmax].
    const bool in_sector = (p.N >= 64 || (s >> p.N) == 0) && __popcll(s) == p.nup;
    out[i] = in_sector ? (int64_t)rank_state<S, R, W, NC>((S)s, tab, p) : (int64_t)-1;
  }
}

// -------------------------------------------------------------- launchers --
static int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
    cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  }
  return cus;
}

static size_t hv_smem_bytes(const lsqed_ham* h) {
  const size_t esz = h->tab.wide ? 8 : 4;
  size_t b = (h->tab.entries.size() + h->tab.binom.size()) * esz;
  b = (b + 15) & ~(size_t)15;
  return b + (64 + 8) * sizeof(double);
}

int hv_grid(const lsqed_ham* h, int64_t ntiles) {
  // Blocks resident per CU limited by LDS (160 KB per CU on gfx950).
  const size_t sm = hv_smem_bytes(h);
  int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / std::max<size_t>(sm, 1)));
  int64_t g = (int64_t)device_cus() * per_cu * 2;
  g = std::min<int64_t>(g, std::max<int64_t>(ntiles, 1));
  g = ((g + 7) / 8) * 8;  // a multiple of 8 so the XCD remap is a permutation
  return (int)g;
}

template <typename S, typename R, int W, int NC, bool DOT, bool CPLX = false>
static int launch_hv_t(lsqed_ham* h, const double* x, double* y, double* dot_part, int grid, hipStream_t st) {
  RowParams p = make_row_params(h);
  const int64_t ntiles = (h->D + kHvThreads - 1) / kHvThreads;
  const size_t sm = hv_smem_bytes(h);
  auto kfn = hv_kernel<S, R, W, NC, DOT, CPLX>;
  if (sm > 64 * 1024) LSQED_HIP(hipFuncSetAttribute((const void*)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
  hipLaunchKernelGGL(kfn, dim3(grid), dim3(kHvThreads), sm, st, p, (const R*)h->d_table, (const R*)h->d_binom,
                     (const uint32_t*)h->d_bonds, x, y, dot_part, ntiles);
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

// Kernel choice for H.v: 3 = specialised row split (hipRTC), 2 = row split,
// 1 = generic chunk table.
static int hv_kind(lsqed_ham* h) {
  if (h->hv_variant >= 1) return h->hv_variant;
  if (jit_enabled() && !h->jit_failed && split_jit_fits(h)) {
    if (jit_prepare(h) == LSQED_OK) return 3;
    h->jit_failed = true;  // keep an unspecialised kernel
  }
  return split_supported(h) ? 2 : 1;
}

// Whether launch_hv can run without building anything (no hipRTC compile,
// no row-split tables): the kernel is resolved and its plan exists.
bool hv_plan_ready(const lsqed_ham* h) {
  int k = h->hv_variant;
  if (k == 0) {
    if (jit_enabled() && !h->jit_failed && split_jit_fits(h))
      k = h->jit ? 3 : -1;
    else
      k = split_supported(h) ? 2 : 1;
  }
  if (k == 3) return h->jit && h->split;
  if (k == 2) return h->split != nullptr;
  return k == 1;
}

int hv_launch_grid(lsqed_ham* h) {
  const int k = hv_kind(h);
  if (k >= 2) {
    if (split_plan_create(h) != LSQED_OK) return 0;
    return k == 3 ? jit_grid(h) : split_grid(h);
  }
  return hv_grid(h, (h->D + kHvThreads - 1) / kHvThreads);
}

// y = H x, and (when dot_partials != null) per-block partial sums of x.y in
// dot_partials[0 .. hv_launch_grid(h)).
int launch_hv(lsqed_ham* h, const double* x, double* y, double* dot_partials, hipStream_t s) {
  if (h->D == 0) return LSQED_OK;
  const int k = hv_kind(h);
  if (k == 3) {
    int rc = split_plan_create(h);
    if (rc) return rc;
    return launch_hv_jit(h, x, y, dot_partials, jit_grid(h), s);
  }
  if (k == 2) return launch_hv_split(h, x, y, dot_partials, 0, s);
  const int64_t ntiles = (h->D + kHvThreads - 1) / kHvThreads;
  const int grid = hv_grid(h, ntiles);
  if (!h->tab.wide) {
    if (dot_partials) return launch_hv_t<uint32_t, uint32_t, 8, 4, true>(h, x, y, dot_partials, grid, s);
    return launch_hv_t<uint32_t, uint32_t, 8, 4, false>(h, x, y, nullptr, grid, s);
  }
  if (dot_partials) return launch_hv_t<uint64_t, uint64_t, 6, kMaxChunks, true>(h, x, y, dot_partials, grid, s);
  return launch_hv_t<uint64_t, uint64_t, 6, kMaxChunks, false>(h, x, y, nullptr, grid, s);
}

// y = H x on interleaved complex vectors (D complex128 entries each): the
// specialised row split when that is the handle's kernel, else the generic
// chunk-table kernel (the unspecialised row split has no complex form).
int launch_hv_c128(lsqed_ham* h, const double* x, double* y, hipStream_t s) {
  if (h->D == 0) return LSQED_OK;
  if (hv_kind(h) == 3) {
    int rc = split_plan_create(h);
    if (rc) return rc;
    return launch_hv_jit_c128(h, x, y, s);
  }
  const int grid = hv_grid(h, (h->D + kHvThreads - 1) / kHvThreads);
  if (!h->tab.wide) return launch_hv_t<uint32_t, uint32_t, 8, 4, false, true>(h, x, y, nullptr, grid, s);
  return launch_hv_t<uint64_t, uint64_t, 6, kMaxChunks, false, true>(h, x, y, nullptr, grid, s);
}

static int aux_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

// ------------------------------------------------------------ error state --
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int ensure_scratch(lsqed_ham* h, size_t doubles) {
  if (h->scratch_doubles >= doubles) return LSQED_OK;
  if (h->d_scratch) LSQED_HIP(hipFree(h->d_scratch));
  h->d_scratch = nullptr;
  LSQED_HIP(hipMalloc(&h->d_scratch, doubles * sizeof(double)));
  h->scratch_doubles = doubles;
  return LSQED_OK;
}

}  // namespace lsqed

using namespace lsqed;

// =================================================================== C-ABI ==
extern "C" {

const char* lsqed_last_error(void) { return g_last_error.c_str(); }
const char* lsqed_version(void) { return "lsqed-mi355x 0.1 (gfx950)"; }

int64_t lsqed_sector_dim(int nspins, int nup) {
  if (nspins <= 0 || nspins > 62 || nup < 0 || nup > nspins) return 0;
  return sector_dim(nspins, nup);
}

int lsqed_ham_create(const lsqed_params* pr, lsqed_ham** out) {
  if (!pr || !out) return fail(LSQED_ERR_ARG, "null argument");
  *out = nullptr;
  // Environment: N even and positive (src/environment.cpp:31-35); Basis:
  // |mag| <= N/2 (src/basis.cpp:66-67).
  if (pr->nspins <= 0 || pr->nspins % 2 != 0) return fail(LSQED_ERR_ARG, "nspins must be even and positive");
  if (pr->nspins > 62) return fail(LSQED_ERR_ARG, "nspins > 62 is not supported");
  if (pr->nup < 0 || pr->nup > pr->nspins)
    return fail(LSQED_ERR_ARG, "The required constant magnetization has to be <= nspins/2.");
  const int nb[2] = {pr->n_nn, pr->n_nnn};
  const int* bi[2] = {pr->nn_i, pr->nnn_i};
  const int* bj[2] = {pr->nn_j, pr->nnn_j};
  for (int k = 0; k < 2; ++k) {
    if (nb[k] < 0 || nb[k] > kMaxBonds) return fail(LSQED_ERR_ARG, "bond count out of range");
    if (nb[k] > 0 && (!bi[k] || !bj[k])) return fail(LSQED_ERR_ARG, "null bond array");
    for (int b = 0; b < nb[k]; ++b)
      if (bi[k][b] < 0 || bi[k][b] >= pr->nspins || bj[k][b] < 0 || bj[k][b] >= pr->nspins)
        return fail(LSQED_ERR_ARG, "bond site out of range");
  }
  lsqed_ham* h = new lsqed_ham;
  h->N = pr->nspins;
  h->nup = pr->nup;
  h->D = sector_dim(h->N, h->nup);
  h->J[0] = pr->J1;
  h->Dz[0] = pr->D1;
  h->J[1] = pr->J2;
  h->Dz[1] = pr->D2;
  std::vector<uint32_t> packed, ordered;
  for (int k = 0; k < 2; ++k) {
    h->bi[k].assign(bi[k], bi[k] + nb[k]);
    h->bj[k].assign(bj[k], bj[k] + nb[k]);
    for (int b = 0; b < nb[k]; ++b) {
      packed.push_back(pack_bond(bi[k][b], bj[k][b]));
      ordered.push_back((uint32_t)bi[k][b] | ((uint32_t)bj[k][b] << 8));
    }
  }
  build_sector_tables(h->N, h->nup, h->tab);
  auto cleanup = [&](int code, const std::string& m) {
    lsqed_ham_destroy(h);
    return fail(code, m);
  };
  if (hipGetDevice(&h->device) != hipSuccess) return cleanup(LSQED_ERR_HIP, "hipGetDevice failed");
  const size_t esz = h->tab.wide ? 8 : 4;
  std::vector<uint32_t> t32, b32;
  const void* tsrc = h->tab.entries.data();
  const void* bsrc = h->tab.binom.data();
  if (!h->tab.wide) {
    t32.assign(h->tab.entries.begin(), h->tab.entries.end());
    b32.assign(h->tab.binom.begin(), h->tab.binom.end());
    tsrc = t32.data();
    bsrc = b32.data();
  }
  if (hipMalloc(&h->d_table, std::max<size_t>(1, h->tab.entries.size() * esz)) != hipSuccess ||
      hipMalloc(&h->d_binom, std::max<size_t>(1, h->tab.binom.size() * esz)) != hipSuccess ||
      hipMalloc((void**)&h->d_bonds, std::max<size_t>(1, packed.size() * 4)) != hipSuccess ||
      hipMalloc((void**)&h->d_bonds_ordered, std::max<size_t>(1, ordered.size() * 4)) != hipSuccess)
    return cleanup(LSQED_ERR_HIP, "hipMalloc failed for rank tables");
  if (hipMemcpy(h->d_table, tsrc, h->tab.entries.size() * esz, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_binom, bsrc, h->tab.binom.size() * esz, hipMemcpyHostToDevice) != hipSuccess ||
      (!packed.empty() && hipMemcpy(h->d_bonds, packed.data(), packed.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
      (!ordered.empty() &&
       hipMemcpy(h->d_bonds_ordered, ordered.data(), ordered.size() * 4, hipMemcpyHostToDevice) != hipSuccess))
    return cleanup(LSQED_ERR_HIP, "hipMemcpy failed for rank tables");
  *out = h;
  return LSQED_OK;
}

int lsqed_ham_destroy(lsqed_ham* h) {
  if (!h) return LSQED_OK;
  (void)hipFree(h->d_table);
  (void)hipFree(h->d_binom);
  (void)hipFree(h->d_bonds);
  (void)hipFree(h->d_bonds_ordered);
  (void)hipFree(h->d_fields);
  (void)hipFree(h->d_scratch);
  split_plan_destroy(h);
  delete h;
  return LSQED_OK;
}

int64_t lsqed_ham_dim(const lsqed_ham* h) { return h ? h->D : 0; }
int lsqed_ham_nspins(const lsqed_ham* h) { return h ? h->N : 0; }
int lsqed_ham_nup(const lsqed_ham* h) { return h ? h->nup : 0; }

int lsqed_ham_get_params(const lsqed_ham* h, double* J1, double* D1, double* J2, double* D2, int* flags, int* n_nn,
                         int* n_nnn) {
  if (!h) return fail(LSQED_ERR_ARG, "null handle");
  if (J1) *J1 = h->J[0];
  if (D1) *D1 = h->Dz[0];
  if (J2) *J2 = h->J[1];
  if (D2) *D2 = h->Dz[1];
  if (flags) *flags = h->build;
  if (n_nn) *n_nn = (int)h->bi[0].size();
  if (n_nnn) *n_nnn = (int)h->bi[1].size();
  return LSQED_OK;
}

int lsqed_ham_set_build(lsqed_ham* h, int flags) {
  if (!h) return fail(LSQED_ERR_ARG, "null handle");
  if (flags & ~(LSQED_BUILD_DIAG | LSQED_BUILD_OFFDIAG | LSQED_BUILD_DISORDER))
    return fail(LSQED_ERR_ARG, "unknown build flag");
  if ((flags & LSQED_BUILD_DISORDER) && !h->d_fields)
    return fail(LSQED_ERR_STATE, "build_disorder needs fields (lsqed_ham_set_fields)");
  if (flags != h->build) h->jit = h->jit_c = nullptr;  // the specialised kernel bakes in the build flags
  h->build = flags;
  return LSQED_OK;
}

int lsqed_ham_set_fields(lsqed_ham* h, const double* hi) {
  if (!h) return fail(LSQED_ERR_ARG, "null handle");
  if (!hi) {
    h->fields.clear();
    if (h->build & LSQED_BUILD_DISORDER) h->jit = h->jit_c = nullptr;
    h->build &= ~LSQED_BUILD_DISORDER;
    return LSQED_OK;
  }
  h->fields.assign(hi, hi + h->N);
  if (!h->d_fields) LSQED_HIP(hipMalloc((void**)&h->d_fields, h->N * sizeof(double)));
  LSQED_HIP(hipMemcpy(h->d_fields, hi, h->N * sizeof(double), hipMemcpyHostToDevice));
  if (!(h->build & LSQED_BUILD_DISORDER)) h->jit = h->jit_c = nullptr;
  h->build |= LSQED_BUILD_DISORDER;
  return LSQED_OK;
}

int lsqed_ham_set_couplings(lsqed_ham* h, double J1, double D1, double J2, double D2) {
  if (!h) return fail(LSQED_ERR_ARG, "null handle");
  if (!std::isfinite(J1) || !std::isfinite(D1) || !std::isfinite(J2) || !std::isfinite(D2))
    return fail(LSQED_ERR_ARG, "couplings must be finite");
  h->J[0] = J1;
  h->Dz[0] = D1;
  h->J[1] = J2;
  h->Dz[1] = D2;
  h->jit = h->jit_c = nullptr;  // the specialised kernel bakes in the couplings
  return LSQED_OK;
}

int lsqed_hv(lsqed_ham* h, const double* x, double* y, lsqed_stream_t stream) {
  if (!h || !x || !y) return fail(LSQED_ERR_ARG, "null argument");
  if (x == y) return fail(LSQED_ERR_ARG, "x and y must not alias");
  return launch_hv(h, x, y, nullptr, (hipStream_t)stream);
}

int lsqed_hv_c128(lsqed_ham* h, const double* x, double* y, lsqed_stream_t stream) {
  if (!h || !x || !y) return fail(LSQED_ERR_ARG, "null argument");
  if (x == y) return fail(LSQED_ERR_ARG, "x and y must not alias");
  return launch_hv_c128(h, x, y, (hipStream_t)stream);
}

int lsqed_ham_get_hv_kernel(lsqed_ham* h) { return h ? hv_kind(h) : 0; }

int lsqed_ham_hv_fingerprint(lsqed_ham* h, uint64_t* fp) {
  if (!h || !fp) return fail(LSQED_ERR_ARG, "null argument");
  const int k = hv_kind(h);
  const int grid = hv_launch_grid(h);
  if (grid <= 0) return fail(LSQED_ERR_HIP, "H.v plan setup failed");
  uint64_t f = fnv1a_val(k, fnv1a_val(grid, fnv1a_val(h->N, fnv1a_val(h->nup, fnv1a_val(h->build, 1469598103934665603ull)))));
  if (k >= 2) {
    // the order in which rows are dealt to blocks decides which gathers meet in cache
    SplitTables t;
    if (split_tables(h, t) != LSQED_OK) return fail(LSQED_ERR_HIP, "row-split plan setup failed");
    f = fnv1a_val(t.order_hash, f);
  }
  if (k == 3) {
    // the generated source holds the bonds, couplings, build mode and tuning
    f = fnv1a_val(jit_source_hash(h), fnv1a_val(jit_threads(h), f));
  } else {
    for (int kind = 0; kind < 2; ++kind) {
      f = fnv1a(h->bi[kind].data(), h->bi[kind].size() * sizeof(int), f);
      f = fnv1a(h->bj[kind].data(), h->bj[kind].size() * sizeof(int), f);
      f = fnv1a_val(h->J[kind], fnv1a_val(h->Dz[kind], f));
    }
    const char* v = lsqed_version();
    f = fnv1a(v, std::strlen(v), f);
  }
  *fp = f;
  return LSQED_OK;
}

int lsqed_ham_set_hv_kernel(lsqed_ham* h, int variant) {
  if (!h) return fail(LSQED_ERR_ARG, "null handle");
  if (variant < 0 || variant > 3)
    return fail(LSQED_ERR_ARG, "variant must be 0 (auto), 1 (generic), 2 (row split) or 3 (specialised row split)");
  if (variant == 2 && !split_supported(h)) return fail(LSQED_ERR_ARG, "the unspecialised row split needs N <= 32");
  if (variant == 3) {
    if (!split_jit_fits(h))
      return fail(LSQED_ERR_ARG, "the specialised row split needs N <= 40 and at most 256 bonds between high sites");
    int rc = jit_prepare(h);
    if (rc) return rc;
  }
  h->hv_variant = variant;
  return LSQED_OK;
}

static int diag_launch(lsqed_ham* h, double* d, int cplx, hipStream_t stream) {
  if (!h || !d) return fail(LSQED_ERR_ARG, "null argument");
  if (h->D == 0) return LSQED_OK;
  RowParams p = make_row_params(h);
  const int g = aux_grid(h->D);
  if (!h->tab.wide)
    hipLaunchKernelGGL((diag_kernel<uint32_t, uint32_t, 8, 4>), dim3(g), dim3(256), 0, stream, p,
                       (const uint32_t*)h->d_table, (const uint32_t*)h->d_binom, h->d_bonds, d, cplx);
  else
    hipLaunchKernelGGL((diag_kernel<uint64_t, uint64_t, 6, kMaxChunks>), dim3(g), dim3(256), 0, stream, p,
                       (const uint64_t*)h->d_table, (const uint64_t*)h->d_binom, h->d_bonds, d, cplx);
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

int lsqed_diag(lsqed_ham* h, double* d, lsqed_stream_t stream) { return diag_launch(h, d, 0, (hipStream_t)stream); }

int lsqed_diag_c128(lsqed_ham* h, double* d, lsqed_stream_t stream) {
  return diag_launch(h, d, 1, (hipStream_t)stream);
}

int lsqed_export_cols(lsqed_ham* h, int kind, int64_t row0, int64_t nrows, int max_cols, int64_t* cols,
                      int32_t* ncol, lsqed_stream_t stream) {
  if (!h || !cols || !ncol) return fail(LSQED_ERR_ARG, "null argument");
  if (kind < 0 || kind > 1) return fail(LSQED_ERR_ARG, "kind must be 0 (nn) or 1 (nnn)");
  if (row0 < 0 || nrows < 0 || row0 + nrows > h->D) return fail(LSQED_ERR_ARG, "row range out of bounds");
  if (max_cols < (int)h->bi[kind].size()) return fail(LSQED_ERR_ARG, "max_cols smaller than the bond count");
  if (nrows == 0) return LSQED_OK;
  RowParams p = make_row_params(h);
  const int g = aux_grid(nrows);
  if (!h->tab.wide)
    hipLaunchKernelGGL((export_kernel<uint32_t, uint32_t, 8, 4>), dim3(g), dim3(256), 0, (hipStream_t)stream, p,
                       (const uint32_t*)h->d_table, (const uint32_t*)h->d_binom, h->d_bonds_ordered, kind, row0,
                       nrows, max_cols, cols, ncol);
  else
    hipLaunchKernelGGL((export_kernel<uint64_t, uint64_t, 6, kMaxChunks>), dim3(g), dim3(256), 0,
                       (hipStream_t)stream, p, (const uint64_t*)h->d_table, (const uint64_t*)h->d_binom,
                       h->d_bonds_ordered, kind, row0, nrows, max_cols, cols, ncol);
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

int lsqed_unrank(lsqed_ham* h, int64_t row0, int64_t n, uint64_t* states, lsqed_stream_t stream) {
  if (!h || !states) return fail(LSQED_ERR_ARG, "null argument");
  if (row0 < 0 || n < 0 || row0 + n > h->D) return fail(LSQED_ERR_ARG, "row range out of bounds");
  if (n == 0) return LSQED_OK;
  RowParams p = make_row_params(h);
  const int g = aux_grid(n);
  if (!h->tab.wide)
    hipLaunchKernelGGL((unrank_kernel<uint32_t, uint32_t, 8, 4>), dim3(g), dim3(256), 0, (hipStream_t)stream, p,
                       (const uint32_t*)h->d_binom, row0, n, states);
  else
    hipLaunchKernelGGL((unrank_kernel<uint64_t, uint64_t, 6, kMaxChunks>), dim3(g), dim3(256), 0,
                       (hipStream_t)stream, p, (const uint64_t*)h->d_binom, row0, n, states);
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

int lsqed_rank(lsqed_ham* h, const uint64_t* states, int64_t n, int64_t* ranks, lsqed_stream_t stream) {
  if (!h || !states || !ranks) return fail(LSQED_ERR_ARG, "null argument");
  if (n < 0) return fail(LSQED_ERR_ARG, "negative count");
  if (n == 0) return LSQED_OK;
  RowParams p = make_row_params(h);
  const int g = aux_grid(n);
  if (!h->tab.wide)
    hipLaunchKernelGGL((rank_kernel<uint32_t, uint32_t, 8, 4>), dim3(g), dim3(256), 0, (hipStream_t)stream, p,
                       (const uint32_t*)h->d_table, states, n, ranks);
  else
    hipLaunchKernelGGL((rank_kernel<uint64_t, uint64_t, 6, kMaxChunks>), dim3(g), dim3(256), 0,
                       (hipStream_t)stream, p, (const uint64_t*)h->d_table, states, n, ranks);
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

}  // extern "C"
