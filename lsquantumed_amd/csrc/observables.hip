// observables.hip -- expectation values on a device-resident state.
//
// Replaces the per-pair operator builds of the reference
// (src/spinOperators/correlation.cpp:86-299: a D x D MATMPIAIJ per (i, j),
// then MatMult + VecDot) and the DSF amplitude loop
// (include/phys.tmpl_func.hpp:166-186, include/spinOperators/spinOPft.tmpl_func.hpp)
// with single passes over psi, one lane per basis state:
//   - pairs kernel: for every i < j, the parallel weight sum_s psi^2 [b_i == b_j]
//     and the flip overlap sum_{b_i != b_j} psi(s) psi(s ^ m_ij), plus
//     m_i = sum_s psi^2 b_i and P = sum_s psi^2.  From these
//       <S^z_i S^z_j> = (2 par_ij - P) / 4,  <S^+_i S^-_j> = flip_ij / 2,
//       <S^+_i S^-_i> = m_i,  <S^z_i> = m_i - P/2.
//   - DSF kernels: a[n][i] = <n| S_i |0> for nev states at once.
// Sums are wavefront shuffles, per-wave LDS slots, per-block partials and a
// fixed-order final reduction (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "lsqed_internal.hpp"
#include "sector_device.hpp"

namespace lsqed {

constexpr int kObsThreads = 256;

__device__ __forceinline__ double obs_wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;  // every lane holds the total
}

// Quantities per block: Q = 2 * npairs + N + 1, laid out
//   [0, npairs)           parallel weight of pair q
//   [npairs, 2 npairs)    flip overlap of pair q
//   [2 npairs, +N)        m_i
//   [2 npairs + N]        P
template <typename S, typename R, int W, int NC>
__global__ __launch_bounds__(kObsThreads) void k_obs_pairs(RowParams p, const R* __restrict__ gtab,
                                                           const R* __restrict__ gbinom,
                                                           const double* __restrict__ psi, int Q,
                                                           double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* acc = reinterpret_cast<double*>(smem);  // [4][Q]
  R* s_tab = reinterpret_cast<R*>(acc + 4 * Q);
  R* s_binom = s_tab + p.table_entries;
  for (int i = threadIdx.x; i < p.table_entries; i += blockDim.x) s_tab[i] = gtab[i];
  for (int i = threadIdx.x; i < p.binom_entries; i += blockDim.x) s_binom[i] = gbinom[i];
  for (int i = threadIdx.x; i < 4 * Q; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double* my = acc + wid * Q;
  const int N = p.N;
  const int npairs = N * (N - 1) / 2;
  const int64_t ntiles = (p.D + kObsThreads - 1) / kObsThreads;
  const int64_t G = gridDim.x, g = blockIdx.x;
  const int64_t t0 = g * ntiles / G, t1 = (g + 1) * ntiles / G;
  for (int64_t tile = t0; tile < t1; ++tile) {
    const int64_t row = tile * kObsThreads + threadIdx.x;
    const bool valid = row < p.D;
    S s = 0;
    double ps = 0.0;
    int kc[NC];
    R tc[NC];
    if (valid) {
      s = unrank_row<S, R>((R)row, N, p.nup, s_binom);
      ps = psi[row];
      chunk_terms<S, R, W, NC>(s, kc, tc, s_tab, p);
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) { kc[c] = 0; tc[c] = 0; }
    }
    const double p2 = ps * ps;
    int q = 0;
    for (int i = 0; i < N; ++i) {
      for (int j = i + 1; j < N; ++j, ++q) {
        const bool anti = valid && (((s >> i) ^ (s >> j)) & 1);
        double par = anti ? 0.0 : p2;
        double flip = 0.0;
        if (anti) {
          const R rn = flip_rank<S, R, W, NC>((R)row, s, i, j, kc, tc, s_tab, p);
          flip = ps * psi[rn];
        }
        par = obs_wsum(par);
        flip = obs_wsum(flip);
        if (lane == 0) {
          my[q] += par;
          my[npairs + q] += flip;
        }
      }
    }
    for (int i = 0; i < N; ++i) {
      double mi = ((s >> i) & 1) ? p2 : 0.0;
      mi = obs_wsum(mi);
      if (lane == 0) my[2 * npairs + i] += mi;
    }
    const double tot = obs_wsum(p2);
    if (lane == 0) my[2 * npairs + N] += tot;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < Q; i += blockDim.x)
    part[(int64_t)i * gridDim.x + blockIdx.x] = acc[i] + acc[Q + i] + acc[2 * Q + i] + acc[3 * Q + i];
}

// DSF with S^z_i (Sqz, same sector): a[n][i] = sum_s e_n(s) * 0.5 * sigma_i(s) * psi0(s).
// Quantities per block: nev * N, index n*N + i.
template <typename S, typename R, int W, int NC>
__global__ __launch_bounds__(kObsThreads) void k_dsf_z(RowParams p, const R* __restrict__ gbinom,
                                                       const double* __restrict__ psi0,
                                                       const double* __restrict__ ev, int nev,
                                                       double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = p.N;
  const int Q = nev * N;
  double* acc = reinterpret_cast<double*>(smem);
  R* s_binom = reinterpret_cast<R*>(acc + 4 * Q);
  for (int i = threadIdx.x; i < p.binom_entries; i += blockDim.x) s_binom[i] = gbinom[i];
  for (int i = threadIdx.x; i < 4 * Q; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double* my = acc + wid * Q;
  const int64_t ntiles = (p.D + kObsThreads - 1) / kObsThreads;
  const int64_t G = gridDim.x, g = blockIdx.x;
  const int64_t t0 = g * ntiles / G, t1 = (g + 1) * ntiles / G;
  for (int64_t tile = t0; tile < t1; ++tile) {
    const int64_t row = tile * kObsThreads + threadIdx.x;
    const bool valid = row < p.D;
    S s = valid ? unrank_row<S, R>((R)row, N, p.nup, s_binom) : (S)0;
    const double a0 = valid ? psi0[row] : 0.0;
    for (int n = 0; n < nev; ++n) {
      const double qn = valid ? ev[(int64_t)n * p.D + row] * a0 : 0.0;
      for (int i = 0; i < N; ++i) {
        // Sqz::OpOnStateVector: rhs = psi * 0.5 * (+1 up / -1 down)
        double c = qn * (((s >> i) & 1) ? 0.5 : -0.5);
        c = obs_wsum(c);
        if (lane == 0) my[n * N + i] += c;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < Q; i += blockDim.x)
    part[(int64_t)i * gridDim.x + blockIdx.x] = acc[i] + acc[Q + i] + acc[2 * Q + i] + acc[3 * Q + i];
}

// DSF with S^+_i (Sqp, sector nup -> nup+1): a[n][i] = sum_{s: b_i = 0} e_n(rank1(s | 1<<i)) psi0(s).
template <typename S, typename R, int W, int NC>
__global__ __launch_bounds__(kObsThreads) void k_dsf_p(RowParams p0, RowParams p1, const R* __restrict__ gbinom0,
                                                       const R* __restrict__ gtab1,
                                                       const double* __restrict__ psi0,
                                                       const double* __restrict__ ev, int nev,
                                                       double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = p0.N;
  const int Q = nev * N;
  double* acc = reinterpret_cast<double*>(smem);
  R* s_binom = reinterpret_cast<R*>(acc + 4 * Q);
  R* s_tab1 = s_binom + p0.binom_entries;
  for (int i = threadIdx.x; i < p0.binom_entries; i += blockDim.x) s_binom[i] = gbinom0[i];
  for (int i = threadIdx.x; i < p1.table_entries; i += blockDim.x) s_tab1[i] = gtab1[i];
  for (int i = threadIdx.x; i < 4 * Q; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double* my = acc + wid * Q;
  const int64_t ntiles = (p0.D + kObsThreads - 1) / kObsThreads;
  const int64_t G = gridDim.x, g = blockIdx.x;
  const int64_t t0 = g * ntiles / G, t1 = (g + 1) * ntiles / G;
  for (int64_t tile = t0; tile < t1; ++tile) {
    const int64_t row = tile * kObsThreads + threadIdx.x;
    const bool valid = row < p0.D;
    S s = valid ? unrank_row<S, R>((R)row, N, p0.nup, s_binom) : (S)0;
    const double a0 = valid ? psi0[row] : 0.0;
    for (int i = 0; i < N; ++i) {
      const bool on = valid && !((s >> i) & 1);
      R r1 = 0;
      if (on) r1 = rank_state<S, R, W, NC>(s | ((S)1 << i), s_tab1, p1);
      for (int n = 0; n < nev; ++n) {
        double c = on ? ev[(int64_t)n * p1.D + r1] * a0 : 0.0;
        c = obs_wsum(c);
        if (lane == 0) my[n * N + i] += c;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < Q; i += blockDim.x)
    part[(int64_t)i * gridDim.x + blockIdx.x] = acc[i] + acc[Q + i] + acc[2 * Q + i] + acc[3 * Q + i];
}

__global__ void k_obs_reduce(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double red[8];
  const double* p = part + (int64_t)blockIdx.x * n;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += p[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    out[blockIdx.x] = t;
  }
}

static int obs_blocks(int64_t D) { return (int)std::max<int64_t>(1, std::min<int64_t>(2048, (D + 255) / 256)); }

// Runs a kernel producing Q columns of G partials into h's scratch, reduces
// them to `out` (host, Q values).
template <typename Launch>
static int run_reduce(lsqed_ham* h, int Q, int G, hipStream_t s, double* out, Launch launch) {
  int rc = ensure_scratch(h, (size_t)Q * G + Q);
  if (rc) return rc;
  double* part = h->d_scratch;
  double* red = h->d_scratch + (size_t)Q * G;
  launch(part);
  LSQED_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_obs_reduce, dim3(Q), dim3(256), 0, s, part, G, red);
  LSQED_HIP(hipGetLastError());
  LSQED_HIP(hipMemcpyAsync(out, red, Q * sizeof(double), hipMemcpyDeviceToHost, s));
  LSQED_HIP(hipStreamSynchronize(s));
  return LSQED_OK;
}

static int pair_sums(lsqed_ham* h, const double* psi, hipStream_t s, std::vector<double>& q) {
  const int N = h->N;
  const int npairs = N * (N - 1) / 2;
  const int Q = 2 * npairs + N + 1;
  const int G = obs_blocks(h->D);
  q.assign(Q, 0.0);
  RowParams p = make_row_params(h);
  const size_t esz = h->tab.wide ? 8 : 4;
  const size_t sm = (size_t)4 * Q * sizeof(double) + (h->tab.entries.size() + h->tab.binom.size()) * esz;
  return run_reduce(h, Q, G, s, q.data(), [&](double* part) {
    if (!h->tab.wide) {
      auto k = k_obs_pairs<uint32_t, uint32_t, 8, 4>;
      if (sm > 64 * 1024) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
      hipLaunchKernelGGL(k, dim3(G), dim3(kObsThreads), sm, s, p, (const uint32_t*)h->d_table,
                         (const uint32_t*)h->d_binom, psi, Q, part);
    } else {
      auto k = k_obs_pairs<uint64_t, uint64_t, 6, kMaxChunks>;
      if (sm > 64 * 1024) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
      hipLaunchKernelGGL(k, dim3(G), dim3(kObsThreads), sm, s, p, (const uint64_t*)h->d_table,
                         (const uint64_t*)h->d_binom, psi, Q, part);
    }
  });
}

}  // namespace lsqed

using namespace lsqed;

extern "C" {

int lsqed_obs_szsz(lsqed_ham* h, const double* psi, double* out, lsqed_stream_t stream) {
  if (!h || !psi || !out) return fail(LSQED_ERR_ARG, "null argument");
  std::vector<double> q;
  int rc = pair_sums(h, psi, (hipStream_t)stream, q);
  if (rc) return rc;
  const int N = h->N, np = N * (N - 1) / 2;
  const double P = q[2 * np + N];
  int k = 0;
  for (int i = 0; i < N; ++i) {
    out[i * N + i] = 0.25 * P;
    for (int j = i + 1; j < N; ++j, ++k) out[i * N + j] = out[j * N + i] = 0.25 * (2.0 * q[k] - P);
  }
  return LSQED_OK;
}

int lsqed_obs_spsm(lsqed_ham* h, const double* psi, double* out, lsqed_stream_t stream) {
  if (!h || !psi || !out) return fail(LSQED_ERR_ARG, "null argument");
  std::vector<double> q;
  int rc = pair_sums(h, psi, (hipStream_t)stream, q);
  if (rc) return rc;
  const int N = h->N, np = N * (N - 1) / 2;
  int k = 0;
  for (int i = 0; i < N; ++i) {
    out[i * N + i] = q[2 * np + i];
    for (int j = i + 1; j < N; ++j, ++k) out[i * N + j] = out[j * N + i] = 0.5 * q[np + k];
  }
  return LSQED_OK;
}

int lsqed_obs_sz(lsqed_ham* h, const double* psi, double* out, lsqed_stream_t stream) {
  if (!h || !psi || !out) return fail(LSQED_ERR_ARG, "null argument");
  std::vector<double> q;
  int rc = pair_sums(h, psi, (hipStream_t)stream, q);
  if (rc) return rc;
  const int N = h->N, np = N * (N - 1) / 2;
  const double P = q[2 * np + N];
  for (int i = 0; i < N; ++i) out[i] = q[2 * np + i] - 0.5 * P;
  return LSQED_OK;
}

int lsqed_obs_dsf(lsqed_ham* h0, lsqed_ham* h1, const double* psi0, const double* evecs, int nev, double* out,
                  lsqed_stream_t stream) {
  if (!h0 || !h1 || !psi0 || !evecs || !out) return fail(LSQED_ERR_ARG, "null argument");
  if (nev <= 0 || nev > 64) return fail(LSQED_ERR_ARG, "nev must be in 1..64");
  if (h1->N != h0->N) return fail(LSQED_ERR_ARG, "sectors of different lattices");
  const bool same = (h1 == h0) || (h1->nup == h0->nup);
  if (!same && h1->nup != h0->nup + 1) return fail(LSQED_ERR_ARG, "target sector must be nup or nup+1");
  if (h0->tab.wide != h1->tab.wide) return fail(LSQED_ERR_ARG, "sector width mismatch");
  hipStream_t s = (hipStream_t)stream;
  const int N = h0->N, Q = nev * N, G = obs_blocks(h0->D);
  RowParams p0 = make_row_params(h0), p1 = make_row_params(h1);
  const size_t esz = h0->tab.wide ? 8 : 4;
  if (same) {
    const size_t sm = (size_t)4 * Q * sizeof(double) + h0->tab.binom.size() * esz;
    return run_reduce(h0, Q, G, s, out, [&](double* part) {
      if (!h0->tab.wide) {
        auto k = k_dsf_z<uint32_t, uint32_t, 8, 4>;
        if (sm > 64 * 1024) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        hipLaunchKernelGGL(k, dim3(G), dim3(kObsThreads), sm, s, p0, (const uint32_t*)h0->d_binom, psi0, evecs,
                           nev, part);
      } else {
        auto k = k_dsf_z<uint64_t, uint64_t, 6, kMaxChunks>;
        if (sm > 64 * 1024) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        hipLaunchKernelGGL(k, dim3(G), dim3(kObsThreads), sm, s, p0, (const uint64_t*)h0->d_binom, psi0, evecs,
                           nev, part);
      }
    });
  }
  const size_t sm = (size_t)4 * Q * sizeof(double) + (h0->tab.binom.size() + h1->tab.entries.size()) * esz;
  return run_reduce(h0, Q, G, s, out, [&](double* part) {
    if (!h0->tab.wide) {
      auto k = k_dsf_p<uint32_t, uint32_t, 8, 4>;
      if (sm > 64 * 1024) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
      hipLaunchKernelGGL(k, dim3(G), dim3(kObsThreads), sm, s, p0, p1, (const uint32_t*)h0->d_binom,
                         (const uint32_t*)h1->d_table, psi0, evecs, nev, part);
    } else {
      auto k = k_dsf_p<uint64_t, uint64_t, 6, kMaxChunks>;
      if (sm > 64 * 1024) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
      hipLaunchKernelGGL(k, dim3(G), dim3(kObsThreads), sm, s, p0, p1, (const uint64_t*)h0->d_binom,
                         (const uint64_t*)h1->d_table, psi0, evecs, nev, part);
    }
  });
}

}  // extern "C"
