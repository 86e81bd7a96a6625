// sector_device.hpp -- device-side rank/unrank and reduction helpers shared
// by the row kernels (H.v, diagonal, export, observables).
#pragma once

#include <hip/hip_runtime.h>

#include "lsqed_internal.hpp"

namespace lsqed {

template <typename S, typename R>
__host__ __device__ __forceinline__ S unrank_row(R row, int N, int nup, const R* __restrict__ binom) {
  S s = 0;
  R rest = row;
  int k = nup;
  for (int pos = N - 1; pos >= 0 && k > 0; --pos) {
    R c = binom[pos * (nup + 1) + k];  // C(pos, k), zero for pos < k
    if (c <= rest) {
      rest -= c;
      s |= (S)1 << pos;
      --k;
    }
  }
  return s;
}

template <typename S, typename R, int W>
__host__ __device__ __forceinline__ R tab_at(const R* __restrict__ tab, const RowParams& p, int c, int k, uint32_t v) {
  return tab[((p.rowbase[c] + (k - p.kmin[c])) << W) + v];
}

template <typename S, typename R, int W, int NC>
__host__ __device__ __forceinline__ R rank_state(S s, const R* __restrict__ tab, const RowParams& p) {
  R r = 0;
  int k = 0;
  const int nch = (p.N + W - 1) / W;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c < nch) {
      uint32_t v = (uint32_t)(s >> (c * W)) & ((1u << W) - 1u);
      r += tab_at<S, R, W>(tab, p, c, k, v);
      k += __builtin_popcount(v);
    }
  }
  return r;
}

// Diagonal element with the reference's operation order, no contraction.
__device__ __forceinline__ double diag_value(const RowParams& p, int c0, int c1, uint64_t s,
                                             const double* __restrict__ f) {
#pragma clang fp contract(off)
  double d = 0.0;
  if (p.diag_mode == 0) {
    // build_diag: ADD D1/4*c_nn then ADD D2/4*c_nnn, each skipped for
    // |D| < 1e-14 or c == 0 (tmpl:144-213).
    if (p.dz_on[0] && c0 != 0) d = p.dq[0] * (double)c0;
    if (p.dz_on[1] && c1 != 0) d = d + p.dq[1] * (double)c1;
  } else if (p.diag_mode == 1) {
    // build_disorder: D1*0.25*c_nn + D2*0.25*c_nnn + 0.5*sum_i(+-h_i)
    // (tmpl:108-111, src/hamiltonian.cpp:31-49).
    double dis = 0.0;
    for (int i = 0; i < p.N; ++i) dis = ((s >> i) & 1) ? dis + f[i] : dis - f[i];
    double a = p.dq[0] * (double)c0;
    double b = p.dq[1] * (double)c1;
    double e = 0.5 * dis;
    d = (a + b) + e;
  }
  return d;
}

// Vector entries: FP64 (real scalars) or interleaved complex FP64 (re, im),
// as a complex-scalar PETSc Vec stores them.  H is real, so both parts see
// the same coefficients; each part is updated with the same expression as
// the real product, so it rounds the same way.
template <bool CPLX>
struct Entry {
  typedef double T;
  static __device__ __forceinline__ T zero() { return 0.0; }
  static __device__ __forceinline__ T load(const double* __restrict__ x, int64_t i) { return x[i]; }
  static __device__ __forceinline__ void store(double* __restrict__ y, int64_t i, T v) { y[i] = v; }
};
template <>
struct Entry<true> {
  typedef double2 T;
  static __device__ __forceinline__ T zero() { return make_double2(0.0, 0.0); }
  static __device__ __forceinline__ T load(const double* __restrict__ x, int64_t i) {
    return reinterpret_cast<const double2*>(x)[i];
  }
  static __device__ __forceinline__ void store(double* __restrict__ y, int64_t i, T v) {
    reinterpret_cast<double2*>(y)[i] = v;
  }
};
__device__ __forceinline__ void madd(double& acc, double c, double v) { acc += c * v; }
__device__ __forceinline__ void madd(double2& acc, double c, double2 v) {
  acc.x += c * v.x;
  acc.y += c * v.y;
}
__device__ __forceinline__ double diag_apply(double d, double x, double acc) { return d * x + acc; }
__device__ __forceinline__ double2 diag_apply(double d, double2 x, double2 acc) {
  return make_double2(d * x.x + acc.x, d * x.y + acc.y);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-wide sum over a 256-thread block; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
  return t;
}

// Logical block index: consecutive logical blocks sit on the same XCD
// (hardware dispatches block b to XCD b % 8), so each XCD walks one
// contiguous range of rows and its L2 sees the nearby-rank gathers.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t G = gridDim.x, b = blockIdx.x;
  return (b & 7) * (G >> 3) + (b >> 3);
}


// Rank of s ^ (bit lo | bit hi) for an antiparallel pair lo < hi, from the
// row's rank r and its per-chunk (k_c, T_c) kept in registers: only chunks
// clo..chi change; k_c moves by -1 (particle lo -> hi) or +1 (hi -> lo) for
// clo < c <= chi.
template <typename S, typename R, int W, int NC>
__host__ __device__ __forceinline__ R flip_rank(R r, S s, int lo, int hi, const int* kc, const R* tc,
                                       const R* __restrict__ tab, const RowParams& p) {
  const S sp = s ^ (((S)1 << lo) | ((S)1 << hi));
  const int dk = ((s >> lo) & 1) ? -1 : 1;
  const int clo = lo / W, chi = hi / W;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c >= clo && c <= chi) {
      const int k = kc[c] + (c > clo ? dk : 0);
      const uint32_t v = (uint32_t)(sp >> (c * W)) & ((1u << W) - 1u);
      r += tab_at<S, R, W>(tab, p, c, k, v) - tc[c];
    }
  }
  return r;
}

// Per-chunk (k_c, T_c) of a state.
template <typename S, typename R, int W, int NC>
__host__ __device__ __forceinline__ void chunk_terms(S s, int* kc, R* tc, const R* __restrict__ tab, const RowParams& p) {
  const int nch = (p.N + W - 1) / W;
  int k = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const uint32_t v = (uint32_t)(s >> (c * W)) & ((1u << W) - 1u);
    kc[c] = k;
    tc[c] = (c < nch) ? tab_at<S, R, W>(tab, p, c, k, v) : (R)0;
    k += __builtin_popcount(v);
  }
}

}  // namespace lsqed
