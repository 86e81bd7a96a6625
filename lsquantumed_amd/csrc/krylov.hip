// krylov.hip -- device-resident Krylov-Schur (thick-restart Lanczos) and the
// plain three-term Lanczos on MI355X.
//
// Replaces SLEPc's EPSKRYLOVSCHUR for EPS_HEP / EPS_SMALLEST_REAL as
// configured by Solver::SolverInit (src/solver.cpp:7-24).  All D-sized data
// (the ncv+1 basis vectors) stays in HBM; only the projected (ncv x ncv)
// problem and O(ncv) scalars cross to the host per step.
//
// Vector kernels are HBM-streaming: coalesced FP64 loads, wavefront shuffle
// reductions, per-block partial sums reduced in a fixed order (results are
// deterministic run to run).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "lsqed_internal.hpp"

namespace lsqed {


constexpr int kThreads = 256;
constexpr int kRedBlocks = 1024;  // partial sums per reduction
constexpr int kMaxNcv = 64;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum over a block of at most 512 threads, valid in thread 0.
__device__ __forceinline__ double bsum(double v) {
  __shared__ double red[8];
  v = wsum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
  return t;
}

__device__ __forceinline__ void block_range(int64_t n, int64_t& a, int64_t& b) {
  const int64_t G = gridDim.x, g = blockIdx.x;
  a = g * n / G;
  b = (g + 1) * n / G;
}

// 16-byte loads and stores of streams that a pass touches once: with NT the
// nontemporal forms, which do not keep the lines in L2 (chosen for vectors
// larger than the 256 MB memory-side cache, where nothing would be reused).
template <bool NT>
__device__ __forceinline__ double2 ld2(const double2* p) {
  if constexpr (NT) {
    typedef double v2 __attribute__((ext_vector_type(2)));
    const v2 t = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return make_double2(t.x, t.y);
  } else {
    return *p;
  }
}
template <bool NT>
__device__ __forceinline__ void st2(double2* p, double2 v) {
  if constexpr (NT) {
    typedef double v2 __attribute__((ext_vector_type(2)));
    v2 t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<v2*>(p));
  } else {
    *p = v;
  }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// v[r] = uniform [-1, 1) from splitmix64(seed, r).
__global__ void k_fill_random(double* __restrict__ v, int64_t n, uint64_t seed) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = splitmix64(seed * 0xD1B54A32D192ED03ull + (uint64_t)r);
    v[r] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

// part[blockIdx] = sum over the block's range of x*y
__global__ void k_dot(const double* __restrict__ x, const double* __restrict__ y, int64_t n,
                      double* __restrict__ part) {
  int64_t a, b;
  block_range(n, a, b);
  double acc = 0.0;
  for (int64_t r = a + threadIdx.x; r < b; r += blockDim.x) acc += x[r] * y[r];
  double t = bsum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out[c] = sum_i part[c*n + i]  (one block per column, fixed order)
// Column sums of partials (n per column), one 512-thread block per column.
// Eight independent accumulators per lane keep eight loads in flight: the
// x.y partials of a large H.v launch (one per block) number 2^17 on 4x8 and
// 2^21 on 6x6, which one load at a time per lane took 3.4 ms to sum on 6x6.
// The order of the sums is fixed, so results are deterministic.
constexpr int kRedColThreads = 512;
__global__ __launch_bounds__(kRedColThreads) void k_reduce_cols(const double* __restrict__ part, int n,
                                                                double* __restrict__ out) {
  const double* p = part + (int64_t)blockIdx.x * n;
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int i = threadIdx.x;
  for (; i + 7 * kRedColThreads < n; i += 8 * kRedColThreads) {
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += p[i + u * kRedColThreads];
  }
  for (; i < n; i += kRedColThreads) acc[0] += p[i];
  const double t = bsum(((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7])));
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// h_j partials: part[j*G + block] = sum_r V[j*ld + r] * w[r] for j in the
// block's column group (blockIdx.y), 16 columns per group.
__global__ __launch_bounds__(kThreads) void k_multidot(const double* __restrict__ V, int64_t ld, int m,
                                                       const double* __restrict__ w, int64_t n,
                                                       double* __restrict__ part) {
  int64_t a, b;
  block_range(n, a, b);
  const int j0 = blockIdx.y * 16;
  double acc[16];
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) acc[jj] = 0.0;
  for (int64_t r = a + threadIdx.x; r < b; r += blockDim.x) {
    const double wr = w[r];
#pragma unroll
    for (int jj = 0; jj < 16; ++jj)
      if (j0 + jj < m) acc[jj] += V[(int64_t)(j0 + jj) * ld + r] * wr;
  }
  __shared__ double red[4][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) {
    double t = wsum(acc[jj]);
    if (lane == 0) red[wid][jj] = t;
  }
  __syncthreads();
  if (threadIdx.x < 16 && j0 + (int)threadIdx.x < m) {
    const int jj = threadIdx.x;
    part[(int64_t)(j0 + jj) * gridDim.x + blockIdx.x] = red[0][jj] + red[1][jj] + red[2][jj] + red[3][jj];
  }
}

// w -= V[:, 0:m] h ; part[block] = partial ||w||^2
__global__ __launch_bounds__(kThreads) void k_update(const double* __restrict__ V, int64_t ld, int m,
                                                     const double* __restrict__ hcoef, double* __restrict__ w,
                                                     int64_t n, double* __restrict__ part) {
  __shared__ double hs[kMaxNcv + 1];
  for (int j = threadIdx.x; j < m; j += blockDim.x) hs[j] = hcoef[j];
  __syncthreads();
  int64_t a, b;
  block_range(n, a, b);
  double acc = 0.0;
  for (int64_t r = a + threadIdx.x; r < b; r += blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < m; ++j) s += V[(int64_t)j * ld + r] * hs[j];
    const double v = w[r] - s;
    w[r] = v;
    acc += v * v;
  }
  double t = bsum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// One pass of classical Gram-Schmidt over rows, m <= MAXM basis columns
// held in registers (all m loads of a row are issued before any is used):
//   MODE 0: part[j] += V_j . w                          (projection)
//   MODE 1: w <- w - V h;  part[j] += V_j . w; part[m] += ||w||^2
//           (the first update fused with the second projection of CGS2)
//   MODE 2: w <- (w - V h) * sc;  part[0] += ||w||^2     (final update; sc
//           normalises the new basis vector in the same pass, 1 keeps it)
// part is column-major over blocks: part[j * gridDim.x + block].
template <int MAXM, int MODE>
__global__ __launch_bounds__(kThreads) void k_cgs_pass(const double* __restrict__ V, int64_t ld, int m,
                                                       const double* __restrict__ hcoef, double* __restrict__ w,
                                                       int64_t n, double* __restrict__ part, double sc) {
  __shared__ double hs[MAXM];
  __shared__ double red[kThreads / 64][MAXM + 1];
  if (MODE != 0)
    for (int j = threadIdx.x; j < MAXM; j += blockDim.x) hs[j] = j < m ? hcoef[j] : 0.0;
  __syncthreads();
  int64_t a, b;
  block_range(n, a, b);
  double acc[MAXM + 1];
#pragma unroll
  for (int j = 0; j <= MAXM; ++j) acc[j] = 0.0;
  for (int64_t r = a + threadIdx.x; r < b; r += blockDim.x) {
    double vin[MAXM];
#pragma unroll
    for (int j = 0; j < MAXM; ++j) vin[j] = j < m ? V[(int64_t)j * ld + r] : 0.0;
    double wr = w[r];
    if (MODE != 0) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < MAXM; ++j) s += vin[j] * hs[j];
      wr -= s;
      if (MODE == 2) wr *= sc;
      w[r] = wr;
      acc[MAXM] += wr * wr;
    }
    if (MODE != 2) {
#pragma unroll
      for (int j = 0; j < MAXM; ++j) acc[j] += vin[j] * wr;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j <= MAXM; ++j) {
    if ((MODE == 2 && j < MAXM) || (MODE == 0 && j == MAXM)) continue;
    const double t = wsum(acc[j]);
    if (lane == 0) red[wid][j] = t;
  }
  __syncthreads();
  const int G = gridDim.x;
  for (int j = threadIdx.x; j <= MAXM; j += blockDim.x) {
    double t = 0.0;
    for (int q = 0; q < kThreads / 64; ++q) t += red[q][j];
    if (MODE == 0 && j < m) part[(int64_t)j * G + blockIdx.x] = t;
    if (MODE == 1 && j < m) part[(int64_t)j * G + blockIdx.x] = t;
    if (MODE == 1 && j == MAXM) part[(int64_t)m * G + blockIdx.x] = t;
    if (MODE == 2 && j == MAXM) part[blockIdx.x] = t;
  }
}

__global__ void k_scale(double* __restrict__ v, int64_t n, double a) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    v[r] *= a;
}

// Delayed CGS2 (one projection pass and one update pass per Krylov step).
// The new column j holds u, the residual of the previous step orthogonalised
// once; y = A u.  Projection: part[i] = V_i . u and part[j + i] = V_i . y
// (i < j), part[2j] = u.u, part[2j+1] = u.y, part[2j+2] = y.y; without y
// (HAS_Y false) only V_i . u and u.u (part[j]).  One row per thread per
// iteration, all column loads of a row in flight; part is column-major over
// blocks: part[q * gridDim.x + block].
template <int MAXM, bool HAS_Y>
__global__ __launch_bounds__(kThreads) void k_dcgs_proj(const double* __restrict__ V, int64_t ld, int j,
                                                        const double* __restrict__ u, const double* __restrict__ y,
                                                        int64_t n, double* __restrict__ part) {
  constexpr int NQ = HAS_Y ? 2 * MAXM + 3 : MAXM + 1;
  __shared__ double red[kThreads / 64][NQ];
  int64_t a, b;
  block_range(n, a, b);
  double cu[MAXM], cy[MAXM];
  double uu = 0.0, uy = 0.0, yy = 0.0;
#pragma unroll
  for (int i = 0; i < MAXM; ++i) cu[i] = cy[i] = 0.0;
  for (int64_t r = a + threadIdx.x; r < b; r += blockDim.x) {
    double vin[MAXM];
#pragma unroll
    for (int i = 0; i < MAXM; ++i) vin[i] = i < j ? V[(int64_t)i * ld + r] : 0.0;
    const double ur = u[r];
    const double yr = HAS_Y ? y[r] : 0.0;
#pragma unroll
    for (int i = 0; i < MAXM; ++i) {
      cu[i] += vin[i] * ur;
      if (HAS_Y) cy[i] += vin[i] * yr;
    }
    uu += ur * ur;
    if (HAS_Y) {
      uy += ur * yr;
      yy += yr * yr;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < MAXM; ++i) {
    const double t = wsum(cu[i]);
    if (lane == 0) red[wid][i] = t;
    if (HAS_Y) {
      const double t2 = wsum(cy[i]);
      if (lane == 0) red[wid][MAXM + i] = t2;
    }
  }
  {
    const double t = wsum(uu);
    if (lane == 0) red[wid][HAS_Y ? 2 * MAXM : MAXM] = t;
  }
  if (HAS_Y) {
    const double t1 = wsum(uy), t2 = wsum(yy);
    if (lane == 0) {
      red[wid][2 * MAXM + 1] = t1;
      red[wid][2 * MAXM + 2] = t2;
    }
  }
  __syncthreads();
  const int G = gridDim.x;
  for (int q = threadIdx.x; q < NQ; q += blockDim.x) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w][q];
    // compact layout: [V.u (j)] [V.y (j)] [u.u] [u.y] [y.y]
    int out = -1;
    if (q < MAXM) out = q < j ? q : -1;
    else if (HAS_Y && q < 2 * MAXM) out = q - MAXM < j ? j + (q - MAXM) : -1;
    else out = (HAS_Y ? 2 * j : j) + (q - (HAS_Y ? 2 * MAXM : MAXM));
    if (out >= 0) part[(int64_t)out * G + blockIdx.x] = t;
  }
}

// Delayed CGS2 update: u <- (u - V c) * inv_nu (the new basis vector) and,
// with y, y <- ya * y + ub * u_old - V z (the next residual, orthogonalised
// once) with part[block] = ||y||^2.  coef = c[0..j) then z[0..j).  With OOP
// the new residual goes to yo instead of over y (a buffer the pass does not
// read: its stores then do not land on lines the same pass has just read,
// which measured 8% faster for the whole pass on 4.8 GB columns).
template <int MAXM, bool HAS_Y, bool OOP>
__global__ __launch_bounds__(kThreads) void k_dcgs_update(const double* __restrict__ V, int64_t ld, int j,
                                                          const double* __restrict__ coef, double inv_nu, double ya,
                                                          double ub, double* __restrict__ u, double* __restrict__ y,
                                                          double* __restrict__ yo, int64_t n,
                                                          double* __restrict__ part) {
  __shared__ double cs[MAXM], zs[MAXM];
  for (int i = threadIdx.x; i < MAXM; i += blockDim.x) {
    cs[i] = i < j ? coef[i] : 0.0;
    zs[i] = (HAS_Y && i < j) ? coef[j + i] : 0.0;
  }
  __syncthreads();
  int64_t a, b;
  block_range(n, a, b);
  double acc = 0.0;
  for (int64_t r = a + threadIdx.x; r < b; r += blockDim.x) {
    double vin[MAXM];
#pragma unroll
    for (int i = 0; i < MAXM; ++i) vin[i] = i < j ? V[(int64_t)i * ld + r] : 0.0;
    const double ur = u[r];
    const double yr = HAS_Y ? y[r] : 0.0;
    double sc = 0.0, sz = 0.0;
#pragma unroll
    for (int i = 0; i < MAXM; ++i) {
      sc += vin[i] * cs[i];
      if (HAS_Y) sz += vin[i] * zs[i];
    }
    u[r] = (ur - sc) * inv_nu;
    if (HAS_Y) {
      const double w = ya * yr + ub * ur - sz;
      if (OOP) yo[r] = w;
      else y[r] = w;
      acc += w * w;
    }
  }
  if (HAS_Y) {
    const double t = bsum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// Two-row forms of the delayed CGS2 passes and the restart: each thread takes
// rows 2t and 2t+1 with 16-byte loads (columns 16-byte aligned: ld even),
// halving the load instructions per byte; the arithmetic of each row is
// that of the one-row kernels.
template <int MAXM, bool HAS_Y, bool NT>
__global__ __launch_bounds__(kThreads) void k_dcgs_proj2(const double* __restrict__ V, int64_t ld, int j,
                                                         const double* __restrict__ u, const double* __restrict__ y,
                                                         int64_t n, double* __restrict__ part) {
  constexpr int NQ = HAS_Y ? 2 * MAXM + 3 : MAXM + 1;
  __shared__ double red[kThreads / 64][NQ];
  int64_t a, b;
  block_range(n / 2, a, b);
  const double2* __restrict__ U2 = reinterpret_cast<const double2*>(u);
  const double2* __restrict__ Y2 = reinterpret_cast<const double2*>(y);
  double cu[MAXM], cy[MAXM];
  double uu = 0.0, uy = 0.0, yy = 0.0;
#pragma unroll
  for (int i = 0; i < MAXM; ++i) cu[i] = cy[i] = 0.0;
  for (int64_t r = a + threadIdx.x; r < b; r += blockDim.x) {
    double2 vin[MAXM];
#pragma unroll
    for (int i = 0; i < MAXM; ++i)
      vin[i] = i < j ? ld2<NT>(reinterpret_cast<const double2*>(V + (int64_t)i * ld) + r) : make_double2(0.0, 0.0);
    const double2 ur = ld2<NT>(U2 + r);
    const double2 yr = HAS_Y ? ld2<NT>(Y2 + r) : make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < MAXM; ++i) {
      cu[i] += vin[i].x * ur.x;
      cu[i] += vin[i].y * ur.y;
      if (HAS_Y) {
        cy[i] += vin[i].x * yr.x;
        cy[i] += vin[i].y * yr.y;
      }
    }
    uu += ur.x * ur.x;
    uu += ur.y * ur.y;
    if (HAS_Y) {
      uy += ur.x * yr.x;
      uy += ur.y * yr.y;
      yy += yr.x * yr.x;
      yy += yr.y * yr.y;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < MAXM; ++i) {
    const double t = wsum(cu[i]);
    if (lane == 0) red[wid][i] = t;
    if (HAS_Y) {
      const double t2 = wsum(cy[i]);
      if (lane == 0) red[wid][MAXM + i] = t2;
    }
  }
  {
    const double t = wsum(uu);
    if (lane == 0) red[wid][HAS_Y ? 2 * MAXM : MAXM] = t;
  }
  if (HAS_Y) {
    const double t1 = wsum(uy), t2 = wsum(yy);
    if (lane == 0) {
      red[wid][2 * MAXM + 1] = t1;
      red[wid][2 * MAXM + 2] = t2;
    }
  }
  __syncthreads();
  const int G = gridDim.x;
  for (int q = threadIdx.x; q < NQ; q += blockDim.x) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w][q];
    int out = -1;
    if (q < MAXM) out = q < j ? q : -1;
    else if (HAS_Y && q < 2 * MAXM) out = q - MAXM < j ? j + (q - MAXM) : -1;
    else out = (HAS_Y ? 2 * j : j) + (q - (HAS_Y ? 2 * MAXM : MAXM));
    if (out >= 0) part[(int64_t)out * G + blockIdx.x] = t;
  }
}

template <int MAXM, bool HAS_Y, bool OOP>
__global__ __launch_bounds__(kThreads) void k_dcgs_update2(const double* __restrict__ V, int64_t ld, int j,
                                                           const double* __restrict__ coef, double inv_nu, double ya,
                                                           double ub, double* __restrict__ u, double* __restrict__ y,
                                                           double* __restrict__ yo, int64_t n,
                                                           double* __restrict__ part) {
  __shared__ double cs[MAXM], zs[MAXM];
  for (int i = threadIdx.x; i < MAXM; i += blockDim.x) {
    cs[i] = i < j ? coef[i] : 0.0;
    zs[i] = (HAS_Y && i < j) ? coef[j + i] : 0.0;
  }
  __syncthreads();
  int64_t a, b;
  block_range(n / 2, a, b);
  double2* __restrict__ U2 = reinterpret_cast<double2*>(u);
  double2* __restrict__ Y2 = reinterpret_cast<double2*>(y);
  double2* __restrict__ YO2 = reinterpret_cast<double2*>(yo);
  double acc = 0.0;
  for (int64_t r = a + threadIdx.x; r < b; r += blockDim.x) {
    double2 vin[MAXM];
#pragma unroll
    for (int i = 0; i < MAXM; ++i)
      vin[i] = i < j ? reinterpret_cast<const double2*>(V + (int64_t)i * ld)[r] : make_double2(0.0, 0.0);
    const double2 ur = U2[r];
    const double2 yr = HAS_Y ? Y2[r] : make_double2(0.0, 0.0);
    double2 sc = make_double2(0.0, 0.0), sz = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < MAXM; ++i) {
      sc.x += vin[i].x * cs[i];
      sc.y += vin[i].y * cs[i];
      if (HAS_Y) {
        sz.x += vin[i].x * zs[i];
        sz.y += vin[i].y * zs[i];
      }
    }
    U2[r] = make_double2((ur.x - sc.x) * inv_nu, (ur.y - sc.y) * inv_nu);
    if (HAS_Y) {
      const double2 w = make_double2(ya * yr.x + ub * ur.x - sz.x, ya * yr.y + ub * ur.y - sz.y);
      if (OOP) YO2[r] = w;
      else Y2[r] = w;
      acc += w.x * w.x;
      acc += w.y * w.y;
    }
  }
  if (HAS_Y) {
    const double t = bsum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

template <int MAXM, bool NT>
__global__ __launch_bounds__(kThreads) void k_restart2(double* __restrict__ V, int64_t ld, int m, int k,
                                                       const double* __restrict__ S, int64_t n) {
  __shared__ double Ss[MAXM * MAXM];
  for (int i = threadIdx.x; i < m * k; i += blockDim.x) Ss[i] = S[i];
  __syncthreads();
  const int64_t n2 = n / 2;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n2; r += (int64_t)gridDim.x * blockDim.x) {
    double2 vin[MAXM];
#pragma unroll
    for (int j = 0; j < MAXM; ++j)
      vin[j] = (j < m) ? ld2<NT>(reinterpret_cast<const double2*>(V + (int64_t)j * ld) + r) : make_double2(0.0, 0.0);
    for (int c = 0; c < k; ++c) {
      double2 o = make_double2(0.0, 0.0);
#pragma unroll
      for (int j = 0; j < MAXM; ++j)
        if (j < m) {
          o.x += vin[j].x * Ss[c * m + j];
          o.y += vin[j].y * Ss[c * m + j];
        }
      st2<NT>(reinterpret_cast<double2*>(V + (int64_t)c * ld) + r, o);
    }
  }
}

// In-place V[:, 0:k] = V[:, 0:m] S (S column-major m x k), row by row.
template <int MAXM>
__global__ __launch_bounds__(kThreads) void k_restart(double* __restrict__ V, int64_t ld, int m, int k,
                                                      const double* __restrict__ S, int64_t n) {
  __shared__ double Ss[MAXM * MAXM];
  for (int i = threadIdx.x; i < m * k; i += blockDim.x) Ss[i] = S[i];
  __syncthreads();
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double vin[MAXM];
#pragma unroll
    for (int j = 0; j < MAXM; ++j) vin[j] = (j < m) ? V[(int64_t)j * ld + r] : 0.0;
    for (int c = 0; c < k; ++c) {
      double o = 0.0;
#pragma unroll
      for (int j = 0; j < MAXM; ++j)
        if (j < m) o += vin[j] * Ss[c * m + j];
      V[(int64_t)c * ld + r] = o;
    }
  }
}

// Three-term Lanczos update: p <- a y + b q + c p and partials of ||p||^2: per block-iteration a tile of
// 4 double2 per lane (all 12 loads in flight), tiles dealt grid-stride.
// On the 4x8 sector (D = 601,080,390; 32 B per row) at 1024 blocks: 3.65 ms
// (5.27 TB/s) against 3.96 ms for one double per lane over contiguous block
// ranges; 1 or 2 double2 per lane, 2048-65536 blocks and contiguous ranges
// with double2 measured 3.59-4.28 ms.  With nontemporal loads and store
// (NT, large vectors) 3.41 ms; 8 double2 per lane, 512-thread blocks, 2048
// or 8192 blocks and 32-byte loads were all slower (3.59-4.09 ms).
constexpr int kUpdU = 4;
// The same update for vectors that are not 16-byte aligned (an offset into a
// workspace of odd length): one double per lane over contiguous block ranges.
__global__ __launch_bounds__(kThreads) void k_lanczos_update_unaligned(const double* __restrict__ y,
                                                                       const double* __restrict__ q,
                                                                       double* __restrict__ p, int64_t n, double a,
                                                                       double b, double c, double* __restrict__ part) {
  int64_t lo, hi;
  block_range(n, lo, hi);
  double acc = 0.0;
  for (int64_t r = lo + threadIdx.x; r < hi; r += blockDim.x) {
    const double v = a * y[r] + b * q[r] + c * p[r];
    p[r] = v;
    acc += v * v;
  }
  double t = bsum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <bool NT>
__global__ __launch_bounds__(kThreads) void k_lanczos_update(const double* __restrict__ y,
                                                             const double* __restrict__ q, double* __restrict__ p,
                                                             int64_t n, double a, double b, double c,
                                                             double* __restrict__ part) {
  const double2* __restrict__ Y = reinterpret_cast<const double2*>(y);
  const double2* __restrict__ Q = reinterpret_cast<const double2*>(q);
  double2* __restrict__ P = reinterpret_cast<double2*>(p);
  const int64_t n2 = n / 2, tile = (int64_t)kThreads * kUpdU;
  double acc = 0.0;
  for (int64_t t0 = (int64_t)blockIdx.x * tile; t0 < n2; t0 += (int64_t)gridDim.x * tile) {
    double2 vy[kUpdU], vq[kUpdU], vp[kUpdU];
#pragma unroll
    for (int u = 0; u < kUpdU; ++u) {
      const int64_t i = t0 + u * kThreads + threadIdx.x;
      if (i < n2) {
        vy[u] = ld2<NT>(Y + i);
        vq[u] = ld2<NT>(Q + i);
        vp[u] = ld2<NT>(P + i);
      }
    }
#pragma unroll
    for (int u = 0; u < kUpdU; ++u) {
      const int64_t i = t0 + u * kThreads + threadIdx.x;
      if (i < n2) {
        double2 v;
        v.x = a * vy[u].x + b * vq[u].x + c * vp[u].x;
        v.y = a * vy[u].y + b * vq[u].y + c * vp[u].y;
        st2<NT>(P + i, v);
        acc += v.x * v.x;
        acc += v.y * v.y;
      }
    }
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {  // the odd last entry
    const int64_t r = n - 1;
    const double v = a * y[r] + b * q[r] + c * p[r];
    p[r] = v;
    acc += v * v;
  }
  double t = bsum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// part = partial ||y - lam x||^2
__global__ void k_resid(const double* __restrict__ y, const double* __restrict__ x, double lam, int64_t n,
                        double* __restrict__ part) {
  int64_t a, b;
  block_range(n, a, b);
  double acc = 0.0;
  for (int64_t r = a + threadIdx.x; r < b; r += blockDim.x) {
    const double v = y[r] - lam * x[r];
    acc += v * v;
  }
  double t = bsum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Streaming-bandwidth probes (measurement only): a read-only stream and a
// STREAM triad, in the access pattern that measured fastest on gfx950
// (16-byte loads, grid-stride, ~4 blocks per CU).
__global__ __launch_bounds__(kThreads) void k_stream_read(const double* __restrict__ b, int64_t n2,
                                                          double* __restrict__ part) {
  const double2* B = reinterpret_cast<const double2*>(b);
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = B[i];
    acc += x.x + x.y;
  }
  const double t = bsum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kThreads) void k_stream_triad(double* __restrict__ a, const double* __restrict__ b,
                                                           const double* __restrict__ c, int64_t n, double s) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = b[i] + s * c[i];
}

// Scattered-line probes (calibrate the HBM byte counters on a gather): entry
// i reads one double at the start of 2^sh-double piece ((i * P) mod m) of b,
// m = n pieces, so every piece of b is touched exactly once, in an order with
// no locality (P = 1000003 is prime, so coprime with any m it does not divide).  sh = 4: one double
// per 128-byte line; sh = 3: one per 64-byte half line.
__global__ __launch_bounds__(kThreads) void k_stream_gather(double* __restrict__ a, const double* __restrict__ b,
                                                            int64_t n, int sh) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = b[((i * 1000003) % n) << sh];
}

static int red_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (n + 255) / 256)); }
static int stream_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256)); }

// Sum `ncol` columns of partials (n each) to host.
static int finish_reduce(lsqed_ham* h, const double* part, int n, int ncol, double* dev_out, double* host_out,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_cols, dim3(ncol), dim3(kRedColThreads), 0, s, part, n, dev_out);
  LSQED_HIP(hipGetLastError());
  LSQED_HIP(hipMemcpyAsync(host_out, dev_out, ncol * sizeof(double), hipMemcpyDeviceToHost, s));
  LSQED_HIP(hipStreamSynchronize(s));
  return LSQED_OK;
}

// ------------------------------------------------- small dense eigensolver --
// Cyclic Jacobi for a symmetric n x n matrix A (row-major, destroyed):
// eigenvalues ascending in w, eigenvectors as columns of Z (row-major n x n).
// Each rotation (p, q) sets a_pq to zero and moves the diagonal by t a_pq,
// t = tan of the rotation angle (Golub & Van Loan, Matrix Computations,
// sec. 8.5); only the entries off rows/columns p, q go through the rotation
// itself.  Eigenvalues come out accurate to a few ulp of ||A||, which the
// Ritz values of a converged Krylov-Schur solve inherit.
static void sym_eig(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z) {
  Z.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
  auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i) {
      diag += a(i, i) * a(i, i);
      for (int j = i + 1; j < n; ++j) off += a(i, j) * a(i, j);
    }
    if (off == 0.0 || off <= 1e-36 * diag) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (apq == 0.0) continue;
        if (std::fabs(apq) <= 1e-18 * (std::fabs(a(p, p)) + std::fabs(a(q, q)))) {
          a(p, q) = a(q, p) = 0.0;
          continue;
        }
        const double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          if (k == p || k == q) continue;
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = a(p, k) = c * akp - s * akq;
          a(k, q) = a(q, k) = s * akp + c * akq;
        }
        a(p, p) -= t * apq;
        a(q, q) += t * apq;
        a(p, q) = a(q, p) = 0.0;
        for (int k = 0; k < n; ++k) {
          const double zkp = Z[(size_t)k * n + p], zkq = Z[(size_t)k * n + q];
          Z[(size_t)k * n + p] = c * zkp - s * zkq;
          Z[(size_t)k * n + q] = s * zkp + c * zkq;
        }
      }
  }
  // unit columns: the accumulated rotations drift from orthonormal by a few
  // ulp per sweep, and a Ritz vector's norm error becomes a relative error
  // of its Ritz value
  for (int c = 0; c < n; ++c) {
    double ss = 0.0;
    for (int r = 0; r < n; ++r) ss += Z[(size_t)r * n + c] * Z[(size_t)r * n + c];
    const double inv = 1.0 / std::sqrt(ss);
    for (int r = 0; r < n; ++r) Z[(size_t)r * n + c] *= inv;
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  std::vector<double> Zs((size_t)n * n);
  w.resize(n);
  for (int c = 0; c < n; ++c) {
    w[c] = a(idx[c], idx[c]);
    for (int r = 0; r < n; ++r) Zs[(size_t)r * n + c] = Z[(size_t)r * n + idx[c]];
  }
  Z.swap(Zs);
}

// Lowest eigenvalue of the symmetric tridiagonal (alpha, beta) by Sturm
// bisection.
static double tridiag_lowest(const std::vector<double>& al, const std::vector<double>& be) {
  const int n = (int)al.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < n; ++i) {
    double r = (i > 0 ? std::fabs(be[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(be[i]) : 0.0);
    lo = std::min(lo, al[i] - r);
    hi = std::max(hi, al[i] + r);
  }
  auto count_below = [&](double x) {
    int cnt = 0;
    double d = 1.0;
    for (int i = 0; i < n; ++i) {
      const double b2 = i > 0 ? be[i - 1] * be[i - 1] : 0.0;
      d = al[i] - x - (i > 0 ? b2 / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0) ++cnt;
    }
    return cnt;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(lo)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) >= 1)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

}  // namespace lsqed

using namespace lsqed;

struct lsqed_eps {
  lsqed_ham* h = nullptr;
  int nev = 1, ncv = 0, mpd = 0, maxit = 0;
  double tol = 1e-8;
  uint64_t seed = 1;
  int verify_rounds = 0;   // multiplicity check rounds (0: off, SLEPc's stopping rule)
  double* V = nullptr;     // vcols columns of D: ncv + 1, or ncv + 2 (a spare for the delayed CGS2)
  int vcols = 0;
  double* part = nullptr;  // kMaxNcv * kRedBlocks partials
  double* dsm = nullptr;   // small device buffer: coefficients / S matrix
  std::vector<double> evals;
  int nconv = 0, its = 0;
  bool solved = false;
  // Renumbered copy of the model whose products the solver runs (relabel.hip;
  // null: h itself), perm[old site] = new site, and on the device perm then
  // its inverse (N entries each).
  lsqed_ham* hk = nullptr;
  std::vector<int> perm;
  int* d_perm = nullptr;
};

// The model the solver's own products use: its renumbered copy when it has one.
static lsqed_ham* kernel_model(const lsqed_eps* e) { return e->hk ? e->hk : e->h; }

namespace {

struct Ctx {
  lsqed_eps* e;
  hipStream_t s;
  int64_t D;
};

// h = V[:, 0:m]^T w (host out), device copy left in e->dsm[0:m].
int multidot(Ctx& c, int m, const double* w, std::vector<double>& h) {
  const int G = red_blocks(c.D);
  const int groups = (m + 15) / 16;
  hipLaunchKernelGGL(k_multidot, dim3(G, groups), dim3(kThreads), 0, c.s, c.e->V, c.D, m, w, c.D, c.e->part);
  LSQED_HIP(hipGetLastError());
  h.resize(m);
  return finish_reduce(c.e->h, c.e->part, G, m, c.e->dsm, h.data(), c.s);
}

// w -= V[:, 0:m] h ; returns ||w|| (h given on host, uploaded).
int update(Ctx& c, int m, double* w, const std::vector<double>& h, double& nrm) {
  const int G = red_blocks(c.D);
  LSQED_HIP(hipMemcpyAsync(c.e->dsm, h.data(), m * sizeof(double), hipMemcpyHostToDevice, c.s));
  hipLaunchKernelGGL(k_update, dim3(G), dim3(kThreads), 0, c.s, c.e->V, c.D, m, c.e->dsm, w, c.D, c.e->part);
  LSQED_HIP(hipGetLastError());
  double ss = 0.0;
  int rc = finish_reduce(c.e->h, c.e->part, G, 1, c.e->dsm + kMaxNcv, &ss, c.s);
  if (rc) return rc;
  nrm = std::sqrt(ss);
  return LSQED_OK;
}

int norm2(Ctx& c, const double* x, double& nrm) {
  const int G = red_blocks(c.D);
  hipLaunchKernelGGL(k_dot, dim3(G), dim3(kThreads), 0, c.s, x, x, c.D, c.e->part);
  LSQED_HIP(hipGetLastError());
  double ss = 0.0;
  int rc = finish_reduce(c.e->h, c.e->part, G, 1, c.e->dsm, &ss, c.s);
  nrm = std::sqrt(ss);
  return rc;
}

int scale(Ctx& c, double* x, double a) {
  hipLaunchKernelGGL(k_scale, dim3(stream_grid(c.D)), dim3(kThreads), 0, c.s, x, c.D, a);
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

// Classical Gram-Schmidt, twice, of w against V[:, 0:m]; h accumulates both
// passes.  nrm1 / nrm2 are ||w|| after the first / second pass; the second
// pass also decides linear dependence the way SLEPc's refined CGS does
// (BV_ORTHOG_REFINE_IFNEEDED): w is dependent on V when the refinement pass
// still removes more than 1 - 1/sqrt(2) of its norm.  A residual made of
// rounding noise is mostly outside span(V), survives the test and becomes the
// next basis vector -- which is how degenerate copies of an eigenvalue enter
// the Krylov space.
template <int MAXM, int MODE>
void launch_cgs_pass(Ctx& c, int G, int m, const double* hdev, double* w, double sc) {
  hipLaunchKernelGGL((k_cgs_pass<MAXM, MODE>), dim3(G), dim3(kThreads), 0, c.s, c.e->V, c.D, m, hdev, w, c.D,
                     c.e->part, sc);
}

// CGS2 in three passes over V for m <= 32 (projection; first update fused
// with the second projection; second update), each with all column loads of
// a row in flight.  With `normalise`, the second update also divides by the
// norm it leaves, known before the pass: w2 = w1 - V h2 with h2 = V^T w1 and
// V orthonormal gives ||w2||^2 = ||w1||^2 - ||h2||^2, without cancellation
// since the refinement coefficients h2 are small next to w1 unless w is
// dependent on V (then the pass does not scale).  *scaled reports which;
// nrm is the divisor, so that w_true = nrm * w exactly.
int cgs2_fused(Ctx& c, int m, double* w, std::vector<double>& h1, std::vector<double>& h2, double& nrm1,
               double& nrm, bool normalise = false, bool* scaled = nullptr) {
  const int G = red_blocks(c.D);
  double* hdev = c.e->dsm + kMaxNcv * kMaxNcv;  // coefficient staging (2 kMaxNcv doubles)
  std::vector<double> out(m + 1);
  int rc;
  auto pass = [&](int mode, const std::vector<double>* hc, double sc) -> int {
    if (hc) LSQED_HIP(hipMemcpyAsync(hdev, hc->data(), m * sizeof(double), hipMemcpyHostToDevice, c.s));
    if (m <= 16) {
      if (mode == 0) launch_cgs_pass<16, 0>(c, G, m, hdev, w, sc);
      if (mode == 1) launch_cgs_pass<16, 1>(c, G, m, hdev, w, sc);
      if (mode == 2) launch_cgs_pass<16, 2>(c, G, m, hdev, w, sc);
    } else if (m <= 24) {
      if (mode == 0) launch_cgs_pass<24, 0>(c, G, m, hdev, w, sc);
      if (mode == 1) launch_cgs_pass<24, 1>(c, G, m, hdev, w, sc);
      if (mode == 2) launch_cgs_pass<24, 2>(c, G, m, hdev, w, sc);
    } else {
      if (mode == 0) launch_cgs_pass<32, 0>(c, G, m, hdev, w, sc);
      if (mode == 1) launch_cgs_pass<32, 1>(c, G, m, hdev, w, sc);
      if (mode == 2) launch_cgs_pass<32, 2>(c, G, m, hdev, w, sc);
    }
    LSQED_HIP(hipGetLastError());
    const int ncol = mode == 0 ? m : (mode == 1 ? m + 1 : 1);
    return finish_reduce(c.e->h, c.e->part, G, ncol, c.e->dsm, out.data(), c.s);
  };
  if ((rc = pass(0, nullptr, 1.0))) return rc;
  h1.assign(out.begin(), out.begin() + m);
  if ((rc = pass(1, &h1, 1.0))) return rc;
  h2.assign(out.begin(), out.begin() + m);
  nrm1 = std::sqrt(out[m]);
  double beta = 0.0;
  if (normalise) {
    double hh = 0.0;
    for (int j = 0; j < m; ++j) hh += h2[j] * h2[j];
    const double b2 = out[m] - hh;
    if (b2 > 0.0 && std::sqrt(b2) >= 0.7071067811865476 * nrm1) beta = std::sqrt(b2);
  }
  if ((rc = pass(2, &h2, beta > 0.0 ? 1.0 / beta : 1.0))) return rc;
  if (scaled) *scaled = beta > 0.0;
  nrm = beta > 0.0 ? beta : std::sqrt(out[0]);
  return LSQED_OK;
}

// With `scaled` non-null, w comes back normalised when *scaled is set (see
// cgs2_fused); otherwise w is left unscaled with norm nrm.
int cgs2(Ctx& c, int m, double* w, std::vector<double>& h, double& nrm, bool* lindep = nullptr,
         bool* scaled = nullptr) {
  std::vector<double> h1, h2;
  int rc;
  if (lindep) *lindep = false;
  if (scaled) *scaled = false;
  if (m == 0) {
    rc = norm2(c, w, nrm);
    if (lindep) *lindep = !(nrm > 0.0);
    return rc;
  }
  double nrm1 = 0.0;
  if (m <= 32) {
    if ((rc = cgs2_fused(c, m, w, h1, h2, nrm1, nrm, scaled != nullptr, scaled))) return rc;
  } else {
    if ((rc = multidot(c, m, w, h1))) return rc;
    if ((rc = update(c, m, w, h1, nrm1))) return rc;
    if ((rc = multidot(c, m, w, h2))) return rc;
    if ((rc = update(c, m, w, h2, nrm))) return rc;
  }
  h.resize(m);
  for (int j = 0; j < m; ++j) h[j] = h1[j] + h2[j];
  if (lindep) *lindep = !(nrm > 0.0 && nrm >= 0.7071067811865476 * nrm1);
  return LSQED_OK;
}

// The two-row kernels apply when every column and vector is 16-byte aligned
// (even leading dimension; LSQED_KS_TWO_ROWS=0 keeps the one-row kernels).
static bool two_rows(const Ctx& c, const double* u, const double* y) {
  static const bool on = [] {
    const char* e = std::getenv("LSQED_KS_TWO_ROWS");
    return !(e && e[0] == '0');
  }();
  const uintptr_t bits = reinterpret_cast<uintptr_t>(c.e->V) | reinterpret_cast<uintptr_t>(u) |
                         reinterpret_cast<uintptr_t>(y);
  return on && (c.D % 2 == 0) && (bits & 15) == 0;
}

// Nontemporal loads and stores for the streaming passes over vectors of
// 2^25 doubles (256 MB, the memory-side cache) or more, which no later pass
// finds in a cache anyway; LSQED_NT_STREAMS=0/1 overrides.  4x8 (601,080,390
// rows) Lanczos update: 3.64 -> 3.41 ms (5.29 -> 5.64 TB/s, standalone sweep:
// nontemporal loads of y and q alone 3.53 ms, all three loads without the
// nontemporal store 3.63 ms); in the traced 4x8 Krylov-Schur solve the
// update 3.65 -> 3.32 ms, the delayed CGS2 projection 11.92 -> 11.36 ms
// (16 columns) and 15.47 -> 15.21 ms (24), the restart 34.27 -> 33.89 ms;
// 6x6 update 55.7 -> 54.0 ms.  The delayed CGS2 update pass, which rewrites
// u and y in place, is left temporal: 17.24 -> 18.15 ms and 19.92 ->
// 22.28 ms with the nontemporal forms.
static bool nt_streams(int64_t n) {
  static const int env = [] {
    const char* e = std::getenv("LSQED_NT_STREAMS");
    return e ? (std::atoi(e) != 0 ? 1 : 0) : -1;
  }();
  return env >= 0 ? env == 1 : n >= ((int64_t)1 << 25);
}

// V[:, 0:k] = V[:, 0:m] S[:, 0:k] (S host row-major m x m, first k columns).
int restart_transform(Ctx& c, int m, int k, const std::vector<double>& S) {
  std::vector<double> Sc((size_t)m * k);
  for (int col = 0; col < k; ++col)
    for (int j = 0; j < m; ++j) Sc[(size_t)col * m + j] = S[(size_t)j * m + col];
  LSQED_HIP(hipMemcpyAsync(c.e->dsm, Sc.data(), Sc.size() * sizeof(double), hipMemcpyHostToDevice, c.s));
  const int g = stream_grid(c.D);
  if (two_rows(c, c.e->V, c.e->V)) {
    const bool nt = nt_streams(c.D);
    if (m <= 16)
      hipLaunchKernelGGL((nt ? k_restart2<16, true> : k_restart2<16, false>), dim3(g), dim3(kThreads), 0, c.s, c.e->V,
                         c.D, m, k, c.e->dsm, c.D);
    else if (m <= 32)
      hipLaunchKernelGGL((nt ? k_restart2<32, true> : k_restart2<32, false>), dim3(g), dim3(kThreads), 0, c.s, c.e->V,
                         c.D, m, k, c.e->dsm, c.D);
    else
      hipLaunchKernelGGL((nt ? k_restart2<64, true> : k_restart2<64, false>), dim3(g), dim3(kThreads), 0, c.s, c.e->V,
                         c.D, m, k, c.e->dsm, c.D);
  } else if (m <= 16)
    hipLaunchKernelGGL(k_restart<16>, dim3(g), dim3(kThreads), 0, c.s, c.e->V, c.D, m, k, c.e->dsm, c.D);
  else if (m <= 32)
    hipLaunchKernelGGL(k_restart<32>, dim3(g), dim3(kThreads), 0, c.s, c.e->V, c.D, m, k, c.e->dsm, c.D);
  else
    hipLaunchKernelGGL(k_restart<64>, dim3(g), dim3(kThreads), 0, c.s, c.e->V, c.D, m, k, c.e->dsm, c.D);
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

// Delayed CGS2 passes (k_dcgs_proj / k_dcgs_update) for j <= 32 basis columns.
template <int MAXM>
void launch_dcgs_proj(Ctx& c, int G, int j, const double* u, const double* y) {
  if (two_rows(c, u, y)) {
    const bool nt = nt_streams(c.D);
    if (y)
      hipLaunchKernelGGL((nt ? k_dcgs_proj2<MAXM, true, true> : k_dcgs_proj2<MAXM, true, false>), dim3(G),
                         dim3(kThreads), 0, c.s, c.e->V, c.D, j, u, y, c.D, c.e->part);
    else
      hipLaunchKernelGGL((nt ? k_dcgs_proj2<MAXM, false, true> : k_dcgs_proj2<MAXM, false, false>), dim3(G),
                         dim3(kThreads), 0, c.s, c.e->V, c.D, j, u, y, c.D, c.e->part);
    return;
  }
  if (y)
    hipLaunchKernelGGL((k_dcgs_proj<MAXM, true>), dim3(G), dim3(kThreads), 0, c.s, c.e->V, c.D, j, u, y, c.D,
                       c.e->part);
  else
    hipLaunchKernelGGL((k_dcgs_proj<MAXM, false>), dim3(G), dim3(kThreads), 0, c.s, c.e->V, c.D, j, u, y, c.D,
                       c.e->part);
}

// out = [V_i.u (j)] [V_i.y (j)] [u.u] [u.y] [y.y], or without y [V_i.u (j)] [u.u].
int dcgs_project(Ctx& c, int j, const double* u, const double* y, std::vector<double>& out) {
  const int G = red_blocks(c.D);
  if (j <= 8) launch_dcgs_proj<8>(c, G, j, u, y);
  else if (j <= 16) launch_dcgs_proj<16>(c, G, j, u, y);
  else if (j <= 24) launch_dcgs_proj<24>(c, G, j, u, y);
  else launch_dcgs_proj<32>(c, G, j, u, y);
  LSQED_HIP(hipGetLastError());
  const int nq = y ? 2 * j + 3 : j + 1;
  out.resize(nq);
  return finish_reduce(c.e->h, c.e->part, G, nq, c.e->dsm, out.data(), c.s);
}

// yo: where the new residual goes (nullptr or y: over y).
template <int MAXM>
void launch_dcgs_update(Ctx& c, int G, int j, const double* coef, double inv_nu, double ya, double ub, double* u,
                        double* y, double* yo) {
  const bool oop = y && yo && yo != y;
  if (two_rows(c, u, y) && two_rows(c, u, yo)) {
    if (oop)
      hipLaunchKernelGGL((k_dcgs_update2<MAXM, true, true>), dim3(G), dim3(kThreads), 0, c.s, c.e->V, c.D, j, coef,
                         inv_nu, ya, ub, u, y, yo, c.D, c.e->part);
    else if (y)
      hipLaunchKernelGGL((k_dcgs_update2<MAXM, true, false>), dim3(G), dim3(kThreads), 0, c.s, c.e->V, c.D, j,
                         coef, inv_nu, ya, ub, u, y, y, c.D, c.e->part);
    else
      hipLaunchKernelGGL((k_dcgs_update2<MAXM, false, false>), dim3(G), dim3(kThreads), 0, c.s, c.e->V, c.D, j,
                         coef, inv_nu, ya, ub, u, y, y, c.D, c.e->part);
    return;
  }
  if (oop)
    hipLaunchKernelGGL((k_dcgs_update<MAXM, true, true>), dim3(G), dim3(kThreads), 0, c.s, c.e->V, c.D, j, coef,
                       inv_nu, ya, ub, u, y, yo, c.D, c.e->part);
  else if (y)
    hipLaunchKernelGGL((k_dcgs_update<MAXM, true, false>), dim3(G), dim3(kThreads), 0, c.s, c.e->V, c.D, j, coef,
                       inv_nu, ya, ub, u, y, y, c.D, c.e->part);
  else
    hipLaunchKernelGGL((k_dcgs_update<MAXM, false, false>), dim3(G), dim3(kThreads), 0, c.s, c.e->V, c.D, j, coef,
                       inv_nu, ya, ub, u, y, y, c.D, c.e->part);
}

// u <- (u - V c) inv_nu; with y: yo <- ya y + ub u_old - V z (yo nullptr:
// over y) and *yy = ||yo||^2.  coef = c (j values) then, with y, z (j values).
int dcgs_update(Ctx& c, int j, const std::vector<double>& coef, double inv_nu, double ya, double ub, double* u,
                double* y, double* yy, double* yo = nullptr) {
  const int G = red_blocks(c.D);
  double* cdev = c.e->dsm + kMaxNcv * kMaxNcv;  // coefficient staging (2 kMaxNcv doubles)
  if (!coef.empty())
    LSQED_HIP(hipMemcpyAsync(cdev, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, c.s));
  if (j <= 8) launch_dcgs_update<8>(c, G, j, cdev, inv_nu, ya, ub, u, y, yo);
  else if (j <= 16) launch_dcgs_update<16>(c, G, j, cdev, inv_nu, ya, ub, u, y, yo);
  else if (j <= 24) launch_dcgs_update<24>(c, G, j, cdev, inv_nu, ya, ub, u, y, yo);
  else launch_dcgs_update<32>(c, G, j, cdev, inv_nu, ya, ub, u, y, yo);
  LSQED_HIP(hipGetLastError());
  if (!y) return LSQED_OK;
  return finish_reduce(c.e->h, c.e->part, G, 1, c.e->dsm, yy, c.s);
}

// Delayed CGS2 replaces the refined CGS2 on sectors of 2^16 states or more:
// two passes over the basis per Krylov step instead of three (and one host
// synchronisation fewer).  Smaller sectors -- all those of the reference's
// solver tests -- keep the refined CGS2 of SLEPc, whose rounding behaviour
// (which copies of a degenerate level enter the space) those tests pin.
// Measured: 4x8 nev=4 ncv=20 29.2 -> 25.0 s, 4x6 nev=10 ncv=32 0.144 ->
// 0.115 s, the per-realisation loop over chains of 18-22 spins 76.5 -> 93.5
// realisations/s.  LSQED_KS_DELAYED=0/1 forces either form.
bool use_delayed(const lsqed_eps* e, int64_t D) {
  if (const char* v = std::getenv("LSQED_KS_DELAYED")) return std::atoi(v) == 1 && e->ncv <= 32;
  return D >= ((int64_t)1 << 16) && e->ncv <= 32;
}

// New random start vector for column k, orthonormal to V[:, 0:k].
int fresh_vector(Ctx& c, int k, uint64_t seed) {
  double* v = c.e->V + (int64_t)k * c.D;
  hipLaunchKernelGGL(k_fill_random, dim3(stream_grid(c.D)), dim3(kThreads), 0, c.s, v, c.D, seed);
  LSQED_HIP(hipGetLastError());
  std::vector<double> h;
  double nrm = 0.0;
  int rc = cgs2(c, k, v, h, nrm);
  if (rc) return rc;
  if (!(nrm > 0.0)) return fail(LSQED_ERR_STATE, "start vector vanished after orthogonalisation");
  return scale(c, v, 1.0 / nrm);
}

}  // namespace

extern "C" {

int lsqed_eps_create(lsqed_ham* h, const lsqed_eps_opts* o, lsqed_eps** out) {
  if (!h || !out) return fail(LSQED_ERR_ARG, "null argument");
  *out = nullptr;
  lsqed_eps_opts d{};
  d.nev = 1;
  if (o) d = *o;
  if (d.nev <= 0) d.nev = 1;
  const int64_t D = h->D;
  if (D <= 0) return fail(LSQED_ERR_ARG, "empty sector");
  if (d.nev > D) return fail(LSQED_ERR_ARG, "nev larger than the sector dimension");
  int ncv = d.ncv > 0 ? d.ncv : std::max(2 * d.nev, d.nev + 15);  // SLEPc default
  ncv = (int)std::min<int64_t>(ncv, D);
  if (ncv > kMaxNcv) return fail(LSQED_ERR_ARG, "ncv > 64 is not supported");
  if (ncv < d.nev) return fail(LSQED_ERR_ARG, "ncv must be >= nev");
  lsqed_eps* e = new lsqed_eps;
  e->h = h;
  e->nev = d.nev;
  e->ncv = ncv;
  e->mpd = d.mpd > 0 ? d.mpd : ncv;
  e->tol = d.tol > 0 ? d.tol : 1e-8;
  e->maxit = d.maxit > 0 ? d.maxit : (int)std::max<int64_t>(100, 2 * D / ncv);
  e->seed = d.seed ? d.seed : 1;
  e->verify_rounds = std::max(0, d.multiplicity_rounds);
  // The delayed CGS2 takes one more column when the device has room for it
  // (1 GiB kept free): A u goes there, so that the update pass writes the
  // next residual to a column it does not read.  Without it the solve runs
  // the same steps in place (LSQED_KS_SPARE=0 forces that).
  const size_t col_bytes = (size_t)D * sizeof(double);
  e->vcols = ncv + 1;
  const char* sp = std::getenv("LSQED_KS_SPARE");
  if (use_delayed(e, D) && !(sp && std::atoi(sp) == 0)) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b >= (size_t)(ncv + 2) * col_bytes + ((size_t)1 << 30))
      e->vcols = ncv + 2;
    else
      (void)hipGetLastError();
  }
  if (hipMalloc((void**)&e->V, (size_t)e->vcols * col_bytes) != hipSuccess) {
    (void)hipGetLastError();
    e->vcols = ncv + 1;
    if (hipMalloc((void**)&e->V, (size_t)e->vcols * col_bytes) != hipSuccess) {
      (void)hipGetLastError();
      delete e;
      return fail(LSQED_ERR_OOM, "cannot allocate the Krylov basis ((ncv+1) x D doubles)");
    }
  }
  if (hipMalloc((void**)&e->part, (size_t)(2 * kMaxNcv + 4) * kRedBlocks * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&e->dsm, (size_t)(kMaxNcv * kMaxNcv + 2 * kMaxNcv) * sizeof(double)) != hipSuccess) {
    lsqed_eps_destroy(e);
    return fail(LSQED_ERR_OOM, "cannot allocate solver scratch");
  }
  // Products on a renumbered copy of the model where that puts more bonds in
  // the specialised kernel's low block (relabel.hip); its kernel is compiled
  // here, not inside the solve.  Kept only when the copy takes the
  // specialised kernel; otherwise the solver runs on h.
  std::vector<int> perm;
  if (relabel_plan(h, perm)) {
    lsqed_ham* hk = nullptr;
    std::vector<int> both(perm);
    both.resize(2 * perm.size());
    for (size_t i = 0; i < perm.size(); ++i) both[perm.size() + perm[i]] = (int)i;
    const bool ok = relabel_handle(h, perm, &hk) == LSQED_OK && lsqed_ham_get_hv_kernel(hk) == 3 &&
                    hv_launch_grid(hk) > 0 &&
                    hipMalloc((void**)&e->d_perm, both.size() * sizeof(int)) == hipSuccess &&
                    hipMemcpy(e->d_perm, both.data(), both.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
      e->hk = hk;
      e->perm = perm;
    } else {
      (void)hipGetLastError();
      if (hk) lsqed_ham_destroy(hk);
      if (e->d_perm) (void)hipFree(e->d_perm);
      e->d_perm = nullptr;
    }
  }
  *out = e;
  return LSQED_OK;
}

int lsqed_eps_destroy(lsqed_eps* e) {
  if (!e) return LSQED_OK;
  (void)hipFree(e->V);
  (void)hipFree(e->part);
  (void)hipFree(e->dsm);
  if (e->hk) lsqed_ham_destroy(e->hk);
  if (e->d_perm) (void)hipFree(e->d_perm);
  delete e;
  return LSQED_OK;
}

int lsqed_eps_iterations(const lsqed_eps* e) { return e ? e->its : 0; }

int64_t lsqed_eps_workspace_bytes(const lsqed_eps* e) {
  if (!e) return 0;
  return (int64_t)e->vcols * e->h->D * (int64_t)sizeof(double) +
         (int64_t)((2 * kMaxNcv + 4) * kRedBlocks + kMaxNcv * kMaxNcv + 2 * kMaxNcv) * (int64_t)sizeof(double);
}

int lsqed_eps_solve(lsqed_eps* e, int* nconv_out, lsqed_stream_t stream) {
  if (!e) return fail(LSQED_ERR_ARG, "null solver");
  Ctx c{e, (hipStream_t)stream, e->h->D};
  const int m = e->ncv;
  if (e->hk) {
    const int rs = relabel_sync(e->h, e->hk, e->perm);  // the model may have changed since create
    if (rs) return rs;
  }
  const int64_t D = c.D;
  int rc;
  auto col = [&](int j) { return e->V + (int64_t)j * D; };

  // start vector (SLEPc draws a random one; ours is a seeded hash), drawn in
  // the model's basis order also when the products run renumbered (then
  // gathered into the renumbered order), so that the solve starts from the
  // same vector either way
  if (e->hk && e->vcols >= 2) {
    hipLaunchKernelGGL(k_fill_random, dim3(stream_grid(D)), dim3(kThreads), 0, c.s, col(1), D, e->seed);
    LSQED_HIP(hipGetLastError());
    if ((rc = relabel_gather(e->h, e->d_perm + e->h->N, col(1), col(0), c.s))) return rc;
  } else {
    hipLaunchKernelGGL(k_fill_random, dim3(stream_grid(D)), dim3(kThreads), 0, c.s, col(0), D, e->seed);
    LSQED_HIP(hipGetLastError());
  }
  double nrm;
  if ((rc = norm2(c, col(0), nrm))) return rc;
  if ((rc = scale(c, col(0), 1.0 / nrm))) return rc;

  std::vector<double> T((size_t)m * m, 0.0), arrow;
  std::vector<double> beta(m + 1, 0.0);
  const bool delayed = use_delayed(e, D);
  int k = 0;
  e->its = 0;
  e->nconv = 0;
  e->evals.clear();
  uint64_t reseed = e->seed;
  // Multiplicity check: a single start vector spans one direction per
  // distinct eigenvalue; further copies of a degenerate level enter only
  // through rounding and can be missed when the wanted values converge
  // first.  Once nev values have converged, they are locked and the
  // iteration continues from a random vector orthogonal to them until one
  // more value has converged; if the lowest nev changed (a missed copy
  // surfaced), the check is repeated.
  int rounds_left = e->verify_rounds;
  int target = e->nev;
  std::vector<double> prev;
  while (true) {
    int meff = m;
    bool breakdown = false;
    if (delayed) {
      // Delayed CGS2.  Column j holds u, the residual of step j - 1
      // orthogonalised once (at j = k the normalised restart vector), with
      // A Q = Q T + u f^T for Q = V[:, 0:j] (f = e_{j-1}, or the arrow at
      // j = k).  One pass gives c = Q^T u, h = Q^T y (y = A u), u.u, u.y; with
      // nu = ||u - Q c|| = sqrt(u.u - c.c) and phi = f.c the step is
      //   v_j   = (u - Q c) / nu                       (second orthogonalisation)
      //   alpha = v_j^T A v_j = (u.y - c.h) / nu^2 - phi
      //   r     = y / nu - ((phi + alpha) / nu) u - Q (h - (phi + alpha) c) / nu
      // (r: the next residual, orthogonalised once), written by one update
      // pass.  beta_{j-1} = nu.  u is dependent on Q (SLEPc's refinement
      // test, nu < ||u|| / sqrt 2): invariant subspace.
      // A u goes to the spare column when there is one (the update then
      // writes r to column j + 1 without reading it), else to column j + 1.
      std::vector<double> q, coef;
      double* au = e->vcols >= m + 2 ? col(m + 1) : nullptr;
      for (int j = k; j < m; ++j) {
        double* y = au ? au : col(j + 1);
        if ((rc = launch_hv(kernel_model(e), col(j), y, nullptr, c.s))) return rc;
        if ((rc = dcgs_project(c, j, col(j), y, q))) return rc;
        const double* cc = q.data();
        const double* hh = q.data() + j;
        const double gam = q[2 * j], del = q[2 * j + 1];
        double cn = 0.0, ch = 0.0;
        for (int i = 0; i < j; ++i) {
          cn += cc[i] * cc[i];
          ch += cc[i] * hh[i];
        }
        const double nu2 = gam - cn;
        const double nu = nu2 > 0.0 ? std::sqrt(nu2) : 0.0;
        if (!(nu > 0.0 && nu >= 0.7071067811865476 * std::sqrt(gam))) {
          meff = j;
          breakdown = true;
          break;
        }
        double phi = 0.0;
        if (j > k) phi = cc[j - 1];
        else
          for (int i = 0; i < k; ++i) phi += arrow[i] * cc[i];
        const double alpha = (del - ch) / nu2 - phi;
        coef.assign(2 * (size_t)j, 0.0);
        for (int i = 0; i < j; ++i) {
          coef[i] = cc[i];
          coef[j + i] = (hh[i] - (phi + alpha) * cc[i]) / nu;
        }
        double rr = 0.0;
        if ((rc = dcgs_update(c, j, coef, 1.0 / nu, 1.0 / nu, -(phi + alpha) / nu, col(j), y, &rr, col(j + 1))))
          return rc;
        T[(size_t)j * m + j] = alpha;
        if (j == k) {
          for (int i = 0; i < k; ++i) T[(size_t)i * m + k] = T[(size_t)k * m + i] = arrow[i];
        } else {
          T[(size_t)j * m + j - 1] = T[(size_t)(j - 1) * m + j] = nu;
          beta[j - 1] = nu;
        }
        if ((int64_t)(j + 1) >= D) {
          meff = j + 1;
          breakdown = true;
          break;
        }
      }
      if (!breakdown) {
        // the last residual, orthogonalised once: second orthogonalisation
        // and normalisation (it becomes the restart vector), beta_{m-1}
        if ((rc = dcgs_project(c, m, col(m), nullptr, q))) return rc;
        double cn = 0.0;
        for (int i = 0; i < m; ++i) cn += q[i] * q[i];
        const double gam = q[m], nu2 = gam - cn;
        const double nu = nu2 > 0.0 ? std::sqrt(nu2) : 0.0;
        if (!(nu > 0.0 && nu >= 0.7071067811865476 * std::sqrt(gam))) {
          breakdown = true;
        } else {
          beta[m - 1] = nu;
          coef.assign(q.begin(), q.begin() + m);
          if ((rc = dcgs_update(c, m, coef, 1.0 / nu, 0.0, 0.0, col(m), nullptr, nullptr))) return rc;
        }
      }
    }
    for (int j = k; j < m && !delayed; ++j) {
      if ((rc = launch_hv(kernel_model(e), col(j), col(j + 1), nullptr, c.s))) return rc;
      std::vector<double> hcoef;
      double b;
      bool lindep = false, scaled = false;
      if ((rc = cgs2(c, j + 1, col(j + 1), hcoef, b, &lindep, &scaled))) return rc;
      T[(size_t)j * m + j] = hcoef[j];
      if (j == k) {
        for (int i = 0; i < k; ++i) T[(size_t)i * m + k] = T[(size_t)k * m + i] = arrow[i];
      } else {
        T[(size_t)j * m + j - 1] = T[(size_t)(j - 1) * m + j] = beta[j - 1];
      }
      beta[j] = b;
      if (lindep || (int64_t)(j + 1) >= D) {
        meff = j + 1;
        breakdown = true;
        break;
      }
      if (!scaled && (rc = scale(c, col(j + 1), 1.0 / b))) return rc;
    }
    ++e->its;
    // projected problem
    std::vector<double> A((size_t)meff * meff), w, S;
    for (int i = 0; i < meff; ++i)
      for (int j = 0; j < meff; ++j) A[(size_t)i * meff + j] = T[(size_t)i * m + j];
    sym_eig(meff, A, w, S);
    std::vector<double> res(meff);
    for (int i = 0; i < meff; ++i)
      res[i] = breakdown ? 0.0 : std::fabs(beta[meff - 1] * S[(size_t)(meff - 1) * meff + i]);
    int nconv = 0;
    while (nconv < meff && res[nconv] <= e->tol * std::max(std::fabs(w[nconv]), 1e-300)) ++nconv;
    bool enough = nconv >= target;
    const bool full = breakdown && meff >= m;
    if (enough && rounds_left > 0 && !full && meff < D && e->nev + 2 <= m && (int64_t)e->nev + 1 < D &&
        e->its < e->maxit) {
      bool changed = prev.empty();
      for (size_t i = 0; i < prev.size() && !changed; ++i)
        changed = std::fabs(w[i] - prev[i]) > std::max(100.0 * e->tol, 1e-10) * std::max(std::fabs(prev[i]), 1.0);
      if (changed) {
        prev.assign(w.begin(), w.begin() + e->nev);
        --rounds_left;
        if ((rc = restart_transform(c, meff, e->nev, S))) return rc;
        k = e->nev;
        arrow.assign(k, 0.0);
        std::fill(T.begin(), T.end(), 0.0);
        for (int i = 0; i < k; ++i) T[(size_t)i * m + i] = w[i];
        if ((rc = fresh_vector(c, k, ++reseed))) return rc;
        target = e->nev + 1;
        continue;
      }
    }
    if (!enough && nconv >= e->nev && e->its >= e->maxit) enough = true;
    if (enough || e->its >= e->maxit || full) {
      int kfin = enough ? nconv : std::max(nconv, std::min(e->nev, meff));
      kfin = std::min(kfin, meff);
      if ((rc = restart_transform(c, meff, kfin, S))) return rc;
      e->nconv = nconv;
      e->evals.assign(w.begin(), w.begin() + kfin);
      break;
    }
    if (breakdown) {
      // Invariant subspace: keep every Ritz vector, continue from a fresh
      // random direction orthogonal to them (SLEPc restarts the same way).
      if ((rc = restart_transform(c, meff, meff, S))) return rc;
      k = meff;
      arrow.assign(k, 0.0);
      std::fill(T.begin(), T.end(), 0.0);
      for (int i = 0; i < k; ++i) T[(size_t)i * m + i] = w[i];
      if ((rc = fresh_vector(c, k, ++reseed))) return rc;
      continue;
    }
    int keep = nconv + std::max(1, (meff - nconv) / 2);
    keep = std::min(keep, meff - 1);
    if ((rc = restart_transform(c, meff, keep, S))) return rc;
    LSQED_HIP(hipMemcpyAsync(col(keep), col(meff), D * sizeof(double), hipMemcpyDeviceToDevice, c.s));
    arrow.resize(keep);
    for (int i = 0; i < keep; ++i) arrow[i] = beta[meff - 1] * S[(size_t)(meff - 1) * meff + i];
    std::fill(T.begin(), T.end(), 0.0);
    for (int i = 0; i < keep; ++i) T[(size_t)i * m + i] = w[i];
    k = keep;
  }
  LSQED_HIP(hipStreamSynchronize(c.s));
  e->solved = true;
  if (nconv_out) *nconv_out = e->nconv;
  return LSQED_OK;
}

int lsqed_eps_solution(lsqed_eps* e, int i, double* lambda, double* x, double* error, lsqed_stream_t stream) {
  if (!e) return fail(LSQED_ERR_ARG, "null solver");
  if (!e->solved) return fail(LSQED_ERR_STATE, "lsqed_eps_solve has not run");
  if (i < 0 || i >= (int)e->evals.size()) return fail(LSQED_ERR_ARG, "eigenpair index out of range");
  Ctx c{e, (hipStream_t)stream, e->h->D};
  const int64_t D = c.D;
  const double lam = e->evals[i];
  if (lambda) *lambda = lam;
  double* vi = e->V + (int64_t)i * D;
  if (x && e->hk) {
    const int rg = relabel_gather(e->h, e->d_perm, vi, x, c.s);  // back to the model's basis order
    if (rg) return rg;
  } else if (x) {
    LSQED_HIP(hipMemcpyAsync(x, vi, D * sizeof(double), hipMemcpyDeviceToDevice, c.s));
  }
  if (error && e->hk) {
    const int rs = relabel_sync(e->h, e->hk, e->perm);
    if (rs) return rs;
  }
  if (error) {
    // EPSComputeError(EPS_ERROR_RELATIVE): ||H x - lambda x|| / |lambda|
    double* y = e->V + (int64_t)e->ncv * D;
    int rc = launch_hv(kernel_model(e), vi, y, nullptr, c.s);
    if (rc) return rc;
    const int G = red_blocks(D);
    hipLaunchKernelGGL(k_resid, dim3(G), dim3(kThreads), 0, c.s, y, vi, lam, D, e->part);
    LSQED_HIP(hipGetLastError());
    double ss = 0.0;
    if ((rc = finish_reduce(e->h, e->part, G, 1, e->dsm, &ss, c.s))) return rc;
    *error = std::sqrt(ss) / std::max(std::fabs(lam), 1e-300);
  }
  LSQED_HIP(hipStreamSynchronize(c.s));
  return LSQED_OK;
}

int lsqed_stream_probe(int kind, double* a, const double* b, const double* c, int64_t n, double* part,
                       lsqed_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n <= 0 || !b) return fail(LSQED_ERR_ARG, "empty stream");
  if (kind == 0) {
    if (!a || !c) return fail(LSQED_ERR_ARG, "triad needs a, b and c");
    hipLaunchKernelGGL(k_stream_triad, dim3(1024), dim3(kThreads), 0, s, a, b, c, n, 0.5);
  } else if (kind == 1) {
    if (!part || (n & 1)) return fail(LSQED_ERR_ARG, "read stream needs an even n and 2048 partials");
    hipLaunchKernelGGL(k_stream_read, dim3(2048), dim3(kThreads), 0, s, b, n / 2, part);
  } else if (kind == 2 || kind == 3) {
    if (!a || n % 1000003 == 0 || n >= ((int64_t)1 << 34))
      return fail(LSQED_ERR_ARG, "gather probe needs a, n < 2^34 and n not a multiple of 1000003");
    hipLaunchKernelGGL(k_stream_gather, dim3(1024), dim3(kThreads), 0, s, a, b, n, kind == 2 ? 4 : 3);
  } else {
    return fail(LSQED_ERR_ARG, "kind must be 0 (triad), 1 (read), 2 or 3 (gathers)");
  }
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

}  // extern "C"

namespace lsqed {

// One three-term Lanczos step (lsqed_lanczos_step) with its partial sums in
// `scratch` (G + R + 8 doubles: G = hv_launch_grid(h), R = red_blocks(D)).
static int lanczos_step_core(lsqed_ham* h, const double* q, double sq, double* p, double sp, double beta, double* y,
                             double* alpha, double* beta_next, double* scratch, int G, hipStream_t s) {
  const int64_t D = h->D;
  const int R = red_blocks(D);
  int rc = LSQED_OK;
  double* dpart = scratch;
  double* upart = scratch + G;
  double* outv = scratch + G + R;
  if ((rc = launch_hv(h, q, y, dpart, s))) return rc;
  double qy = 0.0;
  if ((rc = finish_reduce(h, dpart, G, 1, outv, &qy, s))) return rc;
  const double a = sq * sq * qy;
  const bool aligned = ((reinterpret_cast<uintptr_t>(y) | reinterpret_cast<uintptr_t>(q) |
                         reinterpret_cast<uintptr_t>(p)) & 15) == 0;
  if (aligned)
    hipLaunchKernelGGL((nt_streams(D) ? k_lanczos_update<true> : k_lanczos_update<false>), dim3(R), dim3(kThreads), 0,
                       s, y, q, p, D, sq, -a * sq, -beta * sp, upart);
  else
    hipLaunchKernelGGL(k_lanczos_update_unaligned, dim3(R), dim3(kThreads), 0, s, y, q, p, D, sq, -a * sq,
                       -beta * sp, upart);
  LSQED_HIP(hipGetLastError());
  double ss = 0.0;
  if ((rc = finish_reduce(h, upart, R, 1, outv + 1, &ss, s))) return rc;
  if (alpha) *alpha = a;
  if (beta_next) *beta_next = std::sqrt(ss);
  return LSQED_OK;
}

// Device scratch of lsqed_lanczos_step_noalloc: 2^22 doubles (32 MiB) per
// device, held until the process ends.  Host code never passes `scratch`
// itself to a kernel (on the host the name is a stand-in, not a device
// address): noalloc_scratch() resolves the device copy once per device with
// hipGetSymbolAddress and every use goes through that pointer.
constexpr size_t kNoAllocDoubles = (size_t)1 << 22;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static double* g_noalloc_ptr[kMaxDevices] = {};

static int noalloc_scratch(double** out) {
  int dev = 0;
  LSQED_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(LSQED_ERR_ARG, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* ptr = nullptr;
    LSQED_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(LSQED_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_noalloc_ptr[dev] = (double*)ptr;
  }
  *out = g_noalloc_ptr[dev];
  return LSQED_OK;
}

}  // namespace lsqed

extern "C" {

int lsqed_lanczos_step(lsqed_ham* h, const double* q, double sq, double* p, double sp, double beta, double* y,
                       double* alpha, double* beta_next, lsqed_stream_t stream) {
  if (!h || !q || !p || !y) return fail(LSQED_ERR_ARG, "null argument");
  const int G = hv_launch_grid(h);
  if (G <= 0) return fail(LSQED_ERR_HIP, "H.v plan setup failed");
  int rc = ensure_scratch(h, (size_t)G + red_blocks(h->D) + 8);
  if (rc) return rc;
  return lanczos_step_core(h, q, sq, p, sp, beta, y, alpha, beta_next, h->d_scratch, G, (hipStream_t)stream);
}

int lsqed_lanczos_step_noalloc(lsqed_ham* h, const double* q, double sq, double* p, double sp, double beta,
                               double* y, double* alpha, double* beta_next, lsqed_stream_t stream) {
  if (!h || !q || !p || !y) return fail(LSQED_ERR_ARG, "null argument");
  if (!hv_plan_ready(h))
    return fail(LSQED_ERR_STATE,
                "the H.v plan of this model is not built: call lsqed_hv or lsqed_lanczos_step on it once first "
                "(and again after lsqed_ham_set_fields / set_couplings / set_build / set_hv_kernel)");
  const int G = hv_launch_grid(h);
  if (G <= 0) return fail(LSQED_ERR_HIP, "H.v plan setup failed");
  const size_t need = (size_t)G + red_blocks(h->D) + 8;
  if (need > kNoAllocDoubles)
    return fail(LSQED_ERR_ARG, "lsqed_lanczos_step_noalloc: this sector needs " + std::to_string(need) +
                                   " doubles of scratch, more than the " + std::to_string(kNoAllocDoubles) +
                                   " of the static array; use lsqed_lanczos_step");
  double* sbuf = nullptr;
  int rc = noalloc_scratch(&sbuf);
  if (rc) return rc;
  return lanczos_step_core(h, q, sq, p, sp, beta, y, alpha, beta_next, sbuf, G, (hipStream_t)stream);
}

int lsqed_lanczos_ground(lsqed_ham* h, int maxit, double tol, uint64_t seed, double* work, double* e0, int* steps,
                         lsqed_stream_t stream) {
  if (!h || !work || !e0) return fail(LSQED_ERR_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const int64_t D = h->D;
  double* q = work;
  double* p = work + D;
  double* y = work + 2 * D;
  hipLaunchKernelGGL(k_fill_random, dim3(stream_grid(D)), dim3(kThreads), 0, s, q, D, seed ? seed : 1);
  LSQED_HIP(hipGetLastError());
  LSQED_HIP(hipMemsetAsync(p, 0, D * sizeof(double), s));
  // ||q||
  const int R = red_blocks(D);
  int rc = ensure_scratch(h, (size_t)std::max(R, 2048) + 8);
  if (rc) return rc;
  hipLaunchKernelGGL(k_dot, dim3(R), dim3(kThreads), 0, s, q, q, D, h->d_scratch);
  LSQED_HIP(hipGetLastError());
  double ss = 0.0;
  if ((rc = finish_reduce(h, h->d_scratch, R, 1, h->d_scratch + 2048, &ss, s))) return rc;
  double sq = 1.0 / std::sqrt(ss), sp = 0.0, beta = 0.0;
  std::vector<double> al, be;
  std::vector<double> hist;
  if (maxit <= 0) maxit = 1000;
  if (tol <= 0) tol = 1e-12;
  int it = 0;
  double ev = 0.0;
  for (it = 0; it < maxit; ++it) {
    double a, bn;
    if ((rc = lsqed_lanczos_step(h, q, sq, p, sp, beta, y, &a, &bn, stream))) return rc;
    al.push_back(a);
    ev = tridiag_lowest(al, be);
    hist.push_back(ev);
    const int n = (int)hist.size();
    if (n > 5 && std::fabs(hist[n - 1] - hist[n - 6]) <= tol * std::max(1.0, std::fabs(ev))) {
      ++it;
      break;
    }
    if (!(bn > 1e-14 * std::fabs(a)) || (int64_t)al.size() >= D) {
      ++it;
      break;
    }
    be.push_back(bn);
    // new q = p / bn (scale factor), new p = old q
    std::swap(q, p);
    sp = sq;
    sq = 1.0 / bn;
    beta = bn;
  }
  *e0 = ev;
  if (steps) *steps = it;
  return LSQED_OK;
}

}  // extern "C"
