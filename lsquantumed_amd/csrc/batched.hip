// batched.hip -- many disorder realisations (and chains of uneven length)
// solved together by a batched three-term Lanczos for the ground state.
//
// The reference solves realisations one at a time (examples/main_disorder.cpp
// :82-120: build_disorder, SolverInit(nev=1), solve, solution per
// realisation).  Here the batch's vectors are concatenated per role (q, p, y,
// psi: sum_k D_k doubles each); one batched H.v walks a task list of
// (problem, row range) pairs, each problem with its own lattice, sector
// tables, couplings and fields; the update, dots and norms are batched the
// same way, and every per-problem scalar of the recurrence lives in device
// memory, so a Lanczos step is a fixed sequence of six launches with no host
// round trip.  kSteps steps are captured in one hipGraph (the roles of the
// two Krylov vectors swap every step; kSteps is even so a replay ends where
// it began) and replayed; each replay also copies its alpha/beta history to
// pinned host memory.  The host processes replay r while replay r+1 runs:
// it extends each problem's tridiagonal matrix, tests the lowest Ritz pair's
// residual |beta_m s_m| / |theta| < tol, and retires converged problems
// through a device flag array that every kernel reads (retirement lands one
// replay later; those extra steps only sharpen the tridiagonal matrix).
// Ground-state vectors come from a second pass that repeats the same steps
// bit for bit (same start vectors, same launch shape, deterministic
// reductions) and accumulates psi = sum_j s_j v_j with the Ritz vector s of
// the tridiagonal matrix at the step where the problem first converged.
// The launch shape (rows per task, lanes per state, blocks per CU) is chosen
// per batch shape by timing candidates on the batch itself (autotune) and
// cached for the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lsqed_internal.hpp"
#include "replay_schedule.hpp"
#include "sector_device.hpp"

namespace lsqed {

constexpr int kBThreads = 256;
// Lanczos steps per graph replay (even).  Shorter replays let a realisation
// stop nearer the step where it converges; longer ones cost fewer host
// round trips.  256-chain disorder workload, steps per replay 2 / 4 / 6 / 8 /
// 12: 1084-1090 / 1093-1095 / 1086 / 1074-1077 / 1058 realisations/s
// (mean Lanczos steps 82.1 / 83.0 / 84.2 / 85.0 / 87.0).
#ifndef LSQED_BATCH_STEPS
#define LSQED_BATCH_STEPS 4
#endif
constexpr int kSteps = LSQED_BATCH_STEPS;
static_assert(kSteps >= 2 && kSteps % 2 == 0, "a replay must end where it started (q and p roles)");
constexpr int kObsStride = LSQED_BATCH_OBS_STRIDE;

// One problem of the batch (device-resident descriptor).
struct BatchProb {
  RowParams p;     // sector tables, bonds, fields, couplings; p.D = this problem's dimension
  int64_t off;     // offset of this problem's entries in the batched vectors
  int task0, task1;  // its task range
  int hv0, hv1;      // its range of H.v partial sums (the task range, or the specialised kernel's)
};

// Task t covers rows [row0, row0 + nrows) of problem prob.
struct BTask {
  int prob;
  int row0;
  int nrows;
};

// Tasks are dealt cyclically: logical block lb takes t = lb, lb + G, ...;
// with the XCD-aware logical index each XCD takes a contiguous stretch of
// G/8 tasks per round, and all blocks advance through the task list
// together, so the vectors being gathered at any time are those of a few
// neighbouring problems.
__device__ __forceinline__ int first_task() { return (int)xcd_block(); }

template <int LPS>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = LPS / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <int LPS>
__device__ __forceinline__ int group_isum(int v) {
#pragma unroll
  for (int o = LPS / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Stage problem k's rank tables, binomials and fields in LDS.
__device__ __forceinline__ void stage_tables(const RowParams& p, uint32_t* s_tab, uint32_t* s_binom, double* s_f) {
  const uint32_t* gt = static_cast<const uint32_t*>(p.table);
  const uint32_t* gb = static_cast<const uint32_t*>(p.binom);
  for (int i = threadIdx.x; i < p.table_entries; i += blockDim.x) s_tab[i] = gt[i];
  for (int i = threadIdx.x; i < p.binom_entries; i += blockDim.x) s_binom[i] = gb[i];
  if (p.diag_mode == 1)
    for (int i = threadIdx.x; i < p.N; i += blockDim.x) s_f[i] = p.fields[i];
}

__device__ __forceinline__ size_t fields_offset(const RowParams& p) {
  return ((size_t)(p.table_entries + p.binom_entries) * 4 + 15) & ~(size_t)15;
}

// Batched H.v: y_k = H_k q_k for every active problem k, LPS lanes per
// basis state (lane j of a group takes bonds j, j+LPS, ...); dotpart[t] =
// the task's share of q_k . y_k.
template <int LPS>
__global__ __launch_bounds__(kBThreads) void k_batch_hv(const BatchProb* __restrict__ probs,
                                                        const BTask* __restrict__ tasks, int ntasks,
                                                        const int* __restrict__ active,
                                                        const double* __restrict__ q, double* __restrict__ y,
                                                        double* __restrict__ dotpart) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[kBThreads / 64];
  uint32_t* s_tab = reinterpret_cast<uint32_t*>(smem);
  constexpr int RPB = kBThreads / LPS;
  const int sub = threadIdx.x % LPS, grp = threadIdx.x / LPS;
  int loaded = -1;
  for (int t = first_task(); t < ntasks; t += gridDim.x) {
    const BTask task = tasks[t];
    const int k = task.prob;
    if (!active[k]) continue;  // block-uniform
    const RowParams& p = probs[k].p;
    uint32_t* s_binom = s_tab + p.table_entries;
    double* s_f = reinterpret_cast<double*>(smem + fields_offset(p));
    if (k != loaded) {
      __syncthreads();
      stage_tables(p, s_tab, s_binom, s_f);
      __syncthreads();
      loaded = k;
    }
    const double* qk = q + probs[k].off;
    double* yk = y + probs[k].off;
    const int rend = task.row0 + task.nrows;
    double acc_dot = 0.0;
    for (int r0 = task.row0; r0 < rend; r0 += RPB) {
      const int row = r0 + grp;
      if (row < rend) {  // uniform within each group of LPS lanes
        const uint32_t s = unrank_row<uint32_t, uint32_t>((uint32_t)row, p.N, p.nup, s_binom);
        int kc[4];
        uint32_t tc[4];
        chunk_terms<uint32_t, uint32_t, 8, 4>(s, kc, tc, s_tab, p);
        double acc = 0.0;
        int anti_n[2] = {0, 0};
#pragma unroll
        for (int kind = 0; kind < 2; ++kind) {
          const uint32_t* bk = p.bonds + (kind ? p.nb[0] : 0);
          const int nb = p.nb[kind];
          const bool fl = p.flip[kind] != 0;
          const double jh = p.jh[kind];
          for (int b = sub; b < nb; b += 4 * LPS) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int bb = b + u * LPS;
              v[u] = 0.0;
              if (bb < nb) {
                const uint32_t bd = bk[bb];
                const int lo = bd & 0xff, hi = (bd >> 8) & 0xff;
                const uint32_t m = (1u << lo) | (1u << hi);
                const uint32_t sm = s & m;
                const bool anti = (sm != 0) && (sm != m);
                anti_n[kind] += anti;
                if (fl && anti)
                  v[u] = qk[flip_rank<uint32_t, uint32_t, 8, 4>((uint32_t)row, s, lo, hi, kc, tc, s_tab, p)];
              }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += jh * v[u];
          }
        }
        if (LPS > 1) {
          acc = group_sum<LPS>(acc);
          anti_n[0] = group_isum<LPS>(anti_n[0]);
          anti_n[1] = group_isum<LPS>(anti_n[1]);
        }
        if (sub == 0) {
          const double d = diag_value(p, p.nb[0] - 2 * anti_n[0], p.nb[1] - 2 * anti_n[1], (uint64_t)s, s_f);
          const double xr = qk[row];
          const double yr = d * xr + acc;
          yk[row] = yr;
          acc_dot += xr * yr;
        }
      }
    }
    const double tot = block_sum(acc_dot, red);
    if (threadIdx.x == 0) dotpart[t] = tot;
  }
}

// out[k] = sum of the partials of problem k, over its task range (hv = 0)
// or its H.v partial range (hv = 1), in a fixed order (deterministic).
__global__ void k_batch_reduce(const BatchProb* __restrict__ probs, int nprob, const double* __restrict__ part,
                               double* __restrict__ out, int hv) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  if (k >= nprob) return;
  const int t0 = hv ? probs[k].hv0 : probs[k].task0, t1 = hv ? probs[k].hv1 : probs[k].task1;
  double acc = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) acc += part[t];
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) out[k] = tot;
}

// The per-problem sum of k_batch_reduce fused with the scalar step that
// consumes it (one launch where there were two).  After the H.v (AB = 0):
// alpha_k = sq^2 (q.y) and the coefficients of p <- sq*y - alpha*sq*q -
// beta*sp*p; after the update (AB = 1): beta' = ||p|| and the scales of the
// next step, sp <- sq, sq <- 1/beta'.
template <int AB>
__global__ void k_batch_reduce_ab(const BatchProb* __restrict__ probs, int nprob, const double* __restrict__ part,
                                  double* __restrict__ out, const int* __restrict__ active, double* __restrict__ sq,
                                  double* __restrict__ sp, double* __restrict__ beta, double* __restrict__ hist,
                                  double* __restrict__ coef) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  if (k >= nprob) return;
  const int t0 = AB == 0 ? probs[k].hv0 : probs[k].task0, t1 = AB == 0 ? probs[k].hv1 : probs[k].task1;
  double acc = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) acc += part[t];
  const double tot = block_sum(acc, red);
  if (threadIdx.x != 0) return;
  out[k] = tot;
  if (AB == 0) {
    const double s = sq[k];
    const double a = active[k] ? s * s * tot : 0.0;
    hist[k] = a;
    coef[3 * k + 0] = s;
    coef[3 * k + 1] = -a * s;
    coef[3 * k + 2] = -beta[k] * sp[k];
  } else {
    if (!active[k]) {
      hist[k] = 0.0;
      return;
    }
    const double b = sqrt(tot);
    hist[k] = b;
    sp[k] = sq[k];
    sq[k] = b > 0.0 ? 1.0 / b : 0.0;
    beta[k] = b;
  }
}

// p <- a_k y + b_k q + c_k p ; normpart[t] = the task's share of ||p||^2.
__global__ __launch_bounds__(kBThreads) void k_batch_update(const BatchProb* __restrict__ probs,
                                                            const BTask* __restrict__ tasks, int ntasks,
                                                            const int* __restrict__ active,
                                                            const double* __restrict__ coef,
                                                            const double* __restrict__ y,
                                                            const double* __restrict__ q, double* __restrict__ p,
                                                            double* __restrict__ normpart) {
  __shared__ double red[kBThreads / 64];
  for (int t = first_task(); t < ntasks; t += gridDim.x) {
    const BTask task = tasks[t];
    const int k = task.prob;
    if (!active[k]) continue;
    const int64_t off = probs[k].off;
    const double a = coef[3 * k], b = coef[3 * k + 1], c = coef[3 * k + 2];
    double acc = 0.0;
    // four rows per lane per iteration, their 12 loads issued before any is used
    const int end = task.row0 + task.nrows;
    for (int r0 = task.row0; r0 < end; r0 += 4 * kBThreads) {
      double vy[4], vq[4], vp[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = r0 + u * kBThreads + threadIdx.x;
        if (r < end) {
          vy[u] = y[off + r];
          vq[u] = q[off + r];
          vp[u] = p[off + r];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = r0 + u * kBThreads + threadIdx.x;
        if (r < end) {
          const double v = a * vy[u] + b * vq[u] + c * vp[u];
          p[off + r] = v;
          acc += v * v;
        }
      }
    }
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) normpart[t] = tot;
  }
}

// The same update on pairs of rows (16-byte loads and stores), for batches
// whose problem offsets and task lengths are all even: four pairs per lane
// per iteration, their 12 loads issued before any is used; the task's share of
// ||p||^2 sums the pairs' squares in lane order.
// NT: nontemporal loads and stores, for batches whose vectors exceed the
// 256 MB memory-side cache (2^25 doubles); LSQED_BATCH_NT=0/1 overrides.
// 256-chain disorder workload (842 MB per vector): all-active step 2.559 ->
// 2.474 ms, solve 1009 -> 1042 realisations/s.
template <bool NT>
__device__ __forceinline__ double2 bld2(const double2* p) {
  if constexpr (NT) {
    typedef double v2 __attribute__((ext_vector_type(2)));
    const v2 t = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return make_double2(t.x, t.y);
  } else {
    return *p;
  }
}
template <bool NT>
__device__ __forceinline__ void bst2(double2* p, double2 v) {
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

template <bool NT>
__global__ __launch_bounds__(kBThreads) void k_batch_update2(const BatchProb* __restrict__ probs,
                                                             const BTask* __restrict__ tasks, int ntasks,
                                                             const int* __restrict__ active,
                                                             const double* __restrict__ coef,
                                                             const double* __restrict__ y,
                                                             const double* __restrict__ q, double* __restrict__ p,
                                                             double* __restrict__ normpart) {
  __shared__ double red[kBThreads / 64];
  for (int t = first_task(); t < ntasks; t += gridDim.x) {
    const BTask task = tasks[t];
    const int k = task.prob;
    if (!active[k]) continue;
    const int64_t off = probs[k].off;
    const double a = coef[3 * k], b = coef[3 * k + 1], c = coef[3 * k + 2];
    const double2* __restrict__ y2 = (const double2*)(y + off);
    const double2* __restrict__ q2 = (const double2*)(q + off);
    double2* __restrict__ p2 = (double2*)(p + off);
    double acc = 0.0;
    const int end = (task.row0 + task.nrows) >> 1;
    for (int i0 = task.row0 >> 1; i0 < end; i0 += 4 * kBThreads) {
      double2 vy[4], vq[4], vp[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + u * kBThreads + threadIdx.x;
        if (i < end) {
          vy[u] = bld2<NT>(y2 + i);
          vq[u] = bld2<NT>(q2 + i);
          vp[u] = bld2<NT>(p2 + i);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + u * kBThreads + threadIdx.x;
        if (i < end) {
          double2 v;
          v.x = a * vy[u].x + b * vq[u].x + c * vp[u].x;
          v.y = a * vy[u].y + b * vq[u].y + c * vp[u].y;
          bst2<NT>(p2 + i, v);
          acc += v.x * v.x;
          acc += v.y * v.y;
        }
      }
    }
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) normpart[t] = tot;
  }
}

// psi_k += w_k * sq_k * q_k (second pass: the Ritz vector, one Lanczos
// vector per step; w_k = 0 once the problem's Ritz vector is complete).
__global__ __launch_bounds__(kBThreads) void k_batch_accum(const BatchProb* __restrict__ probs,
                                                           const BTask* __restrict__ tasks, int ntasks,
                                                           const int* __restrict__ active,
                                                           const double* __restrict__ w,
                                                           const double* __restrict__ sq,
                                                           const double* __restrict__ q, double* __restrict__ psi) {
  for (int t = first_task(); t < ntasks; t += gridDim.x) {
    const BTask task = tasks[t];
    const int k = task.prob;
    if (!active[k] || w[k] == 0.0) continue;
    const int64_t off = probs[k].off;
    const double c = w[k] * sq[k];
    for (int r = task.row0 + threadIdx.x; r < task.row0 + task.nrows; r += kBThreads)
      psi[off + r] += c * q[off + r];
  }
}

// Seeded start vectors, uniform in [-1, 1): splitmix64 of (seed_k, row).
__global__ void k_batch_fill_random(const BatchProb* __restrict__ probs, const BTask* __restrict__ tasks,
                                    int ntasks, uint64_t seed, double* __restrict__ v) {
  for (int t = first_task(); t < ntasks; t += gridDim.x) {
    const BTask task = tasks[t];
    const uint64_t sk = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(task.prob + 1);
    const int64_t off = probs[task.prob].off;
    for (int r = task.row0 + threadIdx.x; r < task.row0 + task.nrows; r += blockDim.x) {
      uint64_t z = sk + 0xD1B54A32D192ED03ull * (uint64_t)r;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      v[off + r] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
  }
}

// Zero-fill over the task list (no memset nodes in code a graph may hold).
__global__ void k_batch_zero(const BatchProb* __restrict__ probs, const BTask* __restrict__ tasks, int ntasks,
                             double* __restrict__ v) {
  for (int t = first_task(); t < ntasks; t += gridDim.x) {
    const BTask task = tasks[t];
    const int64_t off = probs[task.prob].off;
    for (int r = task.row0 + threadIdx.x; r < task.row0 + task.nrows; r += blockDim.x) v[off + r] = 0.0;
  }
}

// part[t] = the task's share of ||v_k||^2
__global__ void k_batch_sumsq(const BatchProb* __restrict__ probs, const BTask* __restrict__ tasks, int ntasks,
                              const double* __restrict__ v, double* __restrict__ part) {
  __shared__ double red[kBThreads / 64];
  for (int t = first_task(); t < ntasks; t += gridDim.x) {
    const BTask task = tasks[t];
    const int64_t off = probs[task.prob].off;
    double acc = 0.0;
    for (int r = task.row0 + threadIdx.x; r < task.row0 + task.nrows; r += blockDim.x) {
      const double x = v[off + r];
      acc += x * x;
    }
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) part[t] = tot;
  }
}

// v_k *= 1/sqrt(nn_k)
__global__ void k_batch_normalise(const BatchProb* __restrict__ probs, const BTask* __restrict__ tasks, int ntasks,
                                  const double* __restrict__ nn, double* __restrict__ v) {
  for (int t = first_task(); t < ntasks; t += gridDim.x) {
    const BTask task = tasks[t];
    const int64_t off = probs[task.prob].off;
    const double c = nn[task.prob] > 0.0 ? 1.0 / sqrt(nn[task.prob]) : 0.0;
    for (int r = task.row0 + threadIdx.x; r < task.row0 + task.nrows; r += blockDim.x) v[off + r] *= c;
  }
}

// Observables of the batch's ground states (the backward pass of E0):
// per task, partial sums of
//   [i < N]   <S^z_i>            = sum_r psi_r^2 (bit_i ? +1/2 : -1/2)
//   [32]      <psi|psi>
//   [33 + 2k] sum_{bonds of kind k} <S^z_i S^z_j>
//   [34 + 2k] sum_{bonds of kind k} <(S^+_i S^-_j + S^-_i S^+_j)/2>
//                                   = sum_r psi_r sum_{antiparallel b} psi_{flip_b(r)} / 2
// (SzSz/SpSm::ExpectVal, src/spinOperators/correlation.cpp:86-224, summed
// over the Hamiltonian's bonds).  Lane-private sums, then one wave sum and
// an LDS f64 atomic per value and wave.
__global__ __launch_bounds__(kBThreads) void k_batch_obs(const BatchProb* __restrict__ probs,
                                                         const BTask* __restrict__ tasks, int ntasks,
                                                         const double* __restrict__ psi, double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double s_acc[kObsStride];
  uint32_t* s_tab = reinterpret_cast<uint32_t*>(smem);
  int loaded = -1;
  for (int i = threadIdx.x; i < kObsStride; i += blockDim.x) s_acc[i] = 0.0;
  for (int t = first_task(); t < ntasks; t += gridDim.x) {
    const BTask task = tasks[t];
    const int k = task.prob;
    const RowParams& p = probs[k].p;
    uint32_t* s_binom = s_tab + p.table_entries;
    double* s_f = reinterpret_cast<double*>(smem + fields_offset(p));
    if (k != loaded) {
      __syncthreads();
      stage_tables(p, s_tab, s_binom, s_f);
      loaded = k;
    }
    __syncthreads();
    const double* x = psi + probs[k].off;
    double sz[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) sz[i] = 0.0;
    double nrm = 0.0, zz[2] = {0.0, 0.0}, fl[2] = {0.0, 0.0};
    for (int row = task.row0 + threadIdx.x; row < task.row0 + task.nrows; row += kBThreads) {
      const uint32_t s = unrank_row<uint32_t, uint32_t>((uint32_t)row, p.N, p.nup, s_binom);
      int kc[4];
      uint32_t tc[4];
      chunk_terms<uint32_t, uint32_t, 8, 4>(s, kc, tc, s_tab, p);
      const double xr = x[row];
      const double w = xr * xr;
      nrm += w;
#pragma unroll
      for (int i = 0; i < 32; ++i)
        if (i < p.N) sz[i] += ((s >> i) & 1u) ? 0.5 * w : -0.5 * w;
#pragma unroll
      for (int kind = 0; kind < 2; ++kind) {
        const uint32_t* bk = p.bonds + (kind ? p.nb[0] : 0);
        double g = 0.0;
        int anti_n = 0;
        for (int b = 0; b < p.nb[kind]; ++b) {
          const uint32_t bd = bk[b];
          const int lo = bd & 0xff, hi = (bd >> 8) & 0xff;
          const uint32_t m = (1u << lo) | (1u << hi);
          const uint32_t sm = s & m;
          if ((sm != 0) && (sm != m)) {
            ++anti_n;
            g += x[flip_rank<uint32_t, uint32_t, 8, 4>((uint32_t)row, s, lo, hi, kc, tc, s_tab, p)];
          }
        }
        zz[kind] += 0.25 * (double)(p.nb[kind] - 2 * anti_n) * w;
        fl[kind] += 0.5 * xr * g;
      }
    }
    const bool lead = (threadIdx.x & 63) == 0;
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      if (i < p.N) {
        const double v = wave_sum(sz[i]);
        if (lead) atomicAdd(&s_acc[i], v);
      }
    }
    {
      const double v0 = wave_sum(nrm), v1 = wave_sum(zz[0]), v2 = wave_sum(fl[0]), v3 = wave_sum(zz[1]),
                   v4 = wave_sum(fl[1]);
      if (lead) {
        atomicAdd(&s_acc[32], v0);
        atomicAdd(&s_acc[33], v1);
        atomicAdd(&s_acc[34], v2);
        atomicAdd(&s_acc[35], v3);
        atomicAdd(&s_acc[36], v4);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kObsStride; i += blockDim.x) {
      part[(size_t)t * kObsStride + i] = s_acc[i];
      s_acc[i] = 0.0;
    }
  }
}

// out[k*S + j] = sum over problem k's tasks of part[t*S + j]
__global__ void k_batch_reduce_obs(const BatchProb* __restrict__ probs, int nprob, const double* __restrict__ part,
                                   double* __restrict__ out) {
  const int k = blockIdx.x;
  const int j = threadIdx.x;
  if (k >= nprob || j >= kObsStride) return;
  double acc = 0.0;
  for (int t = probs[k].task0; t < probs[k].task1; ++t) acc += part[(size_t)t * kObsStride + j];
  out[(size_t)k * kObsStride + j] = acc;
}

}  // namespace lsqed

using namespace lsqed;

namespace {

enum { S_QY = 0, S_NN, S_SQ, S_SP, S_BETA, S_COEF, S_W = S_COEF + 3, S_COUNT = S_W + kSteps };

struct Shape {
  int kernel = 0;   // H.v: 0 chunk-table kernel over the task list, 1 specialised row split per lattice group
  int rows = 4096;  // rows per task
  int lps = 1;      // lanes per basis state (kernel 0)
  int bpc = 4;      // blocks per CU
  float step_ms = 0.0f;  // one Lanczos step of the whole batch (all problems active), as timed by the tuner
  bool operator!=(const Shape& o) const { return kernel != o.kernel || rows != o.rows || lps != o.lps || bpc != o.bpc; }
};

// Problems that share a bond structure and build mode (one specialised kernel).
struct Group {
  void* module = nullptr;
  lsqed_ham* h0 = nullptr;  // its split tables serve the whole group
  std::vector<int> idx;     // batch indices
  int ntasks = 0;           // row-split tasks per problem
  int part0 = 0;            // first H.v partial of the group
  int gofs = 0;             // offset of its indices in d_gidx
  // when its problems are one contiguous range: problems [p0, p0 + np) and
  // their update tasks [t0, t0 + nt) (p0 = -1 otherwise)
  int p0 = -1, np = 0, t0 = 0, nt = 0;
};

std::mutex g_tune_mu;
std::map<std::string, Shape> g_tune_cache;

}  // namespace

struct lsqed_batch {
  std::vector<lsqed_ham*> hams;
  lsqed_batch_opts opts{};
  int nprob = 0;
  int64_t total = 0;
  std::vector<int64_t> off, dims;
  Shape shape;
  int grid = 0, ntasks = 0;
  bool pairs = false;  // every problem offset and task length even: the update moves double2
  bool group_chains = false;  // replays run one Lanczos chain per lattice group, on its own stream
  size_t smem = 0;
  int cus = 256;
  // device
  BatchProb* d_probs = nullptr;
  BTask* d_tasks = nullptr;
  int* d_active = nullptr;
  double *d_q = nullptr, *d_p = nullptr, *d_y = nullptr, *d_psi = nullptr;
  double* d_part = nullptr;  // ntasks * kObsStride doubles
  size_t part_cap = 0;
  double* d_s = nullptr;     // S_COUNT * nprob scalars
  double* d_hist = nullptr;  // [2][kSteps][nprob]: alpha then beta of each step of a replay
  double* d_obs = nullptr;   // nprob * kObsStride
  // specialised H.v per lattice group
  std::vector<Group> groups;
  bool jit_ok = false;
  int64_t* d_poff = nullptr;
  double* d_pcoup = nullptr;   // 4 per problem
  double* d_pfields = nullptr; // 32 per problem
  int* d_gidx = nullptr;
  double* d_hvpart = nullptr;
  size_t hvpart_cap = 0;
  std::vector<int> prob_hv0;
  bool use_graph = true;
  // pinned host
  double* h_hist = nullptr;  // [2 buffers][2][kSteps][nprob]
  int* h_active = nullptr;   // ring of activity vectors, [ring][nprob]
  int ring = 0;
  hipStream_t gs = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // side streams (default priority) for the H.v of the smaller lattice
  // groups, forked from and joined back to the launching stream
  static constexpr int kSide = 3;
  hipStream_t side[kSide] = {nullptr, nullptr, nullptr};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[kSide] = {nullptr, nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};  // [0] plain steps, [1] steps + Ritz accumulation
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  // results
  std::vector<double> e0, resid;
  std::vector<int> steps, steps_vec, conv;
  std::vector<std::vector<double>> al, be;
  bool solved = false, have_vectors = false;
};

namespace {

double* sc(lsqed_batch* b, int which) { return b->d_s + (size_t)which * b->nprob; }
double* hist_alpha(lsqed_batch* b, int j) { return b->d_hist + (size_t)j * b->nprob; }
double* hist_beta(lsqed_batch* b, int j) { return b->d_hist + (size_t)(kSteps + j) * b->nprob; }

std::string shape_key(const lsqed_batch* b) {
  std::string key = "k" + std::to_string(b->opts.hv_kernel) + "l" + std::to_string(b->opts.lanes_per_state) +
                    (b->jit_ok ? "j;" : ";");
  for (const lsqed_ham* h : b->hams)
    key += std::to_string(h->N) + "/" + std::to_string(h->bi[0].size()) + "/" + std::to_string(h->bi[1].size()) +
           "/" + std::to_string(h->build) + ";";
  return key;
}

// Lattice groups of the specialised H.v: problems whose generated kernel
// source (bond structure, build mode) is the same share one module.  Rebuilt
// at every solve, since build flags decide the kernel.  When some problem
// cannot use it (N > 32, too many high bonds, hipRTC unavailable) the batch
// keeps the chunk-table kernel only.
int setup_groups(lsqed_batch* b) {
  const int np = b->nprob;
  b->groups.clear();
  b->jit_ok = false;
  b->prob_hv0.assign(np, 0);
  if (!jit_enabled() || b->opts.hv_kernel == 1) return LSQED_OK;
  std::map<std::string, int> gid;
  std::vector<int> pl(np, 0);
  for (int k = 0; k < np; ++k) {
    lsqed_ham* h = b->hams[k];
    if (!split_supported(h) || !split_jit_fits(h, true)) return LSQED_OK;
    void* mod = nullptr;
    std::string key;
    if (batch_jit_prepare(h, &mod, &key) != LSQED_OK) return LSQED_OK;
    auto it = gid.find(key);
    if (it == gid.end()) {
      it = gid.emplace(key, (int)b->groups.size()).first;
      Group g;
      g.module = mod;
      g.h0 = h;
      b->groups.push_back(g);
    }
    Group& g = b->groups[it->second];
    pl[k] = (int)g.idx.size();
    g.idx.push_back(k);
  }
  int part = 0, gofs = 0;
  std::vector<int> gidx;
  for (Group& g : b->groups) {
    const int rc = batch_task_count(g.module, g.h0, &g.ntasks);
    if (rc) return rc;
    g.part0 = part;
    g.gofs = gofs;
    for (int k : g.idx) b->prob_hv0[k] = part + pl[k] * g.ntasks;
    part += g.ntasks * (int)g.idx.size();
    gofs += (int)g.idx.size();
    gidx.insert(gidx.end(), g.idx.begin(), g.idx.end());
  }
  if ((size_t)part > b->hvpart_cap) {
    (void)hipFree(b->d_hvpart);
    b->d_hvpart = nullptr;
    LSQED_HIP(hipMalloc((void**)&b->d_hvpart, (size_t)part * sizeof(double)));
    b->hvpart_cap = part;
  }
  std::vector<double> coup(4 * (size_t)np), fields(32 * (size_t)np, 0.0);
  for (int k = 0; k < np; ++k) {
    const lsqed_ham* h = b->hams[k];
    const RowParams rp = make_row_params(h);
    coup[4 * k + 0] = rp.jh[0];
    coup[4 * k + 1] = rp.jh[1];
    coup[4 * k + 2] = rp.dq[0];
    coup[4 * k + 3] = rp.dq[1];
    if (rp.diag_mode == 1)
      for (int i = 0; i < h->N; ++i) fields[32 * k + i] = h->fields[i];
  }
  LSQED_HIP(hipMemcpy(b->d_gidx, gidx.data(), np * sizeof(int), hipMemcpyHostToDevice));
  LSQED_HIP(hipMemcpy(b->d_poff, b->off.data(), np * sizeof(int64_t), hipMemcpyHostToDevice));
  LSQED_HIP(hipMemcpy(b->d_pcoup, coup.data(), coup.size() * sizeof(double), hipMemcpyHostToDevice));
  LSQED_HIP(hipMemcpy(b->d_pfields, fields.data(), fields.size() * sizeof(double), hipMemcpyHostToDevice));
  b->jit_ok = true;
  return LSQED_OK;
}

int grid_for(const lsqed_batch* b, int64_t work_items) {
  int64_t g = (int64_t)b->cus * (int64_t)b->shape.bpc;
  g = std::min<int64_t>(g, std::max<int64_t>(work_items, 1));
  return (int)std::max<int64_t>(8, ((g + 7) / 8) * 8);
}

// Upload problem descriptors (fields pointers and build modes may have
// changed since the last solve) and the task list for rows-per-task.
int upload_problems(lsqed_batch* b) {
  std::vector<BTask> tasks;
  std::vector<BatchProb> probs(b->nprob);
  const int rows = b->shape.rows;
  const bool spec = b->shape.kernel == 1;
  for (int k = 0; k < b->nprob; ++k) {
    lsqed_ham* h = b->hams[k];
    BatchProb& bp = probs[k];
    std::memset(&bp, 0, sizeof(bp));
    bp.p = make_row_params(h);
    bp.off = b->off[k];
    bp.task0 = (int)tasks.size();
    for (int64_t r = 0; r < h->D; r += rows) tasks.push_back(BTask{k, (int)r, (int)std::min<int64_t>(rows, h->D - r)});
    bp.task1 = (int)tasks.size();
    if (spec) {
      const Group* g = nullptr;
      for (const Group& gg : b->groups)
        if (std::find(gg.idx.begin(), gg.idx.end(), k) != gg.idx.end()) g = &gg;
      bp.hv0 = b->prob_hv0[k];
      bp.hv1 = bp.hv0 + (g ? g->ntasks : 0);
    } else {
      bp.hv0 = bp.task0;
      bp.hv1 = bp.task1;
    }
  }
  if ((int)tasks.size() != b->ntasks) {
    (void)hipFree(b->d_tasks);
    b->d_tasks = nullptr;
    LSQED_HIP(hipMalloc((void**)&b->d_tasks, tasks.size() * sizeof(BTask)));
  }
  b->ntasks = (int)tasks.size();
  const size_t need = (size_t)b->ntasks * kObsStride;
  if (need > b->part_cap) {
    (void)hipFree(b->d_part);
    b->d_part = nullptr;
    LSQED_HIP(hipMalloc((void**)&b->d_part, need * sizeof(double)));
    b->part_cap = need;
  }
  LSQED_HIP(hipMemcpy(b->d_tasks, tasks.data(), tasks.size() * sizeof(BTask), hipMemcpyHostToDevice));
  LSQED_HIP(hipMemcpy(b->d_probs, probs.data(), probs.size() * sizeof(BatchProb), hipMemcpyHostToDevice));
  b->grid = grid_for(b, b->ntasks);
  // Per-group Lanczos chains need each group's problems (hence its tasks) to
  // be one contiguous range and a stream per group (LSQED_BATCH_GROUP_CHAINS=0
  // keeps the lockstep replay).
  bool contig = spec && b->groups.size() >= 2 && b->groups.size() <= (size_t)lsqed_batch::kSide + 1 && b->side[0];
  for (Group& g : b->groups) {
    g.p0 = -1;
    bool c = !g.idx.empty();
    for (size_t i = 0; c && i < g.idx.size(); ++i) c = g.idx[i] == g.idx[0] + (int)i;
    if (c) {
      g.p0 = g.idx[0];
      g.np = (int)g.idx.size();
      g.t0 = probs[g.p0].task0;
      g.nt = probs[g.p0 + g.np - 1].task1 - g.t0;
    }
    contig = contig && c;
  }
  const char* gc = std::getenv("LSQED_BATCH_GROUP_CHAINS");
  b->group_chains = contig && !(gc && std::atoi(gc) == 0);
  const char* pe = std::getenv("LSQED_BATCH_PAIRS");  // 0: the one-row update
  b->pairs = !(pe && std::atoi(pe) == 0);
  for (const BTask& t : tasks) b->pairs = b->pairs && (t.row0 % 2 == 0) && (t.nrows % 2 == 0);
  for (int k = 0; k < b->nprob; ++k) b->pairs = b->pairs && (b->off[k] % 2 == 0);
  return LSQED_OK;
}

// y = H q for every active problem; the per-problem partials of q.y go to
// the array returned by hv_partials().
int launch_hv_batch(lsqed_batch* b, const double* q, double* y, hipStream_t s) {
  if (b->shape.kernel == 1) {
    int64_t work = 0;
    for (const Group& g : b->groups) work += (int64_t)g.ntasks * (int64_t)g.idx.size();
    const int grid = grid_for(b, work);
    auto launch = [&](const Group& g, hipStream_t st) -> int {
      BatchJitLaunch L{q,           y,           b->d_hvpart + g.part0, b->d_poff, b->d_pcoup, b->d_pfields,
                       b->d_active, b->d_gidx + g.gofs, (int)g.idx.size()};
      return launch_hv_batch_jit(g.module, g.h0, L, grid, st);
    };
    const int ng = (int)b->groups.size();
    if (ng < 2 || !b->side[0]) {
      for (const Group& g : b->groups) {
        const int rc = launch(g, s);
        if (rc) return rc;
      }
      return LSQED_OK;
    }
    // The groups write disjoint rows of y and of the partials: the smaller
    // groups run on side streams beside the largest (a small group's launch
    // alone leaves most CUs idle), all joined back before the step goes on.
    // Inside a graph capture the fork and join become graph edges.
    int big = 0;
    for (int gi = 1; gi < ng; ++gi)
      if ((int64_t)b->groups[gi].ntasks * (int64_t)b->groups[gi].idx.size() >
          (int64_t)b->groups[big].ntasks * (int64_t)b->groups[big].idx.size())
        big = gi;
    LSQED_HIP(hipEventRecord(b->fork_ev, s));
    int used = 0, k = 0;
    for (int gi = 0; gi < ng; ++gi) {
      if (gi == big) continue;
      hipStream_t st = b->side[k % lsqed_batch::kSide];
      if (k < lsqed_batch::kSide) {
        LSQED_HIP(hipStreamWaitEvent(st, b->fork_ev, 0));
        used = k + 1;
      }
      const int rc = launch(b->groups[gi], st);
      if (rc) return rc;
      ++k;
    }
    const int rc = launch(b->groups[big], s);
    if (rc) return rc;
    for (int i = 0; i < used; ++i) {
      LSQED_HIP(hipEventRecord(b->join_ev[i], b->side[i]));
      LSQED_HIP(hipStreamWaitEvent(s, b->join_ev[i], 0));
    }
    return LSQED_OK;
  }
  switch (b->shape.lps) {
    case 2:
      hipLaunchKernelGGL(k_batch_hv<2>, dim3(b->grid), dim3(kBThreads), b->smem, s, b->d_probs, b->d_tasks,
                         b->ntasks, b->d_active, q, y, b->d_part);
      break;
    case 4:
      hipLaunchKernelGGL(k_batch_hv<4>, dim3(b->grid), dim3(kBThreads), b->smem, s, b->d_probs, b->d_tasks,
                         b->ntasks, b->d_active, q, y, b->d_part);
      break;
    default:
      hipLaunchKernelGGL(k_batch_hv<1>, dim3(b->grid), dim3(kBThreads), b->smem, s, b->d_probs, b->d_tasks,
                         b->ntasks, b->d_active, q, y, b->d_part);
  }
  return LSQED_OK;
}

double* hv_partials(lsqed_batch* b) { return b->shape.kernel == 1 ? b->d_hvpart : b->d_part; }

// Nontemporal update streams for batches whose vectors exceed the 256 MB
// memory-side cache (LSQED_BATCH_NT=0/1 overrides).
bool batch_nt(const lsqed_batch* b) {
  static const int nt_env = [] {
    const char* e = std::getenv("LSQED_BATCH_NT");
    return e ? (std::atoi(e) == 1 ? 1 : 0) : -1;
  }();
  return nt_env >= 0 ? nt_env == 1 : b->total >= ((int64_t)1 << 25);
}

// One Lanczos step on the current vector q (scale sq) and previous p; the
// new vector overwrites p.  Scalars of the step go to history slot j.
int enqueue_step(lsqed_batch* b, double* q, double* p, int j, bool accumulate, hipStream_t s) {
  const int np = b->nprob;
  if (j < 0 || j >= kSteps) return fail(LSQED_ERR_STATE, "batched Lanczos: history slot out of range");
  const int rc = launch_hv_batch(b, q, b->d_y, s);
  if (rc) return rc;
  if (accumulate)
    hipLaunchKernelGGL(k_batch_accum, dim3(b->grid), dim3(kBThreads), 0, s, b->d_probs, b->d_tasks, b->ntasks,
                       b->d_active, sc(b, S_W + j), sc(b, S_SQ), q, b->d_psi);
  hipLaunchKernelGGL(k_batch_reduce_ab<0>, dim3(np), dim3(256), 0, s, b->d_probs, np, hv_partials(b), sc(b, S_QY),
                     b->d_active, sc(b, S_SQ), sc(b, S_SP), sc(b, S_BETA), hist_alpha(b, j), sc(b, S_COEF));
  const bool aligned = (((uintptr_t)q | (uintptr_t)p | (uintptr_t)b->d_y) & 15u) == 0;
  const bool nt = batch_nt(b);
  hipLaunchKernelGGL(b->pairs && aligned ? (nt ? k_batch_update2<true> : k_batch_update2<false>) : k_batch_update,
                     dim3(b->grid), dim3(kBThreads), 0, s,
                     b->d_probs, b->d_tasks, b->ntasks, b->d_active, sc(b, S_COEF), b->d_y, q, p, b->d_part);
  hipLaunchKernelGGL(k_batch_reduce_ab<1>, dim3(np), dim3(256), 0, s, b->d_probs, np, b->d_part, sc(b, S_NN),
                     b->d_active, sc(b, S_SQ), sc(b, S_SP), sc(b, S_BETA), hist_beta(b, j), nullptr);
  return LSQED_OK;
}

// Lanczos start: q = seeded random, p = 0, sq = 1/||q||, all problems active.
int start(lsqed_batch* b, hipStream_t s) {
  const int np = b->nprob;
  hipLaunchKernelGGL(k_batch_fill_random, dim3(b->grid), dim3(kBThreads), 0, s, b->d_probs, b->d_tasks, b->ntasks,
                     (uint64_t)b->opts.seed, b->d_q);
  hipLaunchKernelGGL(k_batch_zero, dim3(b->grid), dim3(kBThreads), 0, s, b->d_probs, b->d_tasks, b->ntasks, b->d_p);
  hipLaunchKernelGGL(k_batch_sumsq, dim3(b->grid), dim3(kBThreads), 0, s, b->d_probs, b->d_tasks, b->ntasks, b->d_q,
                     b->d_part);
  hipLaunchKernelGGL(k_batch_reduce, dim3(np), dim3(256), 0, s, b->d_probs, np, b->d_part, sc(b, S_NN), 0);
  LSQED_HIP(hipGetLastError());
  std::vector<double> nn(np);
  LSQED_HIP(hipMemcpyAsync(nn.data(), sc(b, S_NN), np * sizeof(double), hipMemcpyDeviceToHost, s));
  LSQED_HIP(hipStreamSynchronize(s));
  std::vector<double> init((size_t)S_COUNT * np, 0.0);
  for (int k = 0; k < np; ++k) {
    if (!(nn[k] > 0.0)) return fail(LSQED_ERR_STATE, "zero start vector");
    init[(size_t)S_SQ * np + k] = 1.0 / std::sqrt(nn[k]);
  }
  std::vector<int> act(np, 1);
  LSQED_HIP(hipMemcpy(b->d_s, init.data(), init.size() * sizeof(double), hipMemcpyHostToDevice));
  LSQED_HIP(hipMemcpy(b->d_active, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice));
  return LSQED_OK;
}

void drop_graphs(lsqed_batch* b) {
  for (int i = 0; i < 2; ++i) {
    if (b->gexec[i]) (void)hipGraphExecDestroy(b->gexec[i]);
    if (b->graph[i]) (void)hipGraphDestroy(b->graph[i]);
    b->gexec[i] = nullptr;
    b->graph[i] = nullptr;
  }
}

// One Lanczos step of lattice group g alone, on stream st: the same kernels
// as enqueue_step over the group's problems [p0, p0 + np) and tasks
// [t0, t0 + nt) (problem indices inside the kernels stay absolute).
int enqueue_group_step(lsqed_batch* b, const Group& g, double* q, double* p, int j, bool accumulate,
                       hipStream_t st) {
  if (j < 0 || j >= kSteps) return fail(LSQED_ERR_STATE, "batched Lanczos: history slot out of range");
  int64_t work = 0;
  for (const Group& gg : b->groups) work += (int64_t)gg.ntasks * (int64_t)gg.idx.size();
  BatchJitLaunch L{q,           b->d_y,      b->d_hvpart + g.part0, b->d_poff, b->d_pcoup, b->d_pfields,
                   b->d_active, b->d_gidx + g.gofs, (int)g.idx.size()};
  const int rc = launch_hv_batch_jit(g.module, g.h0, L, grid_for(b, work), st);
  if (rc) return rc;
  const int p0 = g.p0, np = g.np;
  const int grid = grid_for(b, g.nt);
  if (accumulate)
    hipLaunchKernelGGL(k_batch_accum, dim3(grid), dim3(kBThreads), 0, st, b->d_probs, b->d_tasks + g.t0, g.nt,
                       b->d_active, sc(b, S_W + j), sc(b, S_SQ), q, b->d_psi);
  hipLaunchKernelGGL(k_batch_reduce_ab<0>, dim3(np), dim3(256), 0, st, b->d_probs + p0, np, b->d_hvpart,
                     sc(b, S_QY) + p0, b->d_active + p0, sc(b, S_SQ) + p0, sc(b, S_SP) + p0, sc(b, S_BETA) + p0,
                     hist_alpha(b, j) + p0, sc(b, S_COEF) + 3 * (size_t)p0);
  const bool aligned = (((uintptr_t)q | (uintptr_t)p | (uintptr_t)b->d_y) & 15u) == 0;
  if (b->pairs && aligned) {
    if (batch_nt(b))
      hipLaunchKernelGGL(k_batch_update2<true>, dim3(grid), dim3(kBThreads), 0, st, b->d_probs, b->d_tasks + g.t0,
                         g.nt, b->d_active, sc(b, S_COEF), b->d_y, q, p, b->d_part + g.t0);
    else
      hipLaunchKernelGGL(k_batch_update2<false>, dim3(grid), dim3(kBThreads), 0, st, b->d_probs, b->d_tasks + g.t0,
                         g.nt, b->d_active, sc(b, S_COEF), b->d_y, q, p, b->d_part + g.t0);
  } else {
    hipLaunchKernelGGL(k_batch_update, dim3(grid), dim3(kBThreads), 0, st, b->d_probs, b->d_tasks + g.t0, g.nt,
                       b->d_active, sc(b, S_COEF), b->d_y, q, p, b->d_part + g.t0);
  }
  hipLaunchKernelGGL(k_batch_reduce_ab<1>, dim3(np), dim3(256), 0, st, b->d_probs + p0, np, b->d_part,
                     sc(b, S_NN) + p0, b->d_active + p0, sc(b, S_SQ) + p0, sc(b, S_SP) + p0, sc(b, S_BETA) + p0,
                     hist_beta(b, j) + p0, nullptr);
  return LSQED_OK;
}

// kSteps Lanczos steps (q,p roles alternating), the unit of a replay.  With
// group chains every lattice group runs its own kSteps steps on its own
// stream (the largest on s, the others on the side streams, forked from s
// and joined back at the end): the groups' problems are independent, so a
// small group's update and its next steps run while the largest group's H.v
// still holds most of the GPU, instead of waiting for it at every step.
int enqueue_replay_steps(lsqed_batch* b, bool accumulate, hipStream_t s) {
  if (b->group_chains) {
    const int ng = (int)b->groups.size();
    int big = 0;
    for (int gi = 1; gi < ng; ++gi)
      if (b->groups[gi].nt > b->groups[big].nt) big = gi;
    std::vector<hipStream_t> st(ng, s);
    LSQED_HIP(hipEventRecord(b->fork_ev, s));
    int used = 0;
    for (int gi = 0; gi < ng; ++gi) {
      if (gi == big) continue;
      st[gi] = b->side[used++];
      LSQED_HIP(hipStreamWaitEvent(st[gi], b->fork_ev, 0));
    }
    for (int j = 0; j < kSteps; ++j) {
      double* q = (j & 1) ? b->d_p : b->d_q;
      double* p = (j & 1) ? b->d_q : b->d_p;
      for (int gi = 0; gi < ng; ++gi) {
        const int rc = enqueue_group_step(b, b->groups[gi], q, p, j, accumulate, st[gi]);
        if (rc) return rc;
      }
    }
    for (int i = 0; i < used; ++i) {
      LSQED_HIP(hipEventRecord(b->join_ev[i], b->side[i]));
      LSQED_HIP(hipStreamWaitEvent(s, b->join_ev[i], 0));
    }
    return LSQED_OK;
  }
  for (int j = 0; j < kSteps; ++j) {
    double* q = (j & 1) ? b->d_p : b->d_q;
    double* p = (j & 1) ? b->d_q : b->d_p;
    const int rc = enqueue_step(b, q, p, j, accumulate, s);
    if (rc) return rc;
  }
  return LSQED_OK;
}

// Capture a replay as graph `which` (0 plain, 1 with Ritz accumulation).
// If the runtime cannot capture it, replays are enqueued eagerly instead.
int make_graph(lsqed_batch* b, int which) {
  if (b->gexec[which] || !b->use_graph) return LSQED_OK;
  LSQED_HIP(hipStreamBeginCapture(b->gs, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_replay_steps(b, which == 1, b->gs);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->gs, &g);
  if (rc || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    b->use_graph = false;
    return LSQED_OK;
  }
  b->graph[which] = g;
  if (hipGraphInstantiate(&b->gexec[which], g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    drop_graphs(b);
    b->use_graph = false;
  }
  return LSQED_OK;
}

int run_replay(lsqed_batch* b, int which) {
  if (b->use_graph && b->gexec[which]) {
    LSQED_HIP(hipGraphLaunch(b->gexec[which], b->gs));
    return LSQED_OK;
  }
  const int rc = enqueue_replay_steps(b, which == 1, b->gs);
  if (rc) return rc;
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

// Lowest eigenvalue of the symmetric tridiagonal matrix (al, be[0..n-2])
// by Sturm-sequence bisection; its eigenvector by inverse iteration with a
// shift just below it ((T - shift) is then positive definite, so the
// pivot-free LDL^T solve is stable).
double tri_lowest(const double* al, const double* be, int n, std::vector<double>* vec) {
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(be[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(be[i]) : 0.0);
    lo = std::min(lo, al[i] - r);
    hi = std::max(hi, al[i] + r);
  }
  auto count_below = [&](double x) {
    int cnt = 0;
    double d = 1.0;
    for (int i = 0; i < n; ++i) {
      d = al[i] - x - (i > 0 ? be[i - 1] * be[i - 1] / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0) ++cnt;
    }
    return cnt;
  };
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count_below(mid) >= 1) hi = mid; else lo = mid;
  }
  const double ev = 0.5 * (lo + hi);
  if (vec) {
    std::vector<double>& z = *vec;
    z.assign(n, 1.0);
    std::vector<double> d(n), l(n), w(n);
    const double shift = ev - 1e-10 * std::max(1.0, std::fabs(ev));
    d[0] = al[0] - shift;
    for (int i = 1; i < n; ++i) {
      l[i] = be[i - 1] / d[i - 1];
      d[i] = al[i] - shift - l[i] * be[i - 1];
    }
    for (int pass = 0; pass < 3; ++pass) {
      w[0] = z[0];
      for (int i = 1; i < n; ++i) w[i] = z[i] - l[i] * w[i - 1];
      z[n - 1] = w[n - 1] / d[n - 1];
      for (int i = n - 2; i >= 0; --i) z[i] = (w[i] - be[i] * z[i + 1]) / d[i];
      double nrm = 0.0;
      for (double v : z) nrm += v * v;
      nrm = std::sqrt(nrm);
      for (double& v : z) v /= nrm;
    }
  }
  return ev;
}

// Time a few Lanczos steps (ms per step) with the current shape.
int time_shape(lsqed_batch* b, float* ms) {
  int rc = upload_problems(b);
  if (rc) return rc;
  if ((rc = start(b, b->gs))) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto timed = [&]() -> int {
    LSQED_HIP(hipEventCreate(&e0));
    LSQED_HIP(hipEventCreate(&e1));
    int r = enqueue_step(b, b->d_q, b->d_p, 0, false, b->gs);
    if (r) return r;
    LSQED_HIP(hipEventRecord(e0, b->gs));
    // the history holds kSteps slots (d_hist: 2 kSteps rows of nprob),
    // which the timed steps reuse cyclically
    for (int j = 1; j <= 4; ++j)
      if ((r = enqueue_step(b, (j & 1) ? b->d_p : b->d_q, (j & 1) ? b->d_q : b->d_p, j % kSteps, false, b->gs)))
        return r;
    LSQED_HIP(hipEventRecord(e1, b->gs));
    LSQED_HIP(hipGetLastError());
    LSQED_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    LSQED_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / 4.0f;
    return LSQED_OK;
  };
  rc = timed();
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

// Autotune, timed on the batch itself and cached per batch shape:
//   chunk-table kernel: rows per task x blocks per CU at one lane per state,
//   then 2 and 4 lanes per state at the best tile (or the fixed lane count);
//   specialised row split: blocks per CU, then rows per task (which set the
//   update's tasks there).
int autotune(lsqed_batch* b) {
  const std::string key = shape_key(b);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      b->shape = it->second;
      return LSQED_OK;
    }
  }
  const int fixed_lps = b->opts.lanes_per_state;
  const bool try_generic = b->opts.hv_kernel != 2;
  const bool try_spec = b->opts.hv_kernel != 1 && fixed_lps == 0 && b->jit_ok;
  if (!try_generic && !try_spec)
    return fail(LSQED_ERR_ARG, "the specialised batched H.v is unavailable for this batch (N > 32, too many "
                               "high bonds, or hipRTC disabled)");
  Shape best;
  float best_ms = 1e30f;
  auto trial = [&](Shape sh) -> int {
    b->shape = sh;
    float ms = 0.0f;
    const int rc = time_shape(b, &ms);
    if (rc) return rc;
    if (ms < best_ms) {
      best_ms = ms;
      best = sh;
      best.step_ms = ms;
    }
    return LSQED_OK;
  };
  int rc;
  int rows_best = 4096;
  if (try_generic) {
    const int l0 = fixed_lps > 0 ? fixed_lps : 1;
    for (int rows : {1024, 4096, 16384})
      for (int bpc : {2, 4, 8})
        if ((rc = trial(Shape{0, rows, l0, bpc}))) return rc;
    rows_best = best.rows;
    if (fixed_lps == 0) {
      const Shape base = best;
      for (int lps : {2, 4})
        if ((rc = trial(Shape{0, base.rows, lps, base.bpc}))) return rc;
    }
  }
  if (try_spec) {
    for (int bpc : {2, 4, 8, 32, 1 << 20})  // 1 << 20: one work item per block
      if ((rc = trial(Shape{1, rows_best, 1, bpc}))) return rc;
    // rows per task still set the update's, the partial sums' and the start
    // vector's tasks under the specialised H.v (whose own tasks are rows)
    if (best.kernel == 1) {
      const Shape sb = best;
      for (int rows : {2048, 4096, 8192})
        if (rows != sb.rows && (rc = trial(Shape{1, rows, 1, sb.bpc}))) return rc;
    }
  }
  b->shape = best;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = best;
  return LSQED_OK;
}

int prepare(lsqed_batch* b, hipStream_t caller) {
  // order after the caller's pending work
  hipEvent_t e;
  LSQED_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  LSQED_HIP(hipEventRecord(e, caller));
  LSQED_HIP(hipStreamWaitEvent(b->gs, e, 0));
  (void)hipEventDestroy(e);
  int rc = setup_groups(b);
  if (rc) return rc;
  if ((rc = autotune(b))) return rc;
  // graphs hold the launch arguments (task list, groups, grid): rebuild them
  // every solve, since build flags and groups may have changed
  drop_graphs(b);
  b->use_graph = true;
  return upload_problems(b);  // also restores the chosen shape's task list and grid
}

}  // namespace

extern "C" {

int64_t lsqed_batch_workspace_bytes(int n, lsqed_ham* const* hams, const lsqed_batch_opts* o) {
  if (n <= 0 || !hams) return -1;
  int64_t total = 0;
  for (int k = 0; k < n; ++k) {
    if (!hams[k]) return -1;
    total += hams[k]->D;
  }
  const int nvec = (o && o->keep_vectors) ? 4 : 3;
  return total * 8 * nvec;
}

int lsqed_batch_create(int n, lsqed_ham* const* hams, const lsqed_batch_opts* o, lsqed_batch** out) {
  if (!out) return fail(LSQED_ERR_ARG, "null output handle");
  *out = nullptr;
  if (n <= 0 || !hams) return fail(LSQED_ERR_ARG, "empty batch");
  for (int k = 0; k < n; ++k) {
    const lsqed_ham* h = hams[k];
    if (!h) return fail(LSQED_ERR_ARG, "null Hamiltonian in batch");
    if (h->tab.wide || h->N > 32) return fail(LSQED_ERR_ARG, "the batched solver needs N <= 32");
    if (h->D < 2) return fail(LSQED_ERR_ARG, "sector dimension must be at least 2");
    if (h->D >= ((int64_t)1 << 31)) return fail(LSQED_ERR_ARG, "problem too large for the batched solver");
  }
  if (o && (o->lanes_per_state != 0 && o->lanes_per_state != 1 && o->lanes_per_state != 2 &&
            o->lanes_per_state != 4))
    return fail(LSQED_ERR_ARG, "lanes_per_state must be 0 (autotune), 1, 2 or 4");
  if (o && (o->hv_kernel < 0 || o->hv_kernel > 2))
    return fail(LSQED_ERR_ARG, "hv_kernel must be 0 (autotune), 1 (chunk table) or 2 (specialised row split)");
  if (o && o->hv_kernel == 2 && o->lanes_per_state > 1)
    return fail(LSQED_ERR_ARG, "lanes_per_state applies to the chunk-table kernel only");
  lsqed_batch* b = new lsqed_batch;
  if (o) b->opts = *o;
  if (b->opts.maxit <= 0) b->opts.maxit = 300;
  if (b->opts.tol <= 0) b->opts.tol = 1e-9;
  if (!b->opts.seed) b->opts.seed = 1;
  b->nprob = n;
  size_t smem = 0;
  for (int k = 0; k < n; ++k) {
    lsqed_ham* h = hams[k];
    b->hams.push_back(h);
    b->off.push_back(b->total);
    b->dims.push_back(h->D);
    b->total += h->D;
    const size_t sm = ((h->tab.entries.size() + h->tab.binom.size()) * 4 + 15) / 16 * 16 + 32 * 8;
    smem = std::max(smem, sm);
  }
  b->smem = smem;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      prop.multiProcessorCount > 0)
    b->cus = prop.multiProcessorCount;
  const int nvec = b->opts.keep_vectors ? 4 : 3;
  bool ok = hipMalloc((void**)&b->d_q, b->total * 8) == hipSuccess &&
            hipMalloc((void**)&b->d_p, b->total * 8) == hipSuccess &&
            hipMalloc((void**)&b->d_y, b->total * 8) == hipSuccess &&
            (nvec == 3 || hipMalloc((void**)&b->d_psi, b->total * 8) == hipSuccess) &&
            hipMalloc((void**)&b->d_probs, n * sizeof(BatchProb)) == hipSuccess &&
            hipMalloc((void**)&b->d_active, n * sizeof(int)) == hipSuccess &&
            hipMalloc((void**)&b->d_s, (size_t)S_COUNT * n * sizeof(double)) == hipSuccess &&
            hipMalloc((void**)&b->d_hist, (size_t)2 * kSteps * n * sizeof(double)) == hipSuccess &&
            hipMalloc((void**)&b->d_obs, (size_t)kObsStride * n * sizeof(double)) == hipSuccess &&
            hipMalloc((void**)&b->d_poff, n * sizeof(int64_t)) == hipSuccess &&
            hipMalloc((void**)&b->d_pcoup, (size_t)4 * n * sizeof(double)) == hipSuccess &&
            hipMalloc((void**)&b->d_pfields, (size_t)32 * n * sizeof(double)) == hipSuccess &&
            hipMalloc((void**)&b->d_gidx, n * sizeof(int)) == hipSuccess;
  b->ring = b->opts.maxit / kSteps + 4;
  ok = ok && hipHostMalloc((void**)&b->h_hist, (size_t)2 * 2 * kSteps * n * sizeof(double), 0) == hipSuccess &&
       hipHostMalloc((void**)&b->h_active, (size_t)b->ring * n * sizeof(int), 0) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    lsqed_batch_destroy(b);
    return fail(LSQED_ERR_OOM, "cannot allocate the batched Lanczos workspace");
  }
  if (hipStreamCreateWithFlags(&b->gs, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev[1], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) != hipSuccess) {
    lsqed_batch_destroy(b);
    return fail(LSQED_ERR_HIP, "cannot create the batch stream");
  }
  for (int i = 0; i < lsqed_batch::kSide; ++i)
    if (hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->join_ev[i], hipEventDisableTiming) != hipSuccess) {
      lsqed_batch_destroy(b);
      return fail(LSQED_ERR_HIP, "cannot create the batch side streams");
    }
  *out = b;
  return LSQED_OK;
}

int lsqed_batch_destroy(lsqed_batch* b) {
  if (!b) return LSQED_OK;
  if (b->gs) (void)hipStreamSynchronize(b->gs);
  for (hipStream_t st : b->side)
    if (st) (void)hipStreamSynchronize(st);
  drop_graphs(b);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  for (hipEvent_t e : b->join_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t st : b->side)
    if (st) (void)hipStreamDestroy(st);
  if (b->gs) (void)hipStreamDestroy(b->gs);
  for (void* ptr : {(void*)b->d_q, (void*)b->d_p, (void*)b->d_y, (void*)b->d_psi, (void*)b->d_part, (void*)b->d_s,
                    (void*)b->d_hist, (void*)b->d_obs, (void*)b->d_tasks, (void*)b->d_probs, (void*)b->d_active,
                    (void*)b->d_poff, (void*)b->d_pcoup, (void*)b->d_pfields, (void*)b->d_gidx, (void*)b->d_hvpart})
    (void)hipFree(ptr);
  if (b->h_hist) (void)hipHostFree(b->h_hist);
  if (b->h_active) (void)hipHostFree(b->h_active);
  delete b;
  return LSQED_OK;
}

int lsqed_batch_solve(lsqed_batch* b, double* e0, lsqed_stream_t stream) {
  if (!b || !e0) return fail(LSQED_ERR_ARG, "null argument");
  const int np = b->nprob;
  int rc = prepare(b, (hipStream_t)stream);
  if (rc) return rc;
  if ((rc = make_graph(b, 0))) return rc;
  if ((rc = start(b, b->gs))) return rc;
  b->solved = b->have_vectors = false;
  b->al.assign(np, {});
  b->be.assign(np, {});
  b->e0.assign(np, 0.0);
  b->resid.assign(np, 0.0);
  b->steps.assign(np, 0);
  b->steps_vec.assign(np, 0);
  b->conv.assign(np, 0);
  // Which replays run which model is decided on the host (replay_schedule.hpp).
  ReplaySchedule sch(np);
  const size_t HS = (size_t)2 * kSteps * np;  // doubles per history buffer
  int slot = 0;                               // ring slot of the next activity upload
  std::vector<double> ritz;
  auto enqueue = [&]() -> int {
    const int r = (int)sch.act_at.size();
    if (sch.dirty) {
      int* act = b->h_active + (size_t)(slot % b->ring) * np;
      ++slot;
      for (int k = 0; k < np; ++k) act[k] = sch.cur[k];
      LSQED_HIP(hipMemcpyAsync(b->d_active, act, np * sizeof(int), hipMemcpyHostToDevice, b->gs));
    }
    sch.enqueued();
    const int rc2 = run_replay(b, 0);
    if (rc2) return rc2;
    LSQED_HIP(hipMemcpyAsync(b->h_hist + (r & 1) * HS, b->d_hist, HS * sizeof(double), hipMemcpyDeviceToHost, b->gs));
    LSQED_HIP(hipEventRecord(b->ev[r & 1], b->gs));
    return LSQED_OK;
  };
  if ((rc = enqueue())) return rc;
  while (!sch.finished()) {
    while (sch.want_replay())
      if ((rc = enqueue())) return rc;
    if (sch.stalled()) {
      (void)hipStreamSynchronize(b->gs);
      int npaused = 0;
      for (int k = 0; k < np; ++k) npaused += sch.paused[k];
      return fail(LSQED_ERR_STATE, "batched Lanczos: no replay in flight (" + std::to_string(sch.ndone) + " of " +
                                       std::to_string(np) + " done, " + std::to_string(npaused) + " paused, " +
                                       std::to_string(sch.done_r) + " replays processed)");
    }
    const int r = sch.done_r;
    LSQED_HIP(hipEventSynchronize(b->ev[r & 1]));
    const double* ha = b->h_hist + (r & 1) * HS;
    const double* hb = ha + (size_t)kSteps * np;
    for (int k = 0; k < np; ++k) {
      if (!sch.ran(k)) continue;
      auto& al = b->al[k];
      auto& be = b->be[k];
      bool broke = false;
      for (int j = 0; j < kSteps && !broke; ++j) {
        al.push_back(ha[(size_t)j * np + k]);
        be.push_back(hb[(size_t)j * np + k]);
        double scale = 0.0;
        for (size_t i = 0; i < al.size(); ++i) scale = std::max({scale, std::fabs(al[i]), std::fabs(be[i])});
        // invariant subspace: the Krylov space is exhausted at this step
        if (!(be.back() > 1e-13 * scale)) broke = true;
      }
      const int m = (int)al.size();
      const double theta = tri_lowest(al.data(), be.data(), m, &ritz);
      const double res = std::fabs(be.back() * ritz[m - 1]) / std::max(std::fabs(theta), 1e-300);
      b->e0[k] = theta;
      b->resid[k] = res;
      b->steps[k] = m;
      const bool conv = broke || res < b->opts.tol;
      const bool stop = conv || m + kSteps > b->opts.maxit;
      if (stop) {
        b->steps_vec[k] = m;
        b->conv[k] = conv ? 1 : 0;
      }
      sch.result(k, stop, res, b->opts.tol);
    }
    sch.processed();
  }
  LSQED_HIP(hipStreamSynchronize(b->gs));
  for (int k = 0; k < np; ++k) e0[k] = b->e0[k];
  b->solved = true;
  if (b->opts.keep_vectors) {
    // second pass: repeat the steps, psi += s_j v_j with the Ritz vector of
    // T at steps_vec[k]; problem k runs ceil(steps_vec/kSteps) replays
    if ((rc = make_graph(b, 1))) return rc;
    if ((rc = start(b, b->gs))) return rc;
    hipLaunchKernelGGL(k_batch_zero, dim3(b->grid), dim3(kBThreads), 0, b->gs, b->d_probs, b->d_tasks, b->ntasks,
                       b->d_psi);
    int nrep = 0;
    std::vector<std::vector<double>> s(np);
    for (int k = 0; k < np; ++k) {
      const int m = b->steps_vec[k];
      tri_lowest(b->al[k].data(), b->be[k].data(), m, &s[k]);
      nrep = std::max(nrep, (m + kSteps - 1) / kSteps);
    }
    // per replay: activity vector + kSteps weight rows (pinned, one buffer per replay)
    double* hw = nullptr;
    int* ha = nullptr;
    LSQED_HIP(hipHostMalloc((void**)&hw, (size_t)nrep * kSteps * np * sizeof(double), 0));
    if (hipHostMalloc((void**)&ha, (size_t)nrep * np * sizeof(int), 0) != hipSuccess) {
      (void)hipHostFree(hw);
      return fail(LSQED_ERR_OOM, "pinned host buffer");
    }
    for (int r = 0; r < nrep; ++r)
      for (int k = 0; k < np; ++k) {
        const int m = b->steps_vec[k];
        ha[(size_t)r * np + k] = r * kSteps < m;
        for (int j = 0; j < kSteps; ++j) {
          const int step = r * kSteps + j;
          hw[((size_t)r * kSteps + j) * np + k] = step < m ? s[k][step] : 0.0;
        }
      }
    for (int r = 0; r < nrep && !rc; ++r) {
      if (hipMemcpyAsync(b->d_active, ha + (size_t)r * np, np * sizeof(int), hipMemcpyHostToDevice, b->gs) ||
          hipMemcpyAsync(sc(b, S_W), hw + (size_t)r * kSteps * np, (size_t)kSteps * np * sizeof(double),
                         hipMemcpyHostToDevice, b->gs) ||
          run_replay(b, 1))
        rc = fail(LSQED_ERR_HIP, "second Lanczos pass launch failed");
    }
    if (!rc) {
      hipLaunchKernelGGL(k_batch_sumsq, dim3(b->grid), dim3(kBThreads), 0, b->gs, b->d_probs, b->d_tasks, b->ntasks,
                         b->d_psi, b->d_part);
      hipLaunchKernelGGL(k_batch_reduce, dim3(np), dim3(256), 0, b->gs, b->d_probs, np, b->d_part, sc(b, S_NN), 0);
      hipLaunchKernelGGL(k_batch_normalise, dim3(b->grid), dim3(kBThreads), 0, b->gs, b->d_probs, b->d_tasks,
                         b->ntasks, sc(b, S_NN), b->d_psi);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->gs) != hipSuccess)
        rc = fail(LSQED_ERR_HIP, "second Lanczos pass failed");
    } else {
      (void)hipStreamSynchronize(b->gs);
    }
    (void)hipHostFree(hw);
    (void)hipHostFree(ha);
    if (rc) return rc;
    b->have_vectors = true;
  }
  return LSQED_OK;
}

int lsqed_batch_info(const lsqed_batch* b, int k, int* steps, double* resid, int* converged) {
  if (!b) return fail(LSQED_ERR_ARG, "null handle");
  if (!b->solved) return fail(LSQED_ERR_STATE, "batch not solved");
  if (k < 0 || k >= b->nprob) return fail(LSQED_ERR_ARG, "problem index out of range");
  if (steps) *steps = b->steps[k];
  if (resid) *resid = b->resid[k];
  if (converged) *converged = b->conv[k];
  return LSQED_OK;
}

int lsqed_batch_schedule(const lsqed_batch* b, int* hv_kernel, int* rows_per_task, int* lanes_per_state,
                         int* blocks_per_cu, int* graph, double* step_ms) {
  if (!b) return fail(LSQED_ERR_ARG, "null handle");
  if (step_ms) *step_ms = b->shape.step_ms;
  if (hv_kernel) *hv_kernel = b->shape.kernel + 1;
  if (graph) *graph = b->use_graph && b->gexec[0] ? 1 : 0;
  if (rows_per_task) *rows_per_task = b->shape.rows;
  if (lanes_per_state) *lanes_per_state = b->shape.lps;
  if (blocks_per_cu) *blocks_per_cu = b->shape.bpc;
  return LSQED_OK;
}

int lsqed_batch_ground_state(lsqed_batch* b, int k, double* psi, lsqed_stream_t stream) {
  if (!b || !psi) return fail(LSQED_ERR_ARG, "null argument");
  if (!b->have_vectors) return fail(LSQED_ERR_STATE, "no ground-state vectors (keep_vectors, then solve)");
  if (k < 0 || k >= b->nprob) return fail(LSQED_ERR_ARG, "problem index out of range");
  LSQED_HIP(hipMemcpyAsync(psi, b->d_psi + b->off[k], b->dims[k] * sizeof(double), hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
  return LSQED_OK;
}

int lsqed_batch_observables(lsqed_batch* b, double* out, lsqed_stream_t stream) {
  if (!b || !out) return fail(LSQED_ERR_ARG, "null argument");
  if (!b->have_vectors) return fail(LSQED_ERR_STATE, "no ground-state vectors (keep_vectors, then solve)");
  hipStream_t s = (hipStream_t)stream;
  const int np = b->nprob;
  hipLaunchKernelGGL(k_batch_obs, dim3(b->grid), dim3(kBThreads), b->smem, s, b->d_probs, b->d_tasks, b->ntasks,
                     b->d_psi, b->d_part);
  hipLaunchKernelGGL(k_batch_reduce_obs, dim3(np), dim3(64), 0, s, b->d_probs, np, b->d_part, b->d_obs);
  LSQED_HIP(hipGetLastError());
  LSQED_HIP(hipMemcpyAsync(out, b->d_obs, (size_t)np * kObsStride * sizeof(double), hipMemcpyDeviceToHost, s));
  LSQED_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < np; ++k) {
    double* o = out + (size_t)k * kObsStride;
    const double nrm = o[32];
    if (!(nrm > 0.0)) return fail(LSQED_ERR_STATE, "zero ground-state vector");
    for (int j = 0; j < kObsStride; ++j)
      if (j != 32) o[j] /= nrm;
  }
  return LSQED_OK;
}

}  // extern "C"
