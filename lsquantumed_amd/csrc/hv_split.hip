// hv_split.hip -- row-split matrix-free H.v for N <= 32 (the fast path).
//
// The colex rank separates into a high and a low part.  Write a state as
// s = (h << L) | l with L <= 16 low bits (split_low_bits).  With kl = nup - popc(h),
//   r(s) = B(h) + RL(l),
//   B(h)  = sum_{j-th set bit of h at p} C(L + p, kl + j)   (row base)
//   RL(l) = sum_{j-th set bit of l at q} C(q, j)            (index inside the row)
// so all states sharing h form one contiguous "row" of C(L, kl) entries, in
// the order of the low patterns.  Bonds fall in three classes:
//   high  (both sites >= L): maps row h to row h ^ m entry by entry -- a
//         wave-uniform test and a contiguous, coalesced stream of x;
//   cross (one site on each side): maps row h to row h ^ (1<<p) in the
//         neighbouring popcount class, index RL(l ^ (1<<q)); the up-to-three
//         cross bonds of a high site share that target row, which stays in L2;
//   low   (both sites < L): stays inside row h (L2 resident).
// A work item is (row h, chunk of 2048 entries); each lane owns 8 entries
// of the chunk.  The row bases of the chunk's targets are computed once per
// item, so the per-entry work is table lookups on 16-bit patterns (LDS) and
// FP64 gathers/FMAs.  Same row semantics as hamiltonian.hip's generic kernel
// (the parity reference for H.v is the oracle's CSR).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lsqed_internal.hpp"

namespace lsqed {

constexpr int kSplitThreads = 256;
constexpr int kSplitPerLane = 8;
constexpr int kSplitChunk = kSplitThreads * kSplitPerLane;  // 2048 entries per work item
constexpr int kMaxClass = 64;                               // bonds per class

struct SplitParams {
  const double* fields;
  const uint16_t* U;      // low patterns per popcount class, colex order
  const uint16_t* RL0;    // 256: rank contribution of the low byte
  const uint16_t* RL1;    // 9 x W1: contribution of bits 8..L-1, by popcount of the low byte
  const uint32_t* binom;  // C(n, k), n <= N, k <= nup: [(n)*(nup+1) + k]
  const uint32_t* tasks;  // h | chunk << 24
  int N, nup, L, NH, ntasks, w1;  // w1 = 1 << max(L-8, 0)
  int nlow, ncross, nhigh;
  int offU[17];
  uint32_t low[kMaxClass];    // mask (16 bits) | kind << 16 | type << 17 (0: byte 0, 1: byte 1, 2: split)
  uint32_t cross[kMaxClass];  // q | p << 8 | kind << 16
  uint32_t high[kMaxClass];   // mask on h (NH <= 24 bits) | kind << 31
  double jh[2], dq[2];
  int flip[2], dz_on[2], nbk[2];
  int diag_mode;
};

struct SplitPlan {
  SplitParams p;
  uint16_t* d_U = nullptr;
  uint16_t* d_RL0 = nullptr;
  uint16_t* d_RL1 = nullptr;
  uint32_t* d_binom = nullptr;
  uint64_t* d_binom64 = nullptr;  // the same binomials in 64 bits (N > 32)
  bool classes_ok = false;        // the unspecialised kernel can run (N <= 32, classes fit)
  uint32_t* d_tasks = nullptr;  // (h, chunk) in h order: the batched kernel's list
  int ntasks = 0;
  uint32_t* d_tasks_ord = nullptr;  // the same items grouped for cache reuse, one segment per XCD
  int ntasks_ord = 0;               // a multiple of 8 (padding items process no entries)
  uint32_t* d_rows_ord = nullptr;   // whole rows, same order and padding (row-staged kernel)
  int nrows_ord = 0;
  uint64_t order_hash = 0;          // FNV-1a of the task and row orders (part of the H.v fingerprint)
  uint32_t* d_rows = nullptr;       // whole rows in h order, no padding (row-staged batched kernel)
  int nrows = 0;
  size_t binom_n = 0;
};

// -------------------------------------------------------------- the kernel --
template <bool DOT>
__global__ __launch_bounds__(kSplitThreads) void hv_split_kernel(SplitParams p, const double* __restrict__ x,
                                                                 double* __restrict__ y,
                                                                 double* __restrict__ dot_part) {
  __shared__ uint16_t s_RL0[256];
  __shared__ uint16_t s_RL1[9 * 256];
  __shared__ uint32_t s_binom[33 * 33];
  __shared__ double s_f[64];
  __shared__ uint32_t s_hbase[kMaxClass];
  __shared__ double s_hcoef[kMaxClass];
  __shared__ uint32_t s_cbase[24];  // NH = N - L <= 24 (N <= 32, L >= 8)
  __shared__ int s_nact;
  __shared__ uint32_t s_B;
  __shared__ int s_cnt[2];
  __shared__ double s_dish;
  __shared__ double s_red[8];

  const int tid = threadIdx.x;
  for (int i = tid; i < 256; i += kSplitThreads) s_RL0[i] = p.RL0[i];
  for (int i = tid; i < 9 * p.w1; i += kSplitThreads) s_RL1[i] = p.RL1[i];
  const int nb = (p.N + 1) * (p.nup + 1);
  for (int i = tid; i < nb; i += kSplitThreads) s_binom[i] = p.binom[i];
  if (p.diag_mode == 1)
    for (int i = tid; i < p.N; i += kSplitThreads) s_f[i] = p.fields[i];
  __syncthreads();

  const int K1 = p.nup + 1;
  auto C = [&](int n, int k) -> uint32_t { return (k < 0 || k > p.nup) ? 0u : s_binom[n * K1 + k]; };
  // row base of high pattern hh
  auto rowbase = [&](uint32_t hh) -> uint32_t {
    const int kl = p.nup - __builtin_popcount(hh);
    uint32_t b = 0;
    int j = 0;
    for (int q = 0; q < p.NH; ++q)
      if ((hh >> q) & 1u) {
        ++j;
        b += C(p.L + q, kl + j);
      }
    return b;
  };

  // XCD-aware contiguous task ranges
  const int64_t G = gridDim.x;
  const int64_t lb = (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
  const int64_t ta = lb * p.ntasks / G, tb = (lb + 1) * p.ntasks / G;
  double dotacc = 0.0;

  for (int64_t t = ta; t < tb; ++t) {
    const uint32_t task = p.tasks[t];
    const uint32_t h = task & 0xffffffu;
    const int chunk = (int)(task >> 24);
    const int kh = __builtin_popcount(h);
    const int kl = p.nup - kh;
    const int R = (int)C(p.L, kl);
    const int i0 = chunk * kSplitChunk;
    __syncthreads();  // previous item's LDS reads done
    if (tid < p.nhigh) {
      const uint32_t hd = p.high[tid];
      const uint32_t m = hd & 0xffffffu;
      const bool anti = __builtin_popcount(h & m) == 1;
      s_hbase[tid] = anti ? rowbase(h ^ m) : 0xffffffffu;
    } else if (tid >= 64 && tid < 64 + p.NH) {
      s_cbase[tid - 64] = rowbase(h ^ (1u << (tid - 64)));
    } else if (tid == 128) {
      s_B = rowbase(h);
      int c0 = 0, c1 = 0;
      for (int b = 0; b < p.nhigh; ++b) {
        const uint32_t hd = p.high[b];
        if (__builtin_popcount(h & (hd & 0xffffffu)) == 1) {
          if (hd >> 31) ++c1; else ++c0;
        }
      }
      s_cnt[0] = c0;
      s_cnt[1] = c1;
      double dh = 0.0;
      if (p.diag_mode == 1)
        for (int q = 0; q < p.NH; ++q) dh = ((h >> q) & 1u) ? dh + s_f[p.L + q] : dh - s_f[p.L + q];
      s_dish = dh;
    }
    __syncthreads();
    if (tid == 0) {
      int n = 0;
      for (int b = 0; b < p.nhigh; ++b) {
        const uint32_t base = s_hbase[b];
        const int kind = (int)(p.high[b] >> 31);
        if (base != 0xffffffffu && p.flip[kind]) {
          s_hbase[n] = base;
          s_hcoef[n] = p.jh[kind];
          ++n;
        }
      }
      s_nact = n;
    }
    __syncthreads();
    const uint32_t B = s_B;
    const int nact = s_nact;

    const uint16_t* Uk = p.U + p.offU[kl];
#pragma unroll 1
    for (int k = 0; k < kSplitPerLane; ++k) {
      const int i = i0 + tid + k * kSplitThreads;
      if (i >= R) break;
      // high bonds: the whole row h ^ m, entry i (coalesced across lanes)
      double a = 0.0;
      for (int b = 0; b < nact; ++b) {
        const uint32_t base = __builtin_amdgcn_readfirstlane(s_hbase[b]);
        a += s_hcoef[b] * x[base + (uint32_t)i];
      }
      // cross and low bonds
      const uint32_t l = Uk[i];
      const uint32_t lb0 = l & 255u;
      const int k0 = __builtin_popcount(lb0);
      const uint32_t t1 = s_RL1[k0 * p.w1 + (l >> 8)];
      const uint32_t t0 = (uint32_t)i - t1;
      int cc0 = 0, cc1 = 0;
#pragma unroll
      for (int b = 0; b < kMaxClass; ++b) {
        if (b < p.ncross) {
          const uint32_t cd = p.cross[b];
          const int q = cd & 255, hp = (cd >> 8) & 255, kind = (cd >> 16) & 1;
          const uint32_t bq = (l >> q) & 1u, bp = (h >> hp) & 1u;
          if (bq != bp) {
            if (kind) ++cc1; else ++cc0;
            if (p.flip[kind]) {
              const uint32_t l2 = l ^ (1u << q);
              uint32_t i2;
              if (q < 8) {
                const uint32_t v0 = l2 & 255u;
                i2 = s_RL0[v0] + s_RL1[__builtin_popcount(v0) * p.w1 + (l2 >> 8)];
              } else {
                i2 = t0 + s_RL1[k0 * p.w1 + (l2 >> 8)];
              }
              a += p.jh[kind] * x[s_cbase[hp] + i2];
            }
          }
        }
      }
#pragma unroll
      for (int b = 0; b < kMaxClass; ++b) {
        if (b < p.nlow) {
          const uint32_t ld = p.low[b];
          const uint32_t m = ld & 0xffffu;
          const int kind = (ld >> 16) & 1, type = (ld >> 17) & 3;
          if (__builtin_popcount(l & m) == 1) {
            if (kind) ++cc1; else ++cc0;
            if (p.flip[kind]) {
              const uint32_t l2 = l ^ m;
              uint32_t i2;
              if (type == 0)
                i2 = s_RL0[l2 & 255u] + t1;
              else if (type == 1)
                i2 = t0 + s_RL1[k0 * p.w1 + (l2 >> 8)];
              else {
                const uint32_t v0 = l2 & 255u;
                i2 = s_RL0[v0] + s_RL1[__builtin_popcount(v0) * p.w1 + (l2 >> 8)];
              }
              a += p.jh[kind] * x[B + i2];
            }
          }
        }
      }
      // diagonal
      const int c0 = p.nbk[0] - 2 * (s_cnt[0] + cc0);
      const int c1 = p.nbk[1] - 2 * (s_cnt[1] + cc1);
      double d = 0.0;
      if (p.diag_mode == 0) {
        if (p.dz_on[0] && c0 != 0) d = p.dq[0] * (double)c0;
        if (p.dz_on[1] && c1 != 0) d = d + p.dq[1] * (double)c1;
      } else if (p.diag_mode == 1) {
        double dis = 0.0;
        for (int q = 0; q < p.L; ++q) dis = ((l >> q) & 1u) ? dis + s_f[q] : dis - s_f[q];
        dis += s_dish;
        d = (p.dq[0] * (double)c0 + p.dq[1] * (double)c1) + 0.5 * dis;
      }
      const uint32_t r = B + (uint32_t)i;
      const double xr = x[r];
      const double yr = d * xr + a;
      y[r] = yr;
      if (DOT) dotacc += xr * yr;
    }
  }
  if (DOT) {
    double v = dotacc;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if ((tid & 63) == 0) s_red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      double tot = 0.0;
      for (int w = 0; w < kSplitThreads / 64; ++w) tot += s_red[w];
      dot_part[blockIdx.x] = tot;
    }
  }
}

// ------------------------------------------------------------------- host --
static int count_bits(uint32_t v) { return __builtin_popcount(v); }

bool split_supported(const lsqed_ham* h) {
  if (h->N > 32 || h->N < 2) return false;
  if ((int)h->bi[0].size() + (int)h->bi[1].size() > 3 * kMaxClass) return false;
  return true;
}

// The single-model kernels and the batched kernel use different row splits
// (split_low_bits), so a model holds up to two plans.
static void*& plan_slot(lsqed_ham* h, bool batch) { return batch ? h->split_b : h->split; }

static void destroy_plan(void*& slot) {
  SplitPlan* sp = (SplitPlan*)slot;
  if (!sp) return;
  (void)hipFree(sp->d_U);
  (void)hipFree(sp->d_RL0);
  (void)hipFree(sp->d_RL1);
  (void)hipFree(sp->d_binom);
  (void)hipFree(sp->d_binom64);
  (void)hipFree(sp->d_tasks);
  (void)hipFree(sp->d_tasks_ord);
  (void)hipFree(sp->d_rows_ord);
  (void)hipFree(sp->d_rows);
  delete sp;
  slot = nullptr;
}

void split_plan_destroy(lsqed_ham* h) {
  destroy_plan(h->split);
  destroy_plan(h->split_b);
}

// Order of the work items of one H.v.  A high or cross bond sends the
// entries of row h to row h ^ m, m on the high bits; that row is read from
// HBM unless a recent item already brought it into L2 or the memory-side
// cache.  The high sites are split into `outer` and `inner` ones: items are
// processed group by group, a group being all rows with the same outer bits,
// so the rows a group reads through bonds among inner sites (and from inner
// sites to the low ones) are the group's own rows, a working set of
// 8 D / 2^ko bytes.  Each group's items are cut into 8 equal contiguous
// parts, one per XCD (the kernels hand XCD x the x-th eighth of the list),
// so that the eight XCDs work on the same group at the same time.  The outer
// sites are picked greedily to add the fewest bonds leaving the group
// (degree minus bonds to sites already outer).  ko = 0 keeps h order.  The
// groups serve the memory-side cache, not L2: L2 fills are unchanged
// (FETCH_SIZE equal within 0.3%).  With 2048-entry items the 4x8 J1-J2 H.v
// went from 44.9 to 43.6 ms at ko = 5 (150 MB groups), the 6x6 one from 781.6
// to 780.2 ms; with whole rows staged in LDS (the default specialised kernel)
// h order is best (4x8: 40.96 ms at ko = 0, 41.78 at ko = 5; with each XCD's
// rows dealt longest first and one row per block, 34.75-34.80 ms at ko = 0
// against 35.48 at ko = 3), so grouping is off unless LSQED_SPLIT_OUTER (ko)
// or LSQED_SPLIT_GROUP_MB asks for it.
static int outer_site_count(const lsqed_ham* h, int NH) {
  if (const char* e = std::getenv("LSQED_SPLIT_OUTER")) return std::max(0, std::min(NH, std::atoi(e)));
  const char* g = std::getenv("LSQED_SPLIT_GROUP_MB");
  if (!g) return 0;
  const double mb = std::max(1.0, std::atof(g));
  const double bytes = 8.0 * (double)h->D;
  int ko = 0;
  while (ko < NH && bytes / std::ldexp(1.0, ko) > mb * 1048576.0) ++ko;
  return ko;
}

// Fills `tasks` with the (row, chunk) items and `rows` with the rows (chunk
// field 0; the row-staged kernel takes a whole row per block), both in group
// order with each group cut into eighths, one per XCD, padded with items that
// process no entries (chunk field 0xff).
static void order_items(const lsqed_ham* h, int L, int NH, int ko, std::vector<uint32_t>& tasks,
                        std::vector<uint32_t>& rows) {
  std::vector<int> deg(NH, 0);
  std::vector<std::vector<int>> adj(NH, std::vector<int>(NH, 0));
  for (int k = 0; k < 2; ++k)
    for (size_t b = 0; b < h->bi[k].size(); ++b) {
      const int a = h->bi[k][b], c = h->bj[k][b];
      if (a == c) continue;
      if (a >= L) ++deg[a - L];
      if (c >= L) ++deg[c - L];
      if (a >= L && c >= L) {
        ++adj[a - L][c - L];
        ++adj[c - L][a - L];
      }
    }
  std::vector<int> outer;
  std::vector<char> is_outer(NH, 0);
  for (int step = 0; step < ko; ++step) {
    int best = -1, best_cost = 0;
    for (int q = NH - 1; q >= 0; --q) {
      if (is_outer[q]) continue;
      int cost = deg[q];
      for (int o : outer) cost -= adj[q][o];
      if (best < 0 || cost < best_cost) best = q, best_cost = cost;
    }
    outer.push_back(best);
    is_outer[best] = 1;
  }
  std::vector<int> inner;
  for (int q = 0; q < NH; ++q)
    if (!is_outer[q]) inner.push_back(q);
  // LSQED_SPLIT_FAST=k: within a group, rows follow a key whose k lowest
  // bits are the k inner high sites chosen greedily so that the most bonds
  // stay inside the sub-cube they span (high bonds with both ends in it,
  // cross bonds at one of its sites): consecutive rows, which one XCD holds
  // at the same time, then read more of their targets from each other.
  // Off by default: on 4x8 J1-J2 with rows dealt longest first, k = 4 / 7 /
  // 10 measured 34.77 / 34.61 / 34.75 ms against 34.75-34.80 without, within
  // the run-to-run spread (the key changes which rows share an XCD and their
  // order among rows of equal length).
  if (const char* e = std::getenv("LSQED_SPLIT_FAST")) {
    const int k = std::max(0, std::min((int)inner.size(), std::atoi(e)));
    std::vector<int> fast;
    std::vector<char> in_fast(NH, 0);
    for (int step = 0; step < k; ++step) {
      int best = -1, best_gain = -1;
      for (int q : inner) {
        if (in_fast[q]) continue;
        int gain = 0;
        for (int kk = 0; kk < 2; ++kk)
          for (size_t b = 0; b < h->bi[kk].size(); ++b) {
            const int a = std::min(h->bi[kk][b], h->bj[kk][b]), c = std::max(h->bi[kk][b], h->bj[kk][b]);
            if (a == c || c < L) continue;
            if (a < L) {
              gain += (c - L == q);
            } else if (a - L == q || c - L == q) {
              const int other = a - L == q ? c - L : a - L;
              gain += in_fast[other] ? 1 : 0;
            }
          }
        if (gain > best_gain) best = q, best_gain = gain;
      }
      fast.push_back(best);
      in_fast[best] = 1;
    }
    std::vector<int> order = fast;
    for (int q : inner)
      if (!in_fast[q]) order.push_back(q);
    inner = order;
  }
  const int nup = h->nup;
  uint32_t pad = 0xffffffffu;  // an item with a valid row and no entries
  std::vector<std::vector<uint32_t>> seg(8), rseg(8);
  std::vector<uint32_t> group, grows;
  for (uint32_t o = 0; o < (1u << ko); ++o) {
    group.clear();
    grows.clear();
    uint32_t ho = 0;
    for (int t = 0; t < ko; ++t)
      if ((o >> t) & 1u) ho |= 1u << outer[t];
    for (uint32_t in = 0; in < (1u << (NH - ko)); ++in) {
      uint32_t hh = ho;
      for (int t = 0; t < NH - ko; ++t)
        if ((in >> t) & 1u) hh |= 1u << inner[t];
      const int kl = nup - count_bits(hh);
      if (kl < 0 || kl > L) continue;
      if (pad == 0xffffffffu) pad = hh | (0xffu << 24);
      const int64_t R = (int64_t)binom_u64(L, kl);
      for (int64_t c = 0; c * kSplitChunk < R; ++c) group.push_back(hh | ((uint32_t)c << 24));
      grows.push_back(hh);
    }
    if (group.empty()) continue;
    const size_t per = (group.size() + 7) / 8, rper = (grows.size() + 7) / 8;
    for (size_t x = 0; x < 8; ++x) {
      for (size_t i = x * per; i < (x + 1) * per; ++i) seg[x].push_back(i < group.size() ? group[i] : pad);
      for (size_t i = x * rper; i < (x + 1) * rper; ++i) rseg[x].push_back(i < grows.size() ? grows[i] : pad);
    }
  }
  // Each XCD's rows longest first (stable in the order above), so that the
  // blocks dispatched last hold the shortest rows and the launch ends with
  // less idle time (a block's two rows, half a segment apart, then also pair
  // a longer with a shorter row).  H.v, h order / longest first: 4x8 36.13 /
  // 35.72-35.75 ms, 6x6 674.4 / 650.5 ms, 4x6 0.129 / 0.118 ms; results
  // bit-identical.  LSQED_SPLIT_LPT=0 keeps h order.
  {
    const char* e = std::getenv("LSQED_SPLIT_LPT");
    if (!(e && std::atoi(e) == 0)) {
      auto len = [&](uint32_t r) -> uint64_t {
        if ((r >> 24) != 0) return 0;  // padding
        const int kl = nup - count_bits(r & 0xffffffu);
        return (kl < 0 || kl > L) ? 0 : binom_u64(L, kl);
      };
      for (int x = 0; x < 8; ++x)
        std::stable_sort(rseg[x].begin(), rseg[x].end(), [&](uint32_t a, uint32_t b) { return len(a) > len(b); });
    }
  }
  tasks.clear();
  rows.clear();
  for (int x = 0; x < 8; ++x) {
    tasks.insert(tasks.end(), seg[x].begin(), seg[x].end());
    rows.insert(rows.end(), rseg[x].begin(), rseg[x].end());
  }
}

int split_plan_create(lsqed_ham* h, bool batch) {
  void*& slot = plan_slot(h, batch);
  if (slot) return LSQED_OK;
  if (!split_plan_supported(h)) return fail(LSQED_ERR_ARG, "row-split tables need 2 <= N <= 40");
  const int N = h->N, nup = h->nup;
  const int L = split_low_bits(h, batch), NH = N - L;
  SplitPlan* sp = new SplitPlan;
  SplitParams& p = sp->p;
  std::memset(&p, 0, sizeof(p));
  p.N = N;
  p.nup = nup;
  p.L = L;
  p.NH = NH;
  p.w1 = 1 << std::max(L - 8, 0);
  // low patterns by popcount class, numeric (= colex) order
  std::vector<uint16_t> U;
  for (int k = 0; k <= L; ++k) {
    p.offU[k] = (int)U.size();
    for (uint32_t v = 0; v < (1u << L); ++v)
      if (count_bits(v) == k) U.push_back((uint16_t)v);
  }
  std::vector<uint16_t> RL0(256, 0), RL1((size_t)9 * p.w1, 0);
  for (int v = 0; v < 256; ++v) {
    if (L < 8 && (v >> L)) continue;
    uint32_t r = 0;
    int j = 0;
    for (int q = 0; q < 8; ++q)
      if ((v >> q) & 1) r += (uint32_t)binom_u64(q, ++j);
    RL0[v] = (uint16_t)r;
  }
  for (int k = 0; k <= 8; ++k)
    for (int v = 0; v < p.w1; ++v) {
      uint32_t r = 0;
      int j = 0;
      for (int q = 0; q < L - 8; ++q)
        if ((v >> q) & 1) r += (uint32_t)binom_u64(8 + q, k + (++j));
      RL1[(size_t)k * p.w1 + v] = (uint16_t)r;
    }
  std::vector<uint32_t> bin((size_t)(N + 1) * (nup + 1));
  std::vector<uint64_t> bin64((size_t)(N + 1) * (nup + 1));
  for (int n = 0; n <= N; ++n)
    for (int k = 0; k <= nup; ++k) {
      bin64[(size_t)n * (nup + 1) + k] = binom_u64(n, k);
      bin[(size_t)n * (nup + 1) + k] = (uint32_t)bin64[(size_t)n * (nup + 1) + k];  // N <= 32: exact
    }
  // bond classes of the unspecialised kernel (N <= 32, at most kMaxClass
  // bonds per class; the specialised kernel emits its own classes)
  sp->classes_ok = N <= 32;
  for (int kind = 0; kind < 2 && sp->classes_ok; ++kind) {
    p.nbk[kind] = (int)h->bi[kind].size();
    for (size_t b = 0; b < h->bi[kind].size() && sp->classes_ok; ++b) {
      int a = h->bi[kind][b], c = h->bj[kind][b];
      int lo = std::min(a, c), hi = std::max(a, c);
      if (lo == hi) continue;  // a self bond is always parallel: diagonal only (counted via nbk)
      if (hi < L) {
        const int type = hi < 8 ? 0 : (lo >= 8 ? 1 : 2);
        if (p.nlow == kMaxClass) { sp->classes_ok = false; break; }
        p.low[p.nlow++] = ((1u << lo) | (1u << hi)) | ((uint32_t)kind << 16) | ((uint32_t)type << 17);
      } else if (lo >= L) {
        if (p.nhigh == kMaxClass) { sp->classes_ok = false; break; }
        p.high[p.nhigh++] = ((1u << (lo - L)) | (1u << (hi - L))) | ((uint32_t)kind << 31);
      } else {
        if (p.ncross == kMaxClass) { sp->classes_ok = false; break; }
        p.cross[p.ncross++] = (uint32_t)lo | ((uint32_t)(hi - L) << 8) | ((uint32_t)kind << 16);
      }
    }
  }
  // work items: (h, chunk) for every row with a valid low class
  std::vector<uint32_t> tasks, rows;
  for (uint32_t hh = 0; hh < (1u << NH); ++hh) {
    const int kl = nup - count_bits(hh);
    if (kl < 0 || kl > L) continue;
    const int64_t R = (int64_t)binom_u64(L, kl);
    for (int64_t c = 0; c * kSplitChunk < R; ++c) tasks.push_back(hh | ((uint32_t)c << 24));
    rows.push_back(hh);
  }
  // Batched plans deal a model's rows longest first (stable in h): the
  // blocks that take the last rows of a launch hold short ones, so the
  // launch ends with less idle time.  256-chain disorder workload: all-active
  // step 2.39-2.42 -> 2.29-2.31 ms, 1137-1139 -> 1184-1188 realisations/s.
  // LSQED_BATCH_ROW_ORDER=0 keeps h order.
  if (batch) {
    const char* e = std::getenv("LSQED_BATCH_ROW_ORDER");
    if (!(e && std::atoi(e) == 0))
      std::stable_sort(rows.begin(), rows.end(), [&](uint32_t a, uint32_t b) {
        return binom_u64(L, nup - count_bits(a)) > binom_u64(L, nup - count_bits(b));
      });
  }
  sp->ntasks = (int)tasks.size();
  sp->nrows = (int)rows.size();
  const int ko = outer_site_count(h, NH);
  std::vector<uint32_t> tasks_ord, rows_ord;
  order_items(h, L, NH, ko, tasks_ord, rows_ord);
  sp->ntasks_ord = (int)tasks_ord.size();
  sp->nrows_ord = (int)rows_ord.size();
  sp->order_hash = fnv1a(rows_ord.data(), rows_ord.size() * 4, fnv1a(tasks_ord.data(), tasks_ord.size() * 4));
  p.ntasks = sp->ntasks_ord;
  auto up = [&](void** dst, const void* src, size_t bytes) -> bool {
    if (hipMalloc(dst, std::max<size_t>(bytes, 4)) != hipSuccess) return false;
    return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up((void**)&sp->d_U, U.data(), U.size() * 2) || !up((void**)&sp->d_RL0, RL0.data(), RL0.size() * 2) ||
      !up((void**)&sp->d_RL1, RL1.data(), RL1.size() * 2) || !up((void**)&sp->d_binom, bin.data(), bin.size() * 4) ||
      !up((void**)&sp->d_binom64, bin64.data(), bin64.size() * 8) ||
      !up((void**)&sp->d_tasks, tasks.data(), tasks.size() * 4) ||
      !up((void**)&sp->d_tasks_ord, tasks_ord.data(), tasks_ord.size() * 4) ||
      !up((void**)&sp->d_rows_ord, rows_ord.data(), rows_ord.size() * 4) ||
      !up((void**)&sp->d_rows, rows.data(), rows.size() * 4)) {
    slot = sp;
    destroy_plan(slot);
    return fail(LSQED_ERR_HIP, "row-split table upload failed");
  }
  p.U = sp->d_U;
  p.RL0 = sp->d_RL0;
  p.RL1 = sp->d_RL1;
  p.binom = sp->d_binom;
  p.tasks = sp->d_tasks_ord;
  slot = sp;
  return LSQED_OK;
}

int split_tables(lsqed_ham* h, SplitTables& t, bool batch) {
  int rc = split_plan_create(h, batch);
  if (rc) return rc;
  SplitPlan* sp = (SplitPlan*)plan_slot(h, batch);
  t.U = sp->d_U;
  t.RL0 = sp->d_RL0;
  t.RL1 = sp->d_RL1;
  t.binom = sp->d_binom;
  t.binom64 = sp->d_binom64;
  t.tasks = sp->d_tasks;
  t.ntasks = sp->ntasks;
  t.tasks_ord = sp->d_tasks_ord;
  t.ntasks_ord = sp->ntasks_ord;
  t.rows_ord = sp->d_rows_ord;
  t.nrows_ord = sp->nrows_ord;
  t.rows = sp->d_rows;
  t.nrows = sp->nrows;
  t.offU = sp->p.offU;
  t.order_hash = sp->order_hash;
  return LSQED_OK;
}

// Low bits of the row split.  The batched kernel takes L = min(14, N) (its
// rows are staged in LDS like the single-model kernel's; many problems give
// the parallelism; LSQED_BATCH_SPLIT_L, 8..16, overrides: on the 256-chain
// disorder batch L = 11 / 12 / 13 / 14 / 15 / 16 gave 733 / 840 / 866 / 940 /
// 835 / 793 realisations/s).  The single-model kernels (one
// row-staged block per row) take, for N >= 20, L = clamp(N - 10, 10, 14) on
// lattices with more than 2.5 flippable bonds per site and clamp(N - 11, 10,
// 13) on sparser ones (a further low site moves only ~2 bonds into LDS there
// while halving the rows), min(N, 10) below N = 20, and at least N - 24 (a
// work item holds the high pattern in 24 bits): about 2^11-2^12 rows or more
// to spread over the CUs, rows of at most C(14, 7) = 3432 entries, so that
// four row-staged 256-thread blocks share a CU (LDS) against one at L = 16
// (C(16, 8) = 12870 entries, 103 KB).  Measured (specialised H.v, ms, at
// L = 12 / 13 / 14): 4x6 J1-J2 0.140 / 0.130 / 0.129; 2x12 0.155 / 0.146 /
// 0.132; 4x8 47.3 / 42.4 / 39.4 (40.5 at 16); 6x6 772 / 717 at 13 / 14 (721
// at 16); chains with J2 = 0.2: N = 24 0.0614 / 0.0605 / 0.0682, N = 26
// 0.219 / 0.203 / 0.221, N = 28 0.827 / 0.882 at 13 / 14; N = 18 / 20 / 22
// at L = 10: 9.7 / 11.6 / 22.3 us against 49 / 94 / 105 us at L = 16.
// With one row per block and the one-row kernel form, L = 14 / 15 (512-thread
// blocks at 15): 4x7 1.818-1.832 / 1.834-1.856 ms, 5x6 7.34 / 7.37-7.38, 6x5
// 6.90 / 7.09-7.10, 8x4 29.10-29.17 / 30.15-30.17, 4x9 648.3 / 681.0, 6x6
// 589.8 / 623.3, and only 4x8 the other way, 34.19-34.28 / 33.88-33.92 ms (the
// same torus as 8x4 under another site numbering, which decides how many
// bonds fall among the L lowest sites), so 14 stays.
// LSQED_SPLIT_L (8..16, keeping at most 24 high bits) overrides.
int split_low_bits(const lsqed_ham* h, bool batch) {
  const int N = h->N;
  if (batch) {
    int L = std::min(14, N);
    if (const char* e = std::getenv("LSQED_BATCH_SPLIT_L")) {
      const int v = std::atoi(e);
      if (v >= 8 && v <= 16 && v <= N && N - v <= 24) L = v;
    }
    return L;
  }
  int nb = 0;
  for (int k = 0; k < 2; ++k)
    for (size_t b = 0; b < h->bi[k].size(); ++b) nb += h->bi[k][b] != h->bj[k][b];
  const bool dense = 2 * nb > 5 * N;
  int L = std::min(N, 10);
  if (N >= 20) L = dense ? std::max(10, std::min(14, N - 10)) : std::max(10, std::min(13, N - 11));
  L = std::max(L, N - 24);
  if (const char* e = std::getenv("LSQED_SPLIT_L")) {
    const int v = std::atoi(e);
    if (v >= 8 && v <= 16 && v <= N && N - v <= 24) L = v;
  }
  return L;
}

// The specialised kernel handles up to 64 high bonds (one prologue lane each).
bool split_plan_supported(const lsqed_ham* h) { return h->N >= 2 && h->N <= 40; }

bool split_jit_fits(const lsqed_ham* h, bool batch) {
  if (!split_plan_supported(h)) return false;
  if ((int)h->bi[0].size() + (int)h->bi[1].size() > 2 * kMaxBonds) return false;
  const int L = split_low_bits(h, batch);
  int nhigh = 0;
  for (int k = 0; k < 2; ++k)
    for (size_t b = 0; b < h->bi[k].size(); ++b)
      if (std::min(h->bi[k][b], h->bj[k][b]) >= L && h->bi[k][b] != h->bj[k][b]) ++nhigh;
  return nhigh <= 256;
}

// One work item per block: with the XCD-aware block index, the blocks an
// XCD runs at any moment hold neighbouring work items (chunks of the same and
// adjacent rows), whose target rows then meet in that XCD's L2.  Measured on
// the 4x8 J1-J2 sector: 54.4 ms at 8 blocks per CU (146 items per block),
// 47.9 ms at 128, 45.0 ms at one item per block.
int split_grid(const lsqed_ham* h) {
  const SplitPlan* sp = (const SplitPlan*)h->split;
  int64_t g = std::max(sp ? sp->ntasks_ord : 1, 1);
  g = ((g + 7) / 8) * 8;
  return (int)g;
}

int launch_hv_split(lsqed_ham* h, const double* x, double* y, double* dot_part, int grid, hipStream_t s) {
  int rc = split_plan_create(h);
  if (rc) return rc;
  SplitPlan* sp = (SplitPlan*)h->split;
  if (!sp->classes_ok)
    return fail(LSQED_ERR_ARG, "the unspecialised row split needs N <= 32 and at most 64 bonds per class");
  SplitParams p = sp->p;
  const RowParams rp = make_row_params(h);
  p.fields = h->d_fields;
  for (int k = 0; k < 2; ++k) {
    p.jh[k] = rp.jh[k];
    p.dq[k] = rp.dq[k];
    p.flip[k] = rp.flip[k];
    p.dz_on[k] = rp.dz_on[k];
  }
  p.diag_mode = rp.diag_mode;
  if (grid <= 0) grid = split_grid(h);
  if (dot_part)
    hipLaunchKernelGGL(hv_split_kernel<true>, dim3(grid), dim3(kSplitThreads), 0, s, p, x, y, dot_part);
  else
    hipLaunchKernelGGL(hv_split_kernel<false>, dim3(grid), dim3(kSplitThreads), 0, s, p, x, y, nullptr);
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

}  // namespace lsqed
