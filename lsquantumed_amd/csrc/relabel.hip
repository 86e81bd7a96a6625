// relabel.hip -- site relabelling for the device-resident eigensolver.
//
// The row split of the specialised H.v serves the bonds among the L lowest
// sites from the row staged in LDS, gathers the cross bonds (one low, one high
// site) entry by entry and streams the rows of the high-bond targets.  Which
// bonds fall in which class follows from the site numbering alone: on the
// periodic 4x8 J1-J2 lattice (square2D(4, 8), src/lattice.cpp:86-136) the 14
// lowest sites hold 31 flipping bonds and 50 cross the split, while the same
// torus numbered as square2D(8, 4) holds 43 and 26, and its H.v takes 29.2 ms
// against 34.2.  The Krylov-Schur solver keeps its basis on the device and
// hands out only eigenpairs, so it can run on a renumbered copy of the
// operator: the L sites with the most flipping bonds among themselves become
// 0..L-1 (ascending), the others follow in the order in which they join the
// growing cluster (most bonds to the sites already numbered first), so that
// the high sites bonded to the low block take the low high bits and their
// target rows lie close to the row in memory.  Measured on the 4x8 sector:
// 30.6 ms with the high sites in ascending order, 29.8 in breadth-first
// order, 29.3 in this order; the order alone is no gain (6x6, whose low block
// is already the densest: 589 ms as numbered, 595 breadth-first, 625 in this
// order), so a renumbering must first add low-block bonds.  Eigenvectors go
// back to the reference's basis
// order by one gather pass each (dst[r] = src[rank(pi(unrank(r)))]).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "lsqed_internal.hpp"
#include "sector_device.hpp"

namespace lsqed {

namespace {

// Flipping-bond multiplicities between sites (bonds of kinds with J != 0).
std::vector<int> flip_adjacency(const lsqed_ham* h) {
  const int N = h->N;
  std::vector<int> A((size_t)N * N, 0);
  for (int k = 0; k < 2; ++k) {
    if (h->J[k] == 0.0) continue;
    for (size_t b = 0; b < h->bi[k].size(); ++b) {
      const int i = h->bi[k][b], j = h->bj[k][b];
      if (i == j) continue;
      ++A[(size_t)i * N + j];
      ++A[(size_t)j * N + i];
    }
  }
  return A;
}

int bonds_into(const std::vector<int>& A, int N, int v, const std::vector<char>& in) {
  int s = 0;
  for (int u = 0; u < N; ++u) s += in[u] ? A[(size_t)v * N + u] : 0;
  return s;
}

int bonds_inside(const std::vector<int>& A, int N, const std::vector<char>& in) {
  int s = 0;
  for (int v = 0; v < N; ++v)
    if (in[v]) s += bonds_into(A, N, v, in);
  return s / 2;
}

// The L-site set with the most flipping bonds inside: greedy growth from
// every site, each result improved by single swaps until none helps.
std::vector<char> densest_block(const std::vector<int>& A, int N, int L, int* inside) {
  std::vector<char> best;
  int best_in = -1;
  for (int s0 = 0; s0 < N; ++s0) {
    std::vector<char> in(N, 0);
    in[s0] = 1;
    for (int n = 1; n < L; ++n) {
      int pick = -1, pv = -1;
      for (int v = 0; v < N; ++v) {
        if (in[v]) continue;
        const int c = bonds_into(A, N, v, in);
        if (c > pv) pick = v, pv = c;
      }
      in[pick] = 1;
    }
    for (bool better = true; better;) {
      better = false;
      for (int u = 0; u < N && !better; ++u) {
        if (!in[u]) continue;
        const int du = bonds_into(A, N, u, in);
        for (int v = 0; v < N; ++v) {
          if (in[v]) continue;
          if (bonds_into(A, N, v, in) - A[(size_t)v * N + u] > du) {
            in[u] = 0;
            in[v] = 1;
            better = true;
            break;
          }
        }
      }
    }
    const int c = bonds_inside(A, N, in);
    if (c > best_in) best_in = c, best = in;
  }
  *inside = best_in;
  return best;
}

template <typename S, typename R, int W, int NC>
__global__ void k_relabel_gather(RowParams p, const R* __restrict__ tab, const R* __restrict__ binom,
                                 const int* __restrict__ perm, const double* __restrict__ src,
                                 double* __restrict__ dst) {
  __shared__ int sp[64];
  for (int i = threadIdx.x; i < p.N; i += blockDim.x) sp[i] = perm[i];
  __syncthreads();
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < p.D; r += (int64_t)gridDim.x * blockDim.x) {
    const S s = unrank_row<S, R>((R)r, p.N, p.nup, binom);
    S t = 0;
    for (int i = 0; i < p.N; ++i) t |= ((s >> i) & (S)1) << sp[i];
    dst[r] = src[(int64_t)rank_state<S, R, W, NC>(t, tab, p)];
  }
}

}  // namespace

// perm[old site] = new site, when the renumbering puts at least 4 more
// flipping bonds among the L low sites of the specialised kernel's row split.
// LSQED_RELABEL=0 keeps the reference's numbering, =1 renumbers any sector
// where that holds; by default sectors of 2^24 states or more (where the H.v
// dominates a solve and one gather pass per eigenvector is negligible).
bool relabel_plan(const lsqed_ham* h, std::vector<int>& perm) {
  const char* e = std::getenv("LSQED_RELABEL");
  const int mode = e ? std::atoi(e) : -1;
  if (mode == 0) return false;
  if (mode != 1 && h->D < ((int64_t)1 << 24)) return false;
  if (h->hv_variant != 0 && h->hv_variant != 3) return false;  // a kernel chosen by the caller stays
  if (!split_jit_fits(h, false)) return false;
  const int N = h->N, L = split_low_bits(h, false);
  if (L >= N || L < 1) return false;
  const std::vector<int> A = flip_adjacency(h);
  std::vector<char> ident(N, 0);
  for (int v = 0; v < L; ++v) ident[v] = 1;
  int inside = 0;
  const std::vector<char> in = densest_block(A, N, L, &inside);
  if (inside < bonds_inside(A, N, ident) + 4) return false;
  perm.assign(N, -1);
  int next = 0;
  std::vector<char> placed(N, 0);
  for (int v = 0; v < N; ++v)
    if (in[v]) perm[v] = next++, placed[v] = 1;
  while (next < N) {
    int pick = -1, pv = -1;
    for (int v = 0; v < N; ++v) {
      if (placed[v]) continue;
      const int c = bonds_into(A, N, v, placed);
      if (c > pv) pick = v, pv = c;
    }
    perm[pick] = next++;
    placed[pick] = 1;
  }
  return true;
}

// Bring a renumbered copy in line with its model: build flags, couplings,
// fields (renumbered) and kernel choice; only what differs is set, since the
// setters drop the specialised kernel.
int relabel_sync(const lsqed_ham* h, lsqed_ham* hk, const std::vector<int>& perm) {
  int rc;
  if (h->J[0] != hk->J[0] || h->Dz[0] != hk->Dz[0] || h->J[1] != hk->J[1] || h->Dz[1] != hk->Dz[1])
    if ((rc = lsqed_ham_set_couplings(hk, h->J[0], h->Dz[0], h->J[1], h->Dz[1]))) return rc;
  if (!h->fields.empty()) {
    std::vector<double> f(h->N);
    for (int i = 0; i < h->N; ++i) f[perm[i]] = h->fields[i];
    if (f != hk->fields && (rc = lsqed_ham_set_fields(hk, f.data()))) return rc;
  }
  if (h->build != hk->build && (rc = lsqed_ham_set_build(hk, h->build))) return rc;
  if (h->hv_variant != hk->hv_variant && (rc = lsqed_ham_set_hv_kernel(hk, h->hv_variant))) return rc;
  return LSQED_OK;
}

// The renumbered copy of h (bonds of each kind in the same order, each site
// mapped through perm), synchronised with h.
int relabel_handle(const lsqed_ham* h, const std::vector<int>& perm, lsqed_ham** out) {
  std::vector<int> b[2][2];
  for (int k = 0; k < 2; ++k)
    for (size_t q = 0; q < h->bi[k].size(); ++q) {
      b[k][0].push_back(perm[h->bi[k][q]]);
      b[k][1].push_back(perm[h->bj[k][q]]);
    }
  lsqed_params p{};
  p.nspins = h->N;
  p.nup = h->nup;
  p.n_nn = (int)b[0][0].size();
  p.n_nnn = (int)b[1][0].size();
  p.nn_i = b[0][0].data();
  p.nn_j = b[0][1].data();
  p.nnn_i = b[1][0].data();
  p.nnn_j = b[1][1].data();
  p.J1 = h->J[0];
  p.D1 = h->Dz[0];
  p.J2 = h->J[1];
  p.D2 = h->Dz[1];
  lsqed_ham* hk = nullptr;
  int rc = lsqed_ham_create(&p, &hk);
  if (rc) return rc;
  if ((rc = lsqed_ham_set_build(hk, 0)) || (rc = relabel_sync(h, hk, perm))) {
    lsqed_ham_destroy(hk);
    return rc;
  }
  *out = hk;
  return LSQED_OK;
}

// dst (h's basis order) from src (the renumbered basis): dst[r] =
// src[rank(pi(unrank(r)))], d_perm the N entries of perm on the device; with
// the inverse permutation, the other way round.
int relabel_gather(const lsqed_ham* h, const int* d_perm, const double* src, double* dst, hipStream_t s) {
  if (h->D == 0) return LSQED_OK;
  const RowParams p = make_row_params(h);
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>(8192, (h->D + 255) / 256));
  if (!h->tab.wide)
    hipLaunchKernelGGL((k_relabel_gather<uint32_t, uint32_t, 8, 4>), dim3(g), dim3(256), 0, s, p,
                       (const uint32_t*)h->d_table, (const uint32_t*)h->d_binom, d_perm, src, dst);
  else
    hipLaunchKernelGGL((k_relabel_gather<uint64_t, uint64_t, 6, kMaxChunks>), dim3(g), dim3(256), 0, s, p,
                       (const uint64_t*)h->d_table, (const uint64_t*)h->d_binom, d_perm, src, dst);
  LSQED_HIP(hipGetLastError());
  return LSQED_OK;
}

}  // namespace lsqed
