// hv_jit.cpp -- lattice-specialised row-split H.v, compiled at run time.
//
// The row-split kernel of hv_split.hip reads its bond descriptors from kernel
// arguments; with ~128 bonds the compiler hoists every decoded descriptor and
// spills scalar registers.  Here the same algorithm is emitted as HIP source
// with the bond classes as constexpr arrays (site indices, bond kinds, chunk
// types become immediates, every bond loop fully unrolls), compiled once per
// bond structure with hipRTC for gfx950 and cached for the process.  Couplings,
// fields and build flags stay run-time arguments, so one compiled kernel serves
// every realisation of a lattice.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <utility>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "lsqed_internal.hpp"

namespace lsqed {

// Device code of the specialised kernel.  The generated prologue defines
// N, NUP, L, NH, W1, the bond class arrays (LOW*, CR*, HI*) and their counts.
static const char* kJitBody = R"JIT(
#ifndef MASKED_GATHER
#define MASKED_GATHER 0
#endif
#ifndef ROWSTAGE
#define ROWSTAGE 0
#endif
#ifndef NT_STORE
#define NT_STORE 0
#endif
#ifndef NT_STAGE
#define NT_STAGE 0
#endif
// ONE_ROW (single-model ROWSTAGE, grid of one block per row): the row loop
// runs once, so that nothing computed before it stays live across rows (the
// looped form keeps tables and constants in registers and spills 12 VGPRs at
// four waves per SIMD, 52 bytes per lane of scratch stores per block)
#ifndef ONE_ROW
#define ONE_ROW 0
#endif
// LOWTAB (with ROWSTAGE): the low bonds of an entry come from a host-built
// record per low pattern l -- for each flipping low bond (LOWFK order) the
// partner's index in the row, or RMAX (a zero entry of the staged row) when
// its sites are parallel, then l and the antiparallel counts of both kinds --
// instead of popcounts and rank-table lookups per bond
#ifndef LOWTAB
#define LOWTAB 0
#endif
#ifndef PREFETCH
#define PREFETCH 0
#endif
// HB_VEC: a batch's high-bond row bases are read from LDS together before
// its gathers are issued (the table is padded to a multiple of the batch)
#ifndef HB_VEC
#define HB_VEC 0
#endif
// BUF_CROSS (with ROWSTAGE): cross gathers are raw buffer loads over the
// target row; a lane whose bond does not flip passes an offset past the
// buffer's size, which returns zero without a memory access
#ifndef BUF_CROSS
#define BUF_CROSS 0
#endif
// BUF_HIGH (with BUF_CROSS and HB_VEC): the high-bond loads too, with the
// entry's byte offset shared by all of them
#ifndef BUF_HIGH
#define BUF_HIGH 0
#endif
// ROWSTAGE: blocks of ROW_THREADS threads take whole rows, staging each
// row's x in LDS, so that the low bonds (targets inside the row) read LDS
// instead of issuing global gathers
#ifndef ROW_THREADS
#define ROW_THREADS 1024
#endif
constexpr int kThreads = ROWSTAGE ? ROW_THREADS : 256;
constexpr int kPerLane = 8;
#ifndef GATHER_BATCH
#define GATHER_BATCH 8
#endif
constexpr int kG = GATHER_BATCH;  // gathers issued together before any is used
constexpr int kChunk = kThreads * kPerLane;

#if WIDE
typedef unsigned long long BT;  // row bases and ranks (D >= 2^32)
#else
typedef unsigned int BT;
#endif

// Vector entries: FP64, or interleaved complex FP64 (CPLX: a complex-scalar
// PETSc Vec); H is real, so both parts take the same coefficients, each
// updated with the expression of the real kernel.
#ifndef CPLX
#define CPLX 0
#endif
#if CPLX
typedef double2 XV;
__device__ __forceinline__ XV xzero() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ void xmadd(XV& a, double c, XV v) {
  a.x += c * v.x;
  a.y += c * v.y;
}
__device__ __forceinline__ XV xdiag(double d, XV x, XV acc) { return make_double2(d * x.x + acc.x, d * x.y + acc.y); }
__device__ __forceinline__ double xdot(XV a, XV b) { return a.x * b.x + a.y * b.y; }
#else
typedef double XV;
__device__ __forceinline__ XV xzero() { return 0.0; }
__device__ __forceinline__ void xmadd(XV& a, double c, XV v) { a += c * v; }
__device__ __forceinline__ XV xdiag(double d, XV x, XV acc) { return d * x + acc; }
__device__ __forceinline__ double xdot(XV a, XV b) { return a * b; }
#endif

#if BUF_CROSS
typedef int lsq_i32x4 __attribute__((ext_vector_type(4)));
#if CPLX
typedef int lsq_ldv __attribute__((ext_vector_type(4)));
__device__ lsq_ldv lsq_buf_load(lsq_i32x4 rsrc, int voffset, int soffset, int aux) __asm("llvm.amdgcn.raw.buffer.load.v4i32");
#else
typedef int lsq_ldv __attribute__((ext_vector_type(2)));
__device__ lsq_ldv lsq_buf_load(lsq_i32x4 rsrc, int voffset, int soffset, int aux) __asm("llvm.amdgcn.raw.buffer.load.v2i32");
#endif
// buffer descriptor of a row starting at p (wave-uniform): 48-bit base,
// stride 0, num_records 0x7ffffff0 bytes, data format 32 (gfx9 raw buffer)
__device__ __forceinline__ lsq_i32x4 row_rsrc(const XV* p) {
  const unsigned long long a = (unsigned long long)p;
  lsq_i32x4 r;
  r.x = __builtin_amdgcn_readfirstlane((int)(unsigned)a);
  r.y = __builtin_amdgcn_readfirstlane((int)((unsigned)(a >> 32) & 0xffffu));
  r.z = 0x7ffffff0;
  r.w = 0x00020000;
  return r;
}
constexpr int kBufNone = (int)0x80000000u;  // offset of a non-flipping lane
#endif

__device__ __forceinline__ BT rfl(BT v) {
#if WIDE
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((BT)hi << 32) | lo;
#else
  return __builtin_amdgcn_readfirstlane(v);
#endif
}

struct JitArgs {
  const double* x;
  double* y;
  double* dot;
  const double* fields;
  const unsigned short* U;
  const unsigned short* RL0;
  const unsigned short* RL1;
  const unsigned int* binom;
  const unsigned long long* binom64;
  const unsigned int* tasks;
  int ntasks;
  int offU[17];
  const uint4* lowtab;     // LOWTAB: RECQ x 16 bytes per low pattern, in U's order
#if BATCH
  const long long* poff;   // per problem: offset of its entries in x and y
  const double* pcoup;     // per problem: J1/2, J2/2, D1/4, D2/4
  const double* pfields;   // per problem: 32 fields
  const int* active;       // per problem: still iterating
  const int* gidx;         // the group's problems (batch indices)
  int nprob;               // problems in the group
#endif
};

#if ROWSTAGE
constexpr int kStage = (RMAX + kThreads - 1) / kThreads;  // entries of a row per thread
#endif

template <bool DOT>
__device__ __forceinline__ void hv_body(const JitArgs& a) {
  // the row's bond tables; under ROWSTAGE double buffered by row parity, so
  // that the next row's are built while waves still read the current ones
  constexpr int NBUF = ROWSTAGE ? 2 : 1;
  __shared__ unsigned short s_RL0[256];
  __shared__ unsigned short s_RL1[9 * W1];
  __shared__ BT s_binom[(N + 1) * (NUP + 1)];
  __shared__ double s_f[N];
  constexpr int HBPAD = ((NHIGH + kG) / kG) * kG;  // > NHIGH, a multiple of the gather batch
  __shared__ __attribute__((aligned(16))) BT s_hbase[NBUF][HBPAD];
  __shared__ __attribute__((aligned(16))) double s_hcoef[NBUF][HBPAD];
  __shared__ BT s_cbase[NBUF][NH + 1];
  __shared__ BT s_B;
  __shared__ int s_nact[NBUF], s_cnt0[NBUF], s_cnt1[NBUF];
  __shared__ double s_dish[NBUF];
  __shared__ double s_red[kThreads / 64];
#if ROWSTAGE
  __shared__ XV s_x[RMAX + LOWTAB];
#endif
  const int tid = threadIdx.x;
#if ROWSTAGE && LOWTAB
  if (tid == 0) s_x[RMAX] = xzero();  // rows have at most RMAX entries: never overwritten
#endif
  for (int i = tid; i < 256; i += kThreads) s_RL0[i] = a.RL0[i];
  for (int i = tid; i < 9 * W1; i += kThreads) s_RL1[i] = a.RL1[i];
#if WIDE
  for (int i = tid; i < (N + 1) * (NUP + 1); i += kThreads) s_binom[i] = a.binom64[i];
#else
  for (int i = tid; i < (N + 1) * (NUP + 1); i += kThreads) s_binom[i] = a.binom[i];
#endif
#if !BATCH
  if (DIAG_MODE == 1)
    for (int i = tid; i < N; i += kThreads) s_f[i] = a.fields[i];
#endif
  __syncthreads();
  constexpr BT kNone = ~(BT)0;
  auto C = [&](int n, int k) -> BT { return (k < 0 || k > NUP) ? (BT)0 : s_binom[n * (NUP + 1) + k]; };
  auto rowbase = [&](unsigned int hh) -> BT {
    const int kl = NUP - __builtin_popcount(hh);
    BT b = 0;
    int j = 0;
    for (int q = 0; q < NH; ++q)
      if ((hh >> q) & 1u) { ++j; b += C(L + q, kl + j); }
    return b;
  };
  constexpr int flip[2] = {FLIP0, FLIP1};
  const long long G = gridDim.x;
#if BATCH
  // tasks t = (problem of the group) * ntasks + (row-split task), dealt
  // cyclically so that all blocks work on neighbouring problems together
  const long long lb = (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
  const long long ntot = (long long)a.nprob * a.ntasks;
  int loaded = -1;
#if ROWSTAGE
  int pb = 0;  // table buffer of the block's current row (alternates per row)
#endif
  for (long long t = lb; t < ntot; t += G) {
    const int pl = (int)(t / a.ntasks);
    const int st = (int)(t - (long long)pl * a.ntasks);
    const int prob = a.gidx[pl];
    if (!a.active[prob]) continue;  // block-uniform
    const XV* __restrict__ xv = (const XV*)a.x + a.poff[prob];
    XV* __restrict__ yv = (XV*)a.y + a.poff[prob];
    const double jh_[2] = {a.pcoup[4 * prob], a.pcoup[4 * prob + 1]};
    const double dq0_ = a.pcoup[4 * prob + 2], dq1_ = a.pcoup[4 * prob + 3];
    if (DIAG_MODE == 1 && prob != loaded) {
      __syncthreads();
      for (int i = tid; i < N; i += kThreads) s_f[i] = a.pfields[32 * prob + i];
      loaded = prob;
#if ROWSTAGE
      __syncthreads();  // the row prologue reads the fields before its own barrier
#endif
    }
    double dotacc = 0.0;
#if !ROWSTAGE
    constexpr int pb = 0;
#endif
    const unsigned int task = a.tasks[st];
#elif ROWSTAGE
  // The row list holds eight equal segments, one per XCD; block b runs on
  // XCD b % 8 (the dispatcher deals blocks round-robin over the XCDs) and
  // takes rows b / 8, b / 8 + P, ... of its segment (P blocks per XCD; one
  // row per block when the grid covers the list), so the blocks of one XCD
  // work on adjacent rows at a time and meet in its L2.  Giving a block
  // adjacent rows instead (rows jK .. jK+K-1) measured slower at two rows
  // per block (4x8 37.8 vs 37.2-37.4 ms, 6x6 700 vs 686 ms): neighbouring
  // rows have correlated lengths, so a block's pair is less often balanced.
  const int xcd = blockIdx.x & 7, P = (int)(G >> 3), seg = a.ntasks >> 3;
  double dotacc = 0.0;
  int pb = 0;
  for (int t = (int)(blockIdx.x >> 3); t < seg; t += (ONE_ROW ? seg : P), pb ^= 1) {
    const XV* __restrict__ xv = (const XV*)a.x;
    XV* __restrict__ yv = (XV*)a.y;
    constexpr double jh_[2] = {JH0, JH1};
    constexpr double dq0_ = DQ0, dq1_ = DQ1;
    const unsigned int task = a.tasks[xcd * seg + t];
#else
  const long long lb = (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
  const long long ta = lb * a.ntasks / G, tb = (lb + 1) * a.ntasks / G;
  double dotacc = 0.0;
  constexpr int pb = 0;
  for (long long t = ta; t < tb; ++t) {
    const XV* __restrict__ xv = (const XV*)a.x;
    XV* __restrict__ yv = (XV*)a.y;
    constexpr double jh_[2] = {JH0, JH1};
    constexpr double dq0_ = DQ0, dq1_ = DQ1;
    const unsigned int task = a.tasks[t];
#endif
    const unsigned int h = task & 0xffffffu;
    const int kl = NUP - __builtin_popcount(h);
#if ROWSTAGE
    const int R = (task >> 24) ? 0 : (int)C(L, kl);  // padding items (chunk field 0xff) hold no row
    const int i0 = 0;
    // wave 0 (high bonds, compacted with a ballot) and wave 1 (cross rows,
    // field sum) fill this row's table buffer while the previous row's waves
    // finish; then every thread issues its staging loads, which are in
    // flight across the barrier (issued after the tables, so that the staged
    // values and the table work are not live together)
    const BT B = rfl(rowbase(h));
    if (tid < 64) {
      int nact = 0, c0 = 0, c1 = 0;
#pragma unroll
      for (int g = 0; g < NHIGH; g += 64) {  // lane tid: high bond g + tid
        const int b = g + tid;
        bool anti = false, act = false;
        int kind = 0;
        BT base = 0;
        if (b < NHIGH) {
          const unsigned int m = HIM[b];
          kind = HIK[b];
          anti = __builtin_popcount(h & m) == 1;
          act = anti && flip[kind];
          if (act) base = rowbase(h ^ m);
        }
        const unsigned long long bal = __ballot(act);
        const int pos = nact + __popcll(bal & ((1ull << tid) - 1ull));
        if (act) {
          s_hbase[pb][pos] = base;
          s_hcoef[pb][pos] = jh_[kind];
        }
        nact += __popcll(bal);
        c0 += __popcll(__ballot(anti && kind == 0));
        c1 += __popcll(__ballot(anti && kind == 1));
      }
      if (tid == 0) {
        s_nact[pb] = nact;
        s_cnt0[pb] = c0;
        s_cnt1[pb] = c1;
      }
    } else if (tid < 64 + NH) {
      s_cbase[pb][tid - 64] = rowbase(h ^ (1u << (tid - 64)));
    } else if (tid == 127) {
      double dh = 0.0;
      if (DIAG_MODE == 1)
        for (int q = 0; q < NH; ++q) dh = ((h >> q) & 1u) ? dh + s_f[L + q] : dh - s_f[L + q];
      s_dish[pb] = dh;
    }
    XV stage[kStage];
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int e = tid + u * kThreads < R ? tid + u * kThreads : R - 1;
      stage[u] = xzero();
#if NT_STAGE && !CPLX
      if (u * kThreads < R) stage[u] = __builtin_nontemporal_load(&xv[B + (BT)e]);
#else
      if (u * kThreads < R) stage[u] = xv[B + (BT)e];
#endif
    }
    __syncthreads();  // every wave is done with the previous row's x
#pragma unroll
    for (int u = 0; u < kStage; ++u)
      if (tid + u * kThreads < R) s_x[tid + u * kThreads] = stage[u];
    __syncthreads();
#else
    const int R = (int)C(L, kl);
    const int i0 = (int)(task >> 24) * kChunk;
    __syncthreads();
    if (tid < 64) {
      for (int b = tid; b < NHIGH; b += 64) {
        const bool anti = __builtin_popcount(h & HIM[b]) == 1;
        s_hbase[0][b] = (anti && flip[HIK[b]]) ? rowbase(h ^ HIM[b]) : kNone;
      }
    } else if (tid < 64 + NH) {
      s_cbase[0][tid - 64] = rowbase(h ^ (1u << (tid - 64)));
    } else if (tid == 192) {
      s_B = rowbase(h);
      int c0 = 0, c1 = 0;
#pragma unroll
      for (int b = 0; b < NHIGH; ++b)
        if (__builtin_popcount(h & HIM[b]) == 1) { if (HIK[b]) ++c1; else ++c0; }
      s_cnt0[0] = c0;
      s_cnt1[0] = c1;
      double dh = 0.0;
      if (DIAG_MODE == 1)
        for (int q = 0; q < NH; ++q) dh = ((h >> q) & 1u) ? dh + s_f[L + q] : dh - s_f[L + q];
      s_dish[0] = dh;
    }
    __syncthreads();
    if (tid == 0) {
      int n = 0;
      for (int b = 0; b < NHIGH; ++b)
        if (s_hbase[0][b] != kNone) { s_hbase[0][n] = s_hbase[0][b]; s_hcoef[0][n] = jh_[HIK[b]]; ++n; }
      s_nact[0] = n;
    }
    __syncthreads();
    const BT B = rfl(s_B);
#endif
    const int nact = __builtin_amdgcn_readfirstlane(s_nact[pb]);
    const int cnt0 = __builtin_amdgcn_readfirstlane(s_cnt0[pb]), cnt1 = __builtin_amdgcn_readfirstlane(s_cnt1[pb]);
    BT cb[NH + 1];
#pragma unroll
    for (int q = 0; q < NH; ++q) cb[q] = rfl(s_cbase[pb][q]);
    const unsigned short* Uk = a.U + a.offU[kl];
#if ROWSTAGE && PREFETCH
    // the pattern (or record) of a lane's next entry is loaded one entry
    // ahead, so the cross gathers, which need it, do not wait for its load
#if LOWTAB
    const uint4* __restrict__ recb = a.lowtab + (unsigned int)a.offU[kl] * RECQ;
    uint4 rq[RECQ];
#pragma unroll
    for (int q = 0; q < RECQ; ++q) rq[q] = tid < R ? recb[(unsigned int)tid * RECQ + q] : make_uint4(0u, 0u, 0u, 0u);
#else
    unsigned int l_nx = tid < R ? (unsigned int)Uk[tid] : 0u;
#endif
#endif
#pragma unroll 1
#if ROWSTAGE
    for (int k = 0; k < kStage; ++k) {
      const int i = tid + k * kThreads;
      if (i >= R) break;
#else
    for (int k = 0; k < kPerLane; ++k) {
      const int i = i0 + tid + k * kThreads;
      if (i >= R) break;
#endif
      const BT r = B + (BT)i;
      XV acc = xzero();
      // high bonds: row h ^ m, same entry; loads issued in batches of 8
      for (int b0 = 0; b0 < nact; b0 += kG) {
        XV v[kG];
#if HB_VEC
        BT hb[kG];
#pragma unroll
        for (int u = 0; u < kG; ++u) hb[u] = s_hbase[pb][b0 + u];
#endif
#pragma unroll
        for (int u = 0; u < kG; ++u) {
          v[u] = xzero();
#if HB_VEC && BUF_HIGH
          if (b0 + u < nact) v[u] = __builtin_bit_cast(XV, lsq_buf_load(row_rsrc(xv + rfl(hb[u])), (int)((unsigned)i * (unsigned)sizeof(XV)), 0, 0));
#elif HB_VEC
          if (b0 + u < nact) v[u] = xv[rfl(hb[u]) + (BT)i];
#else
          if (b0 + u < nact) v[u] = xv[rfl(s_hbase[pb][b0 + u]) + (BT)i];
#endif
        }
#pragma unroll
        for (int u = 0; u < kG; ++u)
          if (b0 + u < nact) xmadd(acc, s_hcoef[pb][b0 + u], v[u]);
      }
#if LOWTAB
      unsigned int rw[4 * RECQ];
      {
#if ROWSTAGE && PREFETCH
        const bool more = i + kThreads < R;
#pragma unroll
        for (int q = 0; q < RECQ; ++q) {
          const uint4 w = rq[q];
          if (more) rq[q] = recb[(unsigned int)(i + kThreads) * RECQ + q];
#else
        const uint4* __restrict__ rec = a.lowtab + (unsigned int)(a.offU[kl] + i) * RECQ;
#pragma unroll
        for (int q = 0; q < RECQ; ++q) {
          const uint4 w = rec[q];
#endif
          rw[4 * q] = w.x;
          rw[4 * q + 1] = w.y;
          rw[4 * q + 2] = w.z;
          rw[4 * q + 3] = w.w;
        }
      }
      const unsigned int l = rw[LWORD] & 0xffffu;
#elif ROWSTAGE && PREFETCH
      const unsigned int l = l_nx;
      if (i + kThreads < R) l_nx = Uk[i + kThreads];
#else
      const unsigned int l = Uk[i];
#endif
      const unsigned int v0 = l & 255u;
      const int k0 = __builtin_popcount(v0);
      const unsigned int t1 = s_RL1[k0 * W1 + (l >> 8)];
      const unsigned int t0 = (unsigned int)i - t1;
      int cc0 = 0, cc1 = 0;
      // cross bonds: row h ^ (1 << p), entry RL(l ^ (1 << q)); branch-free
      // index selection (non-flipping lanes read their own entry) so that the
      // gathers of a batch are all in flight together
#pragma unroll
      for (int b0 = 0; b0 < NCROSS; b0 += kG) {
        XV v[kG];
        unsigned int onm = 0;
#pragma unroll
        for (int u = 0; u < kG; ++u) {
          const int b = b0 + u;
          if (b < NCROSS) {
            const int q = CRQ[b], hp = CRP[b], kind = CRK[b];
            const unsigned int anti = ((l >> q) ^ (h >> hp)) & 1u;
            if (kind) cc1 += anti; else cc0 += anti;
            const bool on = anti && flip[kind];
            onm |= (on ? 1u : 0u) << u;
            const unsigned int l2 = l ^ (1u << q);
            unsigned int i2;
            if (q < 8) {
              const unsigned int w0 = l2 & 255u;
              i2 = s_RL0[w0] + s_RL1[__builtin_popcount(w0) * W1 + (l2 >> 8)];
            } else {
              i2 = t0 + s_RL1[k0 * W1 + (l2 >> 8)];
            }
#if MASKED_GATHER
            v[u] = xzero();
            if (on) v[u] = xv[cb[hp] + (BT)i2];
#elif BUF_CROSS && ROWSTAGE
            v[u] = __builtin_bit_cast(XV, lsq_buf_load(row_rsrc(xv + cb[hp]), on ? (int)(i2 * (unsigned)sizeof(XV)) : kBufNone, 0, 0));
#else
            const BT idx = on ? cb[hp] + (BT)i2 : r;
            v[u] = xv[idx];
#endif
          }
        }
#pragma unroll
        for (int u = 0; u < kG; ++u) {
          const int b = b0 + u;
          if (b < NCROSS) xmadd(acc, ((onm >> u) & 1u) ? jh_[CRK[b]] : 0.0, v[u]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
#if LOWTAB == 1
      // low bonds: inside row h, partner entries from the record (a parallel
      // pair reads the zero entry, so every flipping bond is one LDS read and
      // one FMA, summed in bond order as below)
#pragma unroll
      for (int b = 0; b < NLOWF; ++b) {
        const unsigned int w = rw[b >> 1];
        const unsigned int j = (b & 1) ? (w >> 16) : (w & 0xffffu);
        xmadd(acc, jh_[LOWFK[b]], s_x[j]);
      }
      cc0 += (int)((rw[LWORD] >> 16) & 0xffu);
      cc1 += (int)(rw[LWORD] >> 24);
#elif LOWTAB == 2
      // low bonds in bond order: those with a site in each byte of the
      // pattern (two rank lookups) from the record, the others by one lookup;
      // a parallel pair reads the zero entry either way, and the antiparallel
      // counts come from the record
#pragma unroll
      for (int b = 0; b < NLOW; ++b) {
        if (!flip[LOWK[b]]) continue;
        if (LOWREC[b] >= 0) {
          const unsigned int w = rw[LOWREC[b] >> 1];
          const unsigned int j = (LOWREC[b] & 1) ? (w >> 16) : (w & 0xffffu);
          xmadd(acc, jh_[LOWK[b]], s_x[j]);
        } else {
          const unsigned int m = LOWM[b];
          const unsigned int l2 = l ^ m;
          const unsigned int i2 =
              LOWT[b] == 0 ? s_RL0[l2 & 255u] + t1 : t0 + (unsigned int)s_RL1[k0 * W1 + (l2 >> 8)];
          xmadd(acc, jh_[LOWK[b]], s_x[__builtin_popcount(l & m) == 1 ? i2 : (unsigned int)RMAX]);
        }
      }
      cc0 += (int)((rw[LWORD] >> 16) & 0xffu);
      cc1 += (int)(rw[LWORD] >> 24);
#else
      // low bonds: inside row h, entry RL(l ^ m)
#pragma unroll
      for (int b0 = 0; b0 < NLOW; b0 += kG) {
        XV v[kG];
        unsigned int onm = 0;
#pragma unroll
        for (int u = 0; u < kG; ++u) {
          const int b = b0 + u;
          if (b < NLOW) {
            const unsigned int m = LOWM[b];
            const int kind = LOWK[b], type = LOWT[b];
            const unsigned int anti = __builtin_popcount(l & m) == 1 ? 1u : 0u;
            if (kind) cc1 += anti; else cc0 += anti;
            const bool on = anti && flip[kind];
            onm |= (on ? 1u : 0u) << u;
            const unsigned int l2 = l ^ m;
            unsigned int i2;
            if (type == 0) {
              i2 = s_RL0[l2 & 255u] + t1;
            } else if (type == 1) {
              i2 = t0 + s_RL1[k0 * W1 + (l2 >> 8)];
            } else {
              const unsigned int w0 = l2 & 255u;
              i2 = s_RL0[w0] + s_RL1[__builtin_popcount(w0) * W1 + (l2 >> 8)];
            }
#if ROWSTAGE
            v[u] = s_x[on ? i2 : (unsigned int)i];
#elif MASKED_GATHER
            v[u] = xzero();
            if (on) v[u] = xv[B + (BT)i2];
#else
            const BT idx = on ? B + (BT)i2 : r;
            v[u] = xv[idx];
#endif
          }
        }
#pragma unroll
        for (int u = 0; u < kG; ++u) {
          const int b = b0 + u;
          if (b < NLOW) xmadd(acc, ((onm >> u) & 1u) ? jh_[LOWK[b]] : 0.0, v[u]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
#endif
      const int c0 = NB0 - 2 * (cnt0 + cc0);
      const int c1 = NB1 - 2 * (cnt1 + cc1);
      double d = 0.0;
      {
        // no contraction: one rounding per operation, whichever kernel
        // variant (real, complex, batched) is compiled
#pragma clang fp contract(off)
        if (DIAG_MODE == 0) {
          if (DZ0 && c0 != 0) d = dq0_ * (double)c0;
          if (DZ1 && c1 != 0) d = d + dq1_ * (double)c1;
        } else if (DIAG_MODE == 1) {
          double dis = 0.0;
#pragma unroll
          for (int q = 0; q < L; ++q) dis = ((l >> q) & 1u) ? dis + s_f[q] : dis - s_f[q];
          dis += s_dish[pb];
          d = (dq0_ * (double)c0 + dq1_ * (double)c1) + 0.5 * dis;
        }
      }
#if ROWSTAGE
      const XV xr = s_x[i];
#else
      const XV xr = xv[r];
#endif
      const XV yr = xdiag(d, xr, acc);
#if NT_STORE && !CPLX
      __builtin_nontemporal_store(yr, &yv[r]);
#else
      yv[r] = yr;
#endif
      if (DOT) dotacc += xdot(xr, yr);
    }
#if BATCH
    if (DOT) {
      double v = dotacc;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      __syncthreads();
      if ((tid & 63) == 0) s_red[tid >> 6] = v;
      __syncthreads();
      if (tid == 0) {
        double tot = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) tot += s_red[w];
        a.dot[t] = tot;
      }
    }
#if ROWSTAGE
    pb ^= 1;
#endif
#endif
  }
#if !BATCH
  if (DOT) {
    double v = dotacc;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if ((tid & 63) == 0) s_red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      double tot = 0.0;
      for (int w = 0; w < kThreads / 64; ++w) tot += s_red[w];
      a.dot[blockIdx.x] = tot;
    }
  }
#endif
}

#ifndef WAVES_PER_EU
#define WAVES_PER_EU 4
#endif
#if BATCH
#define KNAME_NODOT hv_bjit_nodot
#define KNAME_DOT hv_bjit_dot
#elif CPLX
#define KNAME_NODOT hv_jit_c128_nodot
#define KNAME_DOT hv_jit_c128_dot
#else
#define KNAME_NODOT hv_jit_nodot
#define KNAME_DOT hv_jit_dot
#endif
extern "C" __global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WAVES_PER_EU)))
void KNAME_NODOT(JitArgs a) { hv_body<false>(a); }
extern "C" __global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WAVES_PER_EU)))
void KNAME_DOT(JitArgs a) { hv_body<true>(a); }
)JIT";

struct JitArgs {
  const double* x;
  double* y;
  double* dot;
  const double* fields;
  const uint16_t* U;
  const uint16_t* RL0;
  const uint16_t* RL1;
  const uint32_t* binom;
  const uint64_t* binom64;
  const uint32_t* tasks;
  int ntasks;
  int offU[17];
  const void* lowtab;
};

// Argument block of the batch variant (same leading layout).
struct BatchJitArgs {
  JitArgs base;
  const int64_t* poff;
  const double* pcoup;
  const double* pfields;
  const int* active;
  const int* gidx;
  int nprob;
};

struct JitModule {
  hipModule_t mod = nullptr;
  hipFunction_t fn_dot = nullptr, fn_nodot = nullptr;
  bool ok = false;
  bool rowstage = false;  // one block per row (rows_ord), else 256 threads per item
  bool one_row = false;   // rowstage kernel whose blocks take exactly one row
  int threads = 256;      // block size
  uint64_t src_hash = 0;  // FNV-1a of the generated source
  void* d_lowtab = nullptr;  // LOWTAB records on the module's device (kept for the process)
  std::string log;
};

// The single-model kernel stages each row of x in LDS (blocks of 256-1024
// threads, each taking whole rows) unless LSQED_JIT_ROWSTAGE=0.  The low bonds then read LDS instead
// of issuing global gathers: 37% fewer vector memory instructions on the
// 4x8 J1-J2 sector, whose H.v was bound by address processing (TA busy 92%);
// 43.7 -> 41.0 ms there, 779 -> 736 ms on 6x6, 0.185 -> 0.152 ms on 4x6,
// bit-identical output.
static bool jit_rowstage() {
  const char* e = getenv("LSQED_JIT_ROWSTAGE");
  return !(e && atoi(e) == 0);
}

// Rows a single-model row-staged block takes (LSQED_JIT_ROWS_PER_BLOCK,
// 1..64; default 1, measured at the grid below).
static int jit_rows_per_block() {
  const char* e = getenv("LSQED_JIT_ROWS_PER_BLOCK");
  return e ? std::max(1, std::min(64, atoi(e))) : 1;
}

// Threads of a row-staged block: 256 for rows of at most C(14, 7) entries
// (four blocks per CU), 512 for C(15, 7) (two), 1024 above (one block per CU
// holds a C(16, 8) row); LSQED_JIT_ROW_THREADS (256, 512 or 1024) overrides.
static int jit_row_threads(int L) {
  const char* e = getenv("LSQED_JIT_ROW_THREADS");
  const int v = e ? atoi(e) : (L <= 14 ? 256 : (L == 15 ? 512 : 1024));
  return (v == 256 || v == 512) ? v : 1024;
}

// Compiled modules by (device, generated source): a module is loaded into one
// device's context and cannot launch on another.
static std::mutex g_jit_mu;
static std::map<std::pair<int, std::string>, std::shared_ptr<JitModule>> g_jit_cache;

template <typename T>
static void emit_array(std::ostringstream& o, const char* name, const std::vector<T>& v, const char* type) {
  o << "constexpr " << type << " " << name << "[" << std::max<size_t>(v.size(), 1) << "] = {";
  if (v.empty()) o << "0";
  for (size_t i = 0; i < v.size(); ++i) o << (i ? "," : "") << v[i];
  o << "};\n";
}

static std::string dbl(double v) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.17g", v);
  std::string s(b);
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
  return s;
}

// Low-bond records (LOWTAB) for row-staged kernels with L <= 14: on for the
// batched kernel, whose chains of N = 18-22 spend most VALU issue on the
// per-bond rank arithmetic of their 25 low bonds (VALU 66% busy on the N = 22
// launch).  Mode 1 records every flipping low bond (64 B per pattern for an
// N = 22 chain), mode 2 -- the batched default -- only those with a site in
// each byte of the pattern (two rank lookups each; 3 of the 25, 16 B per
// pattern), the others costing one lookup: the 256-chain disorder solve 1099
// (mode 1) / 1127 (mode 2) realisations/s, the record loads having moved the
// bound to the texture addresser (74% busy).  Off for single models: the 4x8
// H.v took 38.92 ms with mode-1 records against 37.2 ms without, 6x6 726
// against 686 ms, and 46.5 ms with mode 2 against 36.1 (extra loads on an
// HBM-bound kernel).  LSQED_JIT_LOWTAB=0/1/2 overrides for both kernels.
static int jit_lowtab(bool batch, int L) {
  if (L > 14) return 0;
  const char* e = getenv("LSQED_JIT_LOWTAB");
  const int v = e ? atoi(e) : (batch ? 2 : 0);
  return (v == 1 || v == 2) ? v : 0;
}

// Records of the LOWTAB kernel, one per low pattern l in U's order (popcount
// class, then numeric = colex order), `recq` x 4 words each: words
// 0..(nf+1)/2-1 hold the row indices of l ^ m for the low bonds m listed in
// `f` (flipping bonds: all of them, or with LOWTAB = 2 those with a site in
// each byte of l; two 16-bit halves per word, bond order), or `zslot` when
// the bond's sites are parallel; word (nf+1)/2 holds l | cc0 << 16 | cc1 << 24,
// the counts of antiparallel low bonds of kind 0 and 1 (every low bond,
// flipping or not).
static std::vector<uint32_t> lowtab_records(int L, const std::vector<unsigned>& lowm, const std::vector<int>& lowk,
                                            const std::vector<int>& f, int recq, unsigned zslot) {
  const int lword = ((int)f.size() + 1) / 2;
  auto rank = [&](unsigned v) {
    uint32_t r = 0;
    int j = 0;
    for (int q = 0; q < L; ++q)
      if ((v >> q) & 1u) r += (uint32_t)binom_u64(q, ++j);
    return r;
  };
  std::vector<uint32_t> rec;
  for (int k = 0; k <= L; ++k)
    for (unsigned v = 0; v < (1u << L); ++v) {
      if (__builtin_popcount(v) != k) continue;
      std::vector<uint32_t> w((size_t)4 * recq, 0);
      for (size_t t = 0; t < f.size(); ++t) {
        const unsigned m = lowm[f[t]];
        const uint32_t j = __builtin_popcount(v & m) == 1 ? rank(v ^ m) : zslot;
        w[t / 2] |= (t & 1) ? j << 16 : j;
      }
      uint32_t cc[2] = {0, 0};
      for (size_t b = 0; b < lowm.size(); ++b) cc[lowk[b]] += __builtin_popcount(v & lowm[b]) == 1;
      w[lword] = v | cc[0] << 16 | cc[1] << 24;
      rec.insert(rec.end(), w.begin(), w.end());
    }
  return rec;
}

// Source of the kernel specialised to the bond classes and couplings of `h`
// (row split with L = split_low_bits(h, batch) low bits, see hv_split.hip).  Fields stay a
// run-time array, so disorder realisations share one compiled kernel.
std::string jit_source(const lsqed_ham* h, bool batch, bool cplx, std::vector<uint32_t>* lowtab) {
  const int N = h->N, nup = h->nup, L = split_low_bits(h, batch), NH = N - L;
  std::vector<unsigned> lowm, him;
  std::vector<int> lowk, lowt, crq, crp, crk, hik;
  for (int kind = 0; kind < 2; ++kind)
    for (size_t b = 0; b < h->bi[kind].size(); ++b) {
      const int a = h->bi[kind][b], c = h->bj[kind][b];
      const int lo = std::min(a, c), hi = std::max(a, c);
      if (lo == hi) continue;
      if (hi < L) {
        lowm.push_back((1u << lo) | (1u << hi));
        lowk.push_back(kind);
        lowt.push_back(hi < 8 ? 0 : (lo >= 8 ? 1 : 2));
      } else if (lo >= L) {
        him.push_back((1u << (lo - L)) | (1u << (hi - L)));
        hik.push_back(kind);
      } else {
        crq.push_back(lo);
        crp.push_back(hi - L);
        crk.push_back(kind);
      }
    }
  std::ostringstream o;
  o << "// generated by liblsqed (hv_jit.cpp)\n#define BATCH " << (batch ? 1 : 0) << "\n#define WIDE "
    << (N > 32 ? 1 : 0) << "\n";
  // Complex entries (16 B): the staged row doubles in LDS (two blocks per CU
  // at L = 14), so the kernel asks for two waves per SIMD and its registers.
  if (cplx) o << "#define CPLX 1\n";
  // occupancy hint of the kernel (waves per SIMD); LSQED_JIT_WAVES_PER_EU overrides the default 4
  if (const char* w = getenv("LSQED_JIT_WAVES_PER_EU")) {
    if (atoi(w) >= 1 && atoi(w) <= 8) o << "#define WAVES_PER_EU " << atoi(w) << "\n";
  } else if (cplx) {
    o << "#define WAVES_PER_EU 2\n";
  }
  // Gathers issued together: 16 for real single-model kernels with at least
  // ~55 global gathers per entry (cross bonds plus about half the high
  // bonds), else 8.  Results are bit-identical (the sums run in bond order
  // either way).  H.v at 8 / 16: 4x8 38.9 / 37.2 ms, 8x4 35.2 / 33.5, 4x7
  // 2.11 / 2.03, 2x12 0.132 / 0.128; 4x6 (49 gathers) 0.129 / 0.130; chain
  // N=26 0.199 / 0.227; the batched disorder solve 839 / 810 realisations/s.
  // With the cross gathers as buffer loads: 4x8 37.33 / 36.07 ms (12: 36.68),
  // 6x6 (64-bit ranks) 656.1 / 638.9 ms (12: 661.3; before the buffer loads
  // 686 / 691).  With the one-row kernel form (no spills at 16): 4x8 12 / 16 /
  // 20 / 24: 35.40 / 34.88-35.05 / 37.16 / 38.87 ms, 6x6 16 / 20: 595.6 /
  // 625.3 ms.  LSQED_JIT_GATHER_BATCH (2..32) overrides.
  int gather_batch = (!batch && !cplx && 2 * crq.size() + him.size() >= 110) ? 16 : 8;
  if (const char* w = getenv("LSQED_JIT_GATHER_BATCH"))
    if (atoi(w) >= 2 && atoi(w) <= 32) gather_batch = atoi(w);
  o << "#define GATHER_BATCH " << gather_batch << "\n";
  if (const char* w = getenv("LSQED_JIT_MASKED_GATHER"))
    if (atoi(w) == 1) o << "#define MASKED_GATHER 1\n";
  if (const char* w = getenv("LSQED_JIT_NT_STORE"))
    if (atoi(w) == 1) o << "#define NT_STORE 1\n";
  if (const char* w = getenv("LSQED_JIT_NT_STAGE"))
    if (atoi(w) == 1) o << "#define NT_STAGE 1\n";
  // Row-staged kernels: cross gathers as raw buffer loads whose
  // non-flipping lanes read nothing (BUF_CROSS), a batch's high-bond row bases
  // read together (HB_VEC) and, for single models, the high-bond loads as
  // buffer loads too (BUF_HIGH).  H.v with / without the three: 4x8
  // 36.18 / 36.97 ms, 6x6 660.9 / 677.7 ms, 4x6 0.130 / 0.137 ms, chain N=22
  // 24.3 / 24.3 us; the batched disorder solve 999 (without BUF_HIGH) / 986
  // realisations/s (985 with BUF_HIGH).  Results are bit-identical.  Two
  // alternatives measured slower or equal: non-flipping lanes reading entry 0 of the
  // target row with ordinary loads (4x8 36.58 ms, 6x6 695 ms), and the row
  // staged by direct global->LDS loads after the barrier instead of through
  // registers issued before it (36.91 ms: fewer spilled registers, but the
  // staging latency is no longer hidden).  LSQED_JIT_BUF_CROSS,
  // LSQED_JIT_HB_VEC and LSQED_JIT_BUF_HIGH (0/1) override.
  auto knob = [](const char* name, bool dflt) {
    const char* w = getenv(name);
    return w ? atoi(w) == 1 : dflt;
  };
  const bool buf_cross = knob("LSQED_JIT_BUF_CROSS", true), hb_vec = knob("LSQED_JIT_HB_VEC", true);
  if (buf_cross) o << "#define BUF_CROSS 1\n";
  if (hb_vec) o << "#define HB_VEC 1\n";
  if (buf_cross && hb_vec && knob("LSQED_JIT_BUF_HIGH", !batch)) o << "#define BUF_HIGH 1\n";
  // The next entry's pattern loaded one entry ahead (row-staged kernels;
  // LSQED_JIT_PREFETCH=0 turns it off): 4x8 H.v 37.13 -> 36.82-36.87 ms, the
  // batched disorder solve 938 -> 957 realisations/s (980 with LOWTAB records,
  // 959 with records alone).
  {
    const char* w = getenv("LSQED_JIT_PREFETCH");
    if (!(w && atoi(w) == 0)) o << "#define PREFETCH 1\n";
  }
  const RowParams rp = make_row_params(h);
  if (jit_rowstage()) {
    const uint64_t rmax = binom_u64(L, L / 2);
    o << "#define ROWSTAGE 1\n#define ROW_THREADS " << jit_row_threads(L) << "\nconstexpr int RMAX = " << rmax
      << ";\n";
    if (!batch && jit_rows_per_block() == 1) o << "#define ONE_ROW 1\n";
    if (const int mode = jit_lowtab(batch, L)) {
      // mode 1: every flipping low bond in the record; mode 2: only those
      // with a site in each byte of the pattern (lowt 2), the others computed
      std::vector<int> f, lowfk, lowrec(lowm.size(), -1);
      for (size_t b = 0; b < lowm.size(); ++b)
        if (rp.flip[lowk[b]] && (mode == 1 || lowt[b] == 2)) {
          lowrec[b] = (int)f.size();
          f.push_back((int)b);
          lowfk.push_back(lowk[b]);
        }
      const int nf = (int)f.size(), lword = (nf + 1) / 2, recq = (lword + 1 + 3) / 4;
      o << "#define LOWTAB " << mode << "\nconstexpr int NLOWF = " << nf << ", LWORD = " << lword
        << ", RECQ = " << recq << ";\n";
      emit_array(o, "LOWFK", lowfk, "int");
      emit_array(o, "LOWREC", lowrec, "int");
      if (lowtab) *lowtab = lowtab_records(L, lowm, lowk, f, recq, (unsigned)rmax);
    }
  }
  o << "constexpr int N = " << N << ", NUP = " << nup << ", L = " << L << ", NH = " << NH
    << ", W1 = " << (1 << std::max(L - 8, 0)) << ";\n";
  o << "constexpr int NB0 = " << h->bi[0].size() << ", NB1 = " << h->bi[1].size() << ";\n";
  // the batch variant reads couplings per problem at run time
  o << "constexpr double JH0 = " << dbl(batch ? 0.0 : rp.jh[0]) << ", JH1 = " << dbl(batch ? 0.0 : rp.jh[1])
    << ", DQ0 = " << dbl(batch ? 0.0 : rp.dq[0]) << ", DQ1 = " << dbl(batch ? 0.0 : rp.dq[1]) << ";\n";
  o << "constexpr int FLIP0 = " << rp.flip[0] << ", FLIP1 = " << rp.flip[1] << ", DZ0 = " << rp.dz_on[0]
    << ", DZ1 = " << rp.dz_on[1] << ", DIAG_MODE = " << rp.diag_mode << ";\n";
  o << "constexpr int NLOW = " << lowm.size() << ", NCROSS = " << crq.size() << ", NHIGH = " << him.size()
    << ";\n";
  emit_array(o, "LOWM", lowm, "unsigned");
  emit_array(o, "LOWK", lowk, "int");
  emit_array(o, "LOWT", lowt, "int");
  emit_array(o, "CRQ", crq, "int");
  emit_array(o, "CRP", crp, "int");
  emit_array(o, "CRK", crk, "int");
  emit_array(o, "HIM", him, "unsigned");
  emit_array(o, "HIK", hik, "int");
  o << kJitBody;
  return o.str();
}

static std::shared_ptr<JitModule> jit_compile(const std::string& src, bool batch) {
  auto m = std::make_shared<JitModule>();
  m->src_hash = fnv1a(src.data(), src.size());
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "lsqed_hv_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    m->log = "hiprtcCreateProgram failed";
    return m;
  }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"};
  hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  if (ls > 1) {
    std::string log(ls, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    m->log = log;
  }
  if (rc != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return m;
  }
  size_t sz = 0;
  hiprtcGetCodeSize(prog, &sz);
  std::vector<char> code(sz);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  if (hipModuleLoadData(&m->mod, code.data()) != hipSuccess) {
    m->log += "\nhipModuleLoadData failed";
    return m;
  }
  // kernel names: hv_jit_* (one model), hv_jit_c128_* (complex128 entries), hv_bjit_* (batch)
  const bool cplx = src.find("#define CPLX 1") != std::string::npos;
  const char* kdot = batch ? "hv_bjit_dot" : (cplx ? "hv_jit_c128_dot" : "hv_jit_dot");
  const char* knodot = batch ? "hv_bjit_nodot" : (cplx ? "hv_jit_c128_nodot" : "hv_jit_nodot");
  if (hipModuleGetFunction(&m->fn_dot, m->mod, kdot) != hipSuccess ||
      hipModuleGetFunction(&m->fn_nodot, m->mod, knodot) != hipSuccess) {
    m->log += "\nhipModuleGetFunction failed";
    return m;
  }
  m->ok = true;
  m->rowstage = src.find("#define ROWSTAGE 1") != std::string::npos;
  m->one_row = m->rowstage && src.find("#define ONE_ROW 1\n") != std::string::npos;
  const size_t rt = src.find("#define ROW_THREADS ");
  m->threads = m->rowstage && rt != std::string::npos ? atoi(src.c_str() + rt + 20) : 256;
  return m;
}



bool jit_enabled() {
  const char* e = getenv("LSQED_JIT");
  return !(e && e[0] == '0');
}

// The module of (h's device, generated source), compiled on first use; a
// LOWTAB module's records are built and uploaded once, with the module, to the
// current device (the handle's, as for the module itself).
static std::shared_ptr<JitModule> jit_module(const lsqed_ham* h, bool batch, bool cplx, std::string* key) {
  const std::string src = jit_source(h, batch, cplx, nullptr);
  if (key) *key = src;
  std::lock_guard<std::mutex> lk(g_jit_mu);
  const auto keyd = std::make_pair(h->device, src);
  auto it = g_jit_cache.find(keyd);
  if (it != g_jit_cache.end()) return it->second;
  std::shared_ptr<JitModule> m = jit_compile(src, batch);
  std::vector<uint32_t> tab;
  // (the kernel body's own fallback reads "#define LOWTAB 0")
  const bool lowtab = src.find("#define LOWTAB 1\n") != std::string::npos ||
                      src.find("#define LOWTAB 2\n") != std::string::npos;
  if (m->ok && lowtab) {
    (void)jit_source(h, batch, cplx, &tab);
    const size_t bytes = tab.size() * sizeof(uint32_t);
    if (tab.empty() || hipMalloc(&m->d_lowtab, bytes) != hipSuccess ||
        hipMemcpy(m->d_lowtab, tab.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
      m->ok = false;
      m->log += "\nupload of the low-bond records failed";
    }
  }
  g_jit_cache[keyd] = m;
  return m;
}

int jit_prepare(lsqed_ham* h) {
  if (h->jit) return LSQED_OK;
  std::shared_ptr<JitModule> m = jit_module(h, false, false, nullptr);
  if (!m->ok) return fail(LSQED_ERR_HIP, "hipRTC compile of the specialised H.v failed: " + m->log);
  h->jit = m.get();
  return LSQED_OK;
}

// Blocks of one specialised H.v (one per work item, or a persistent block per
// CU when the kernel stages rows); also the number of x.y partial sums it
// writes.
static int jit_grid_of(lsqed_ham* h, const JitModule* m) {
  SplitTables t;
  if (split_tables(h, t) != LSQED_OK) return 0;
  if (m->rowstage) {
    // one row per block (LSQED_JIT_ROWS_PER_BLOCK overrides).  With each
    // XCD's rows dealt longest first the dispatcher itself then schedules
    // longest-first, and the launch ends on the shortest rows: 4x8 34.73-
    // 34.84 ms against 35.67-35.72 with two rows per block, 6x6 625.1 against
    // 636.2 ms, 8x4 29.33 against 30.01 ms.  In h order two rows per block had
    // been better (a block's once-per-block work -- its tables and the stores
    // of the 12 VGPRs the looped kernel spills at four waves per SIMD --
    // shared by two rows; the one-row form spills nothing: 4x8 34.75 -> 34.20
    // ms, 6x6 625 -> 589 ms, writes per 4x8 product 8.2 -> 5.0 GB): rows per
    // block on 4x8 / 6x6 1: 39.5 / 701.7 ms, 2: 39.1 /
    // 688.8 ms, 4: 38.6 / 714.1 ms, 8: 40.8 ms on 4x8 at the time.  A
    // persistent grid (one block per CU dealing rows cyclically) measured 7%
    // slower: a fixed share of rows per CU is unevenly long.
    const int k = m->one_row ? 1 : jit_rows_per_block();
    const int per_xcd = t.nrows_ord / 8;
    return 8 * std::max(1, (per_xcd + k - 1) / k);
  }
  return std::max(8, ((t.ntasks_ord + 7) / 8) * 8);
}

int jit_grid(lsqed_ham* h) {
  if (jit_prepare(h) != LSQED_OK) return 0;
  return jit_grid_of(h, (const JitModule*)h->jit);
}

uint64_t jit_source_hash(const lsqed_ham* h) { return h->jit ? ((const JitModule*)h->jit)->src_hash : 0; }
int jit_threads(const lsqed_ham* h) { return h->jit ? ((const JitModule*)h->jit)->threads : 0; }

// The complex128 module of the handle (same bond structure and couplings).
static int jit_prepare_c128(lsqed_ham* h) {
  if (h->jit_c) return LSQED_OK;
  std::shared_ptr<JitModule> m = jit_module(h, false, true, nullptr);
  if (!m->ok) return fail(LSQED_ERR_HIP, "hipRTC compile of the complex specialised H.v failed: " + m->log);
  h->jit_c = m.get();
  return LSQED_OK;
}

int launch_hv_jit_c128(lsqed_ham* h, const double* x, double* y, hipStream_t s) {
  int rc = jit_prepare_c128(h);
  if (rc) return rc;
  SplitTables t;
  if ((rc = split_tables(h, t))) return rc;
  JitModule* m = (JitModule*)h->jit_c;
  const int grid = jit_grid_of(h, m);
  if (grid <= 0) return fail(LSQED_ERR_HIP, "row-split plan of the complex H.v failed");
  JitArgs a;
  std::memset(&a, 0, sizeof(a));
  a.x = x;
  a.y = y;
  a.dot = nullptr;
  a.fields = h->d_fields;
  a.U = t.U;
  a.RL0 = t.RL0;
  a.RL1 = t.RL1;
  a.binom = t.binom;
  a.binom64 = t.binom64;
  a.tasks = m->rowstage ? t.rows_ord : t.tasks_ord;
  a.ntasks = m->rowstage ? t.nrows_ord : t.ntasks_ord;
  for (int k = 0; k < 17; ++k) a.offU[k] = t.offU[k];
  a.lowtab = m->d_lowtab;
  void* args[] = {&a};
  LSQED_HIP(hipModuleLaunchKernel(m->fn_nodot, (unsigned)grid, 1, 1, m->threads, 1, 1, 0, s, args, nullptr));
  return LSQED_OK;
}

int launch_hv_jit(lsqed_ham* h, const double* x, double* y, double* dot_part, int grid, hipStream_t s) {
  int rc = jit_prepare(h);
  if (rc) return rc;
  SplitTables t;
  if ((rc = split_tables(h, t))) return rc;
  JitModule* m = (JitModule*)h->jit;
  JitArgs a;
  std::memset(&a, 0, sizeof(a));
  a.x = x;
  a.y = y;
  a.dot = dot_part;
  a.fields = h->d_fields;
  a.U = t.U;
  a.RL0 = t.RL0;
  a.RL1 = t.RL1;
  a.binom = t.binom;
  a.binom64 = t.binom64;
  a.tasks = m->rowstage ? t.rows_ord : t.tasks_ord;
  a.ntasks = m->rowstage ? t.nrows_ord : t.ntasks_ord;
  for (int k = 0; k < 17; ++k) a.offU[k] = t.offU[k];
  a.lowtab = m->d_lowtab;
  void* args[] = {&a};
  LSQED_HIP(hipModuleLaunchKernel(dot_part ? m->fn_dot : m->fn_nodot, (unsigned)grid, 1, 1, m->threads,
                                  1, 1, 0, s, args, nullptr));
  return LSQED_OK;
}

// Batch variant: one module per bond structure and build mode (couplings and
// fields per problem at run time).  *key receives the generated source, which
// identifies the group of problems that can share the kernel.
int batch_jit_prepare(lsqed_ham* h, void** module, std::string* key) {
  std::shared_ptr<JitModule> m = jit_module(h, true, false, key);
  if (!m->ok) return fail(LSQED_ERR_HIP, "hipRTC compile of the batched H.v failed: " + m->log);
  *module = m.get();
  return LSQED_OK;
}

// Tasks per problem of a batch module: rows when it stages rows, else
// 2048-entry work items.
int batch_task_count(void* module, lsqed_ham* h, int* ntasks) {
  SplitTables t;
  const int rc = split_tables(h, t, true);
  if (rc) return rc;
  *ntasks = ((JitModule*)module)->rowstage ? t.nrows : t.ntasks;
  return LSQED_OK;
}

int launch_hv_batch_jit(void* module, lsqed_ham* h0, const BatchJitLaunch& L, int grid, hipStream_t s) {
  SplitTables t;
  int rc = split_tables(h0, t, true);
  if (rc) return rc;
  JitModule* m = (JitModule*)module;
  BatchJitArgs a;
  std::memset(&a, 0, sizeof(a));
  a.base.x = L.x;
  a.base.y = L.y;
  a.base.dot = L.dot;
  a.base.U = t.U;
  a.base.RL0 = t.RL0;
  a.base.RL1 = t.RL1;
  a.base.binom = t.binom;
  a.base.binom64 = t.binom64;
  a.base.tasks = m->rowstage ? t.rows : t.tasks;
  a.base.ntasks = m->rowstage ? t.nrows : t.ntasks;
  for (int k = 0; k < 17; ++k) a.base.offU[k] = t.offU[k];
  a.base.lowtab = m->d_lowtab;
  a.poff = L.poff;
  a.pcoup = L.pcoup;
  a.pfields = L.pfields;
  a.active = L.active;
  a.gidx = L.gidx;
  a.nprob = L.nprob;
  void* args[] = {&a};
  LSQED_HIP(hipModuleLaunchKernel(L.dot ? m->fn_dot : m->fn_nodot, (unsigned)grid, 1, 1, m->threads, 1, 1, 0, s,
                                  args, nullptr));
  return LSQED_OK;
}

}  // namespace lsqed
