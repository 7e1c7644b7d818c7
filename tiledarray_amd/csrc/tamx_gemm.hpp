// The persistent grouped GEMM kernel (FP64 and FP32 MFMA), shared by the
// plain entry points (tamx_gemm.hip) and the gathered ones
// (tamx_gemm_gather.hip).
//
// One launch computes every result tile of a contraction.  A result tile of
// m x n is cut into BM x BN macro tiles ("items"); a workgroup takes items
// from a device queue, and for its item walks all contributing tile pairs in
// order, K in chunks of BK, with A and B chunks staged through LDS (double
// buffered, one barrier per chunk) and accumulated in MFMA registers.  The
// result is written once: alpha*acc, plus the old C when the task asks for
// beta = 1.  Output-stationary accumulation keeps every result bit-for-bit
// reproducible (the pair order is the work-list order) and needs no atomics
// on C.
//
// MFMA fragment layout of v_mfma_{f64,f32}_16x16x4: lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16].  The 16x16 result is held in
// column l % 16; accumulator component v is row 4 * (l / 16) + v for f32 and
// row (l / 16) + 4 * v for f64.
#pragma once

#include <algorithm>
#include <mutex>
#include <unordered_map>

#include "tamx_internal.hpp"
#include "tamx_map.hpp"

namespace {

using tamx::gather_offset;
using tamx::load_modes;

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = d4_t;
  __device__ static inline acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Mfma<float> {
  using acc_t = f4_t;
  __device__ static inline acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int v) { return 4 * lk + v; }
};

template <typename T>
struct GemmParams {
  const T* __restrict__ A;
  const T* __restrict__ B;
  T* __restrict__ C;
  const tamx_gemm_task* __restrict__ tasks;
  const tamx_gemm_pair* __restrict__ pairs;
  const int64_t* __restrict__ counters;
  const int32_t* __restrict__ prefix;
  const int32_t* __restrict__ item_task;
  int32_t* queue;
  int64_t task_capacity;
  int64_t item_capacity;
  T alpha;
  // gathered operands (null: plain row-major tiles with the pair's ld)
  const tamx::GatherTile* __restrict__ a_gather;
  const tamx::GatherTile* __restrict__ b_gather;
  // result tiles stored through a scatter map (null: row-major with the
  // task's ldc); task_tile[t] is task t's position in the result grid
  const tamx::GatherTile* __restrict__ c_scatter;
  const int32_t* __restrict__ task_tile;
};

constexpr int kThreads = 256;

// Buffer loads of plain operands: descriptor word 3 for raw 32-bit-format
// buffers on this family, the offset that lies outside every range, and the
// largest leading dimension whose per-element offsets stay below it.
constexpr int kBufFlags = 0x00020000;
constexpr uint32_t kOutside = 0x80000000u;
constexpr int kBufLd = 1 << 20;

template <typename T>
__device__ inline T buffer_load(__amdgpu_buffer_rsrc_t d, uint32_t off);
template <>
__device__ inline double buffer_load<double>(__amdgpu_buffer_rsrc_t d, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(d, off, 0, 0));
}
template <>
__device__ inline float buffer_load<float>(__amdgpu_buffer_rsrc_t d, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(d, off, 0, 0));
}
constexpr int kRegions = 8;   // XCDs of an MI355X; the queue holds kRegions + 1 counters
constexpr int kPairWin = 64;  // pair descriptors staged in LDS at a time

// the address out-of-tile lanes load from
__device__ const double kZeroF64 = 0.0;
__device__ const float kZeroF32 = 0.0f;
template <typename T>
__device__ inline const T* zero_of();
template <>
__device__ inline const double* zero_of<double>() { return &kZeroF64; }
template <>
__device__ inline const float* zero_of<float>() { return &kZeroF32; }

__device__ inline int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ inline int64_t uni64(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// GM: which operands the launch gathers (bit 0 A, bit 1 B; 0 for the plain
// kernel).  A gathered operand's tile t is read through p.a_gather[t] (or
// p.b_gather[t]), element (row, k) at the pair's tile offset + rowoff(row)
// + koff(k), instead of through the pair's leading dimension.  TA / TB then
// pick only which index the staging lanes run along (A: NoTrans k, Trans m;
// B: NoTrans n, Trans k).
template <typename T, int BM, int BN, int BK, bool TA, bool TB, int GM>
__global__ __launch_bounds__(kThreads, (sizeof(T) == 8 && BM == 64 && BN == 64 && BK == 16) ? 4 : 1) void grouped_gemm_kernel(GemmParams<T> p) {
  constexpr bool GA = (GM & 1) != 0, GB = (GM & 2) != 0, G = GM != 0;
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr int FM = BM / 32, FN = BN / 32;  // 16x16 blocks per wave (2x2 waves)
  constexpr int A_LD = BM * BK / kThreads, B_LD = BK * BN / kThreads;
  constexpr int SA = BM + 1, SB = BN + 1;  // padded LDS rows
  static_assert(BM % 32 == 0 && BN % 32 == 0 && BK % 4 == 0, "tile shape");
  static_assert((BM * BK) % kThreads == 0 && (BK * BN) % kThreads == 0, "load split");
  // per-thread coordinates of the elements it stages (fixed for the kernel)
  // A NoTrans: k fastest (rows of op(A) contiguous along k); A Trans: m fastest
  constexpr int A_STEP = TA ? kThreads / BM : kThreads / BK;  // row (or k) step per e
  constexpr int B_STEP = TB ? kThreads / BK : kThreads / BN;
  static_assert(TA ? kThreads % BM == 0 : kThreads % BK == 0, "A staging layout");
  static_assert(TB ? kThreads % BK == 0 : kThreads % BN == 0, "B staging layout");

  __shared__ T sA[2][BK][SA];
  __shared__ T sB[2][BK][SB];
  __shared__ tamx_gemm_pair s_pairs[kPairWin];
  __shared__ int s_item;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = uni(tid >> 6);
  const int wr = (wave >> 1) * (BM / 2), wc = (wave & 1) * (BN / 2);
  const int lrow = lane & 15, lk = lane >> 4;
  // A staging: element e is (row a_r + e*A_STEP or a_r, k a_k or a_k + e*A_STEP)
  const int a_fast = TA ? tid % BM : tid % BK;  // m (Trans) or k (NoTrans)
  const int a_slow = TA ? tid / BM : tid / BK;  // k (Trans) or m (NoTrans)
  const int b_fast = TB ? tid % BK : tid % BN;  // k (Trans) or n (NoTrans)
  const int b_slow = TB ? tid / BK : tid / BN;  // n (Trans) or k (NoTrans)

  int64_t ntasks64 = p.counters[TAMX_WL_NTASKS];
  // same validation as the planner: an invalid list has no items
  if (ntasks64 < 0 || ntasks64 > p.task_capacity || p.counters[TAMX_WL_STATUS] != 0)
    ntasks64 = 0;
  const int32_t ntasks = uni((int32_t)ntasks64);
  const int32_t total = uni(p.prefix[ntasks > 0 ? ntasks : 0]);

  // Items are split into kRegions contiguous ranges, one per XCD: a
  // workgroup drains the range of its own XCD (workgroups are dispatched to
  // XCDs round-robin by blockIdx), so the macro tiles of one result tile,
  // which share A and B panels, run on one L2; then it steals from the
  // other ranges in turn.
  const int home = blockIdx.x % kRegions;
  int tried = 0;  // used by thread 0 only
  for (;;) {
    if (tid == 0) {
      int it = -1;
      while (tried < kRegions) {
        const int r = (home + tried) % kRegions;
        const int rb = (int)((int64_t)total * r / kRegions);
        const int re = (int)((int64_t)total * (r + 1) / kRegions);
        const int q = rb < re ? atomicAdd(&p.queue[r], 1) : 0;
        if (rb + q < re) {
          it = rb + q;
          break;
        }
        ++tried;
      }
      s_item = it;
    }
    __syncthreads();
    const int item = uni(s_item);
    __syncthreads();
    if (item < 0) break;

    int lo;
    if (total <= p.item_capacity) {
      lo = uni(p.item_task[item]);
    } else {
      lo = 0;
      int hi = ntasks - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (uni(p.prefix[mid]) <= item) lo = mid; else hi = mid - 1;
      }
    }
    const tamx_gemm_task* tp = p.tasks + lo;
    const int64_t t_coff = uni64(tp->c_off);
    const int t_m = uni(tp->m), t_n = uni(tp->n), t_ldc = uni(tp->ldc);
    const int t_pb = uni(tp->pair_begin), t_pc = uni(tp->pair_count);
    const int t_flags = uni(tp->flags);
    const int local = item - uni(p.prefix[lo]);
    const int tiles_n = (t_n + BN - 1) / BN;
    const int r0 = (local / tiles_n) * BM, c0 = (local % tiles_n) * BN;
    const int mrem = min(BM, t_m - r0), nrem = min(BN, t_n - c0);
    // a wave whose quadrant lies wholly outside a ragged edge skips its
    // MFMAs, and inside a quadrant every 16 x 16 block that does (the
    // conditions are uniform across the wave)
    const bool wave_active = (wr < mrem) && (wc < nrem);
    // 16-row (column) blocks of the quadrant inside the tile, as scalars
    const int ni = uni(min(FM, max(0, (mrem - wr + 15) >> 4)));
    const int nj = uni(min(FN, max(0, (nrem - wc + 15) >> 4)));

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[i][j][v] = T(0);

    // Per-pair staging state of the load cursor: the thread's element
    // offsets (relative to the pair's tile) at k0 = 0, and whether its
    // row/column is inside the tile.
    int64_t a_base = 0, b_base = 0;
    int a_lda = 0, b_ldb = 0, pk = 0;
    int64_t a_off[A_LD], b_off[B_LD];
    bool a_ok[A_LD], b_ok[B_LD];
    // Plain operands are read with buffer loads: per chunk one descriptor
    // per operand based at the macro tile's first row (column) and the
    // chunk's first k, and per element a 32-bit byte offset from it that is
    // fixed for the pair.  An element outside the tile carries kOutside,
    // beyond every descriptor's range, and loads zero; so do the k rows past
    // the pair's end where k is the slow index (the range stops there).
    // Pairs with a leading dimension above kBufLd take the pointer path.
    uint32_t a_vo[A_LD], b_vo[B_LD];
    bool buf_ok = false;
    // k-mode maps of the current pair's gathered tiles
    tamx::GatherMode ga_k[4], gb_k[4];
    int at = 0, bt = 0;
    // pair descriptors of the task: in LDS when they fit, else read from
    // global memory
    const int pend = t_pb + t_pc;
    const bool pairs_in_lds = t_pc <= kPairWin;
    if (pairs_in_lds && t_pc > 0) {
      const int32_t* src = reinterpret_cast<const int32_t*>(p.pairs + t_pb);
      int32_t* dst = reinterpret_cast<int32_t*>(s_pairs);
      constexpr int W = sizeof(tamx_gemm_pair) / sizeof(int32_t);
      for (int x = tid; x < t_pc * W; x += kThreads) dst[x] = src[x];
    }
    __syncthreads();
    auto set_pair = [&](int pi) {
      if (pairs_in_lds) {
        const tamx_gemm_pair& pp = s_pairs[pi - t_pb];
        a_base = uni64(pp.a_off);
        b_base = uni64(pp.b_off);
        pk = uni(pp.k);
        a_lda = uni(pp.lda);
        b_ldb = uni(pp.ldb);
        if constexpr (G) {
          at = uni(pp.a_tile);
          bt = uni(pp.b_tile);
        }
      } else {
        const auto& pp = tamx::scalar_ptr(p.pairs)[pi];
        a_base = uni64(pp.a_off);
        b_base = uni64(pp.b_off);
        pk = uni(pp.k);
        a_lda = uni(pp.lda);
        b_ldb = uni(pp.ldb);
        if constexpr (G) {
          at = uni(pp.a_tile);
          bt = uni(pp.b_tile);
        }
      }
      if constexpr (!G) {
        buf_ok = a_lda <= kBufLd && b_ldb <= kBufLd;
#pragma unroll
        for (int e = 0; e < A_LD; ++e) {
          const int slow = a_slow + e * A_STEP;
          const bool in = TA ? a_fast < mrem : slow < mrem;
          a_vo[e] = in ? (uint32_t)(slow * a_lda + a_fast) * (uint32_t)sizeof(T) : kOutside;
        }
#pragma unroll
        for (int e = 0; e < B_LD; ++e) {
          const int slow = b_slow + e * B_STEP;
          const bool in = TB ? slow < nrem : b_fast < nrem;
          b_vo[e] = in ? (uint32_t)(slow * b_ldb + b_fast) * (uint32_t)sizeof(T) : kOutside;
        }
      } else {
#pragma unroll
        for (int e = 0; e < A_LD; ++e) {
          const int slow = a_slow + e * A_STEP;
          if (TA) {  // (k = slow, m = a_fast)
            a_ok[e] = a_fast < mrem;
            a_off[e] = (int64_t)slow * a_lda + r0 + a_fast;
          } else {  // (m = slow, k = a_fast)
            a_ok[e] = slow < mrem;
            a_off[e] = (int64_t)(r0 + slow) * a_lda + a_fast;
          }
        }
#pragma unroll
        for (int e = 0; e < B_LD; ++e) {
          const int slow = b_slow + e * B_STEP;
          if (TB) {  // (n = slow, k = b_fast)
            b_ok[e] = slow < nrem;
            b_off[e] = (int64_t)(c0 + slow) * b_ldb + b_fast;
          } else {  // (k = slow, n = b_fast)
            b_ok[e] = b_fast < nrem;
            b_off[e] = (int64_t)slow * b_ldb + c0 + b_fast;
          }
        }
      }
      // gathered operands: row offsets of the thread's rows through the
      // tile's row modes (the maps are read with scalar loads)
      if constexpr (GA) {
        const auto& g = tamx::scalar_ptr(p.a_gather)[at];
        tamx::GatherMode rm[4];
        load_modes(g.r, rm);
        load_modes(g.k, ga_k);
#pragma unroll
        for (int e = 0; e < A_LD; ++e) {
          const int row = r0 + (TA ? a_fast : a_slow + e * A_STEP);
          a_off[e] = a_ok[e] ? gather_offset(row, rm) : 0;
        }
      }
      if constexpr (GB) {
        const auto& g = tamx::scalar_ptr(p.b_gather)[bt];
        tamx::GatherMode rm[4];
        load_modes(g.r, rm);
        load_modes(g.k, gb_k);
#pragma unroll
        for (int e = 0; e < B_LD; ++e) {
          const int col = c0 + (TB ? b_slow + e * B_STEP : b_fast);
          b_off[e] = b_ok[e] ? gather_offset(col, rm) : 0;
        }
      }
    };

    // load cursor (pair lp, offset lk0) runs two chunks ahead of compute
    int lp = t_pb, lk0 = 0;
    if (lp < pend) set_pair(lp);
    auto advance = [&]() {
      lk0 += BK;
      if (lk0 >= pk) {
        ++lp;
        lk0 = 0;
        if (lp < pend) set_pair(lp);
      }
    };
    // loads of chunk lk0 of the cursor's pair; out-of-tile elements read a
    // zero constant, so the loaded registers need no select and nothing
    // waits for the loads before they are stored to LDS.  Once the cursor
    // has passed the last pair the loads repeat its first chunk, unused:
    // every step issues the same loads, which keeps the wait counts exact.
    // returns the number of 4-k MFMA steps the chunk holds (the rest of a
    // pair's last chunk is zero padding, skipped)
    auto load_chunk = [&](T (&ra)[A_LD], T (&rb)[B_LD]) -> int {
      const int kc = min(BK, pk - lk0);
      if constexpr (!G) {
        if (buf_ok) {
          // chunk origins (elements) and ranges (bytes): k-slow operands
          // end at the pair's last k row, k-fast ones rely on the marks
          const int64_t a_org = uni64(a_base + (TA ? (int64_t)lk0 * a_lda + r0
                                                    : (int64_t)r0 * a_lda + lk0));
          const int64_t b_org = uni64(b_base + (TB ? (int64_t)c0 * b_ldb + lk0
                                                    : (int64_t)lk0 * b_ldb + c0));
          const int a_rng = TA ? uni(kc * a_lda * (int)sizeof(T)) : (int)(kOutside - 1);
          const int b_rng = TB ? (int)(kOutside - 1) : uni(kc * b_ldb * (int)sizeof(T));
          const auto ra_d = __builtin_amdgcn_make_buffer_rsrc((void*)(p.A + a_org), (short)0,
                                                              a_rng, kBufFlags);
          const auto rb_d = __builtin_amdgcn_make_buffer_rsrc((void*)(p.B + b_org), (short)0,
                                                              b_rng, kBufFlags);
          if (kc == BK || (TA && !TB)) {
#pragma unroll
            for (int e = 0; e < A_LD; ++e) ra[e] = buffer_load<T>(ra_d, a_vo[e]);
#pragma unroll
            for (int e = 0; e < B_LD; ++e) rb[e] = buffer_load<T>(rb_d, b_vo[e]);
          } else {
            // a pair's last chunk: k-fast operands mask the k past its end
            const bool a_kok = TA || a_fast < kc, b_kok = !TB || b_fast < kc;
#pragma unroll
            for (int e = 0; e < A_LD; ++e) ra[e] = buffer_load<T>(ra_d, a_kok ? a_vo[e] : kOutside);
#pragma unroll
            for (int e = 0; e < B_LD; ++e) rb[e] = buffer_load<T>(rb_d, b_kok ? b_vo[e] : kOutside);
          }
        } else {
          // 64-bit addresses, recomputed per chunk (no state kept per pair)
          const T* __restrict__ a = p.A + a_base;
          const T* __restrict__ b = p.B + b_base;
#pragma unroll
          for (int e = 0; e < A_LD; ++e) {
            const int slow = a_slow + e * A_STEP;
            const bool ok = TA ? (a_fast < mrem && slow < kc) : (slow < mrem && a_fast < kc);
            const int64_t off = TA ? (int64_t)(lk0 + slow) * a_lda + r0 + a_fast
                                   : (int64_t)(r0 + slow) * a_lda + lk0 + a_fast;
            ra[e] = *(ok ? tamx::global_ptr(a + off) : tamx::global_ptr(zero_of<T>()));
          }
#pragma unroll
          for (int e = 0; e < B_LD; ++e) {
            const int slow = b_slow + e * B_STEP;
            const bool ok = TB ? (slow < nrem && b_fast < kc) : (b_fast < nrem && slow < kc);
            const int64_t off = TB ? (int64_t)(c0 + slow) * b_ldb + lk0 + b_fast
                                   : (int64_t)(lk0 + slow) * b_ldb + c0 + b_fast;
            rb[e] = *(ok ? tamx::global_ptr(b + off) : tamx::global_ptr(zero_of<T>()));
          }
        }
        if (lp < pend) advance();
        return uni((kc + 3) >> 2);
      }
      const T* __restrict__ a = p.A + a_base;
      const T* __restrict__ b = p.B + b_base;
      const int64_t a_k = TA ? (int64_t)lk0 * a_lda : lk0;
      const int64_t b_k = TB ? lk0 : (int64_t)lk0 * b_ldb;
      // gathered operands: one k offset when the lanes run along k
      int32_t ga_ko = 0, gb_ko = 0;
      if constexpr (GA && !TA) ga_ko = gather_offset(lk0 + a_fast, ga_k);
      if constexpr (GB && TB) gb_ko = gather_offset(lk0 + b_fast, gb_k);

#pragma unroll
      for (int e = 0; e < A_LD; ++e) {
        const int kk = TA ? a_slow + e * A_STEP : a_fast;
        const bool ok = a_ok[e] && kk < kc;
        int64_t ko = a_k;
        if constexpr (GA) ko = TA ? gather_offset(lk0 + kk, ga_k) : ga_ko;
        ra[e] = *(ok ? tamx::global_ptr(a + a_off[e] + ko) : tamx::global_ptr(zero_of<T>()));
      }
#pragma unroll
      for (int e = 0; e < B_LD; ++e) {
        const int kk = TB ? b_fast : b_slow + e * B_STEP;
        const bool ok = b_ok[e] && kk < kc;
        int64_t ko = b_k;
        if constexpr (GB) ko = TB ? gb_ko : gather_offset(lk0 + kk, gb_k);
        rb[e] = *(ok ? tamx::global_ptr(b + b_off[e] + ko) : tamx::global_ptr(zero_of<T>()));
      }
      if (lp < pend) advance();
      return uni((kc + 3) >> 2);
    };
    // stores part `part` of NP parts of a staged chunk (all of it: part 0 of 1)
    auto store_part = [&](int buf, const T (&ra)[A_LD], const T (&rb)[B_LD], int part, int NP) {
#pragma unroll
      for (int e = 0; e < A_LD; ++e) {
        if (e * NP / A_LD != part) continue;
        if (TA) sA[buf][a_slow + e * A_STEP][a_fast] = ra[e];
        else sA[buf][a_fast][a_slow + e * A_STEP] = ra[e];
      }
#pragma unroll
      for (int e = 0; e < B_LD; ++e) {
        if (e * NP / B_LD != part) continue;
        if (TB) sB[buf][b_fast][b_slow + e * B_STEP] = rb[e];
        else sB[buf][b_slow + e * B_STEP][b_fast] = rb[e];
      }
    };
    auto store_chunk = [&](int buf, const T (&ra)[A_LD], const T (&rb)[B_LD]) {
      store_part(buf, ra, rb, 0, 1);
    };
    // nks 4-k steps of LDS buffer `buf`; the fragments of step s + 1 are
    // read while the MFMAs of step s run (the reads of blocks past a ragged
    // edge are harmless: they stay inside the buffer)
    // `stage(st)` runs after the MFMAs of step st (every step, also those
    // skipped): the caller stages the next chunk into the other buffer there,
    // so the wave's LDS stores overlap its own MFMAs.
    auto compute_chunk = [&](int buf, int nks, auto&& stage) {
      if (!wave_active) {
#pragma unroll
        for (int st = 0; st < BK / 4; ++st) stage(st);
        return;
      }
      T af[2][FM], bf[2][FN];
      auto frag = [&](int ks, T (&a)[FM], T (&b)[FN]) {
#pragma unroll
        for (int i = 0; i < FM; ++i) a[i] = sA[buf][ks + lk][wr + i * 16 + lrow];
#pragma unroll
        for (int j = 0; j < FN; ++j) b[j] = sB[buf][ks + lk][wc + j * 16 + lrow];
      };
      frag(0, af[0], bf[0]);
#pragma unroll
      for (int st = 0; st < BK / 4; ++st) {
        if (st < nks) {
          if (st + 1 < BK / 4) frag(4 * (st + 1), af[(st + 1) & 1], bf[(st + 1) & 1]);
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
              if (i < ni && j < nj) acc[i][j] = M::op(af[st & 1][i], bf[st & 1][j], acc[i][j]);
        }
        stage(st);
      }
    };

    // Pipeline: registers R0/R1 hold chunks c+1 and c+2 in flight while
    // chunk c is multiplied out of LDS buffer c % 2.  One barrier per chunk.
    // (n0, n1: 4-k steps of the chunks in R0 / R1)
    T ra0[A_LD], rb0[B_LD], ra1[A_LD], rb1[B_LD];
    if (lp < pend) {
      int n0 = load_chunk(ra0, rb0);           // chunk 0
      const bool have1 = lp < pend;
      int n1 = load_chunk(ra1, rb1);           // chunk 1 (if have1)
      store_chunk(0, ra0, rb0);
      int nb = n0;  // steps of the chunk in the buffer being computed
      __syncthreads();
      bool more = have1;  // chunk c + 1 exists
      // (the stores of chunk c + 1 run between the MFMA steps of chunk c;
      // past the last chunk they store the unused repeat, harmlessly)
      constexpr int NS = BK / 4;
      for (;;) {
        // even step: compute buffer 0, chunk c+1 is in R1
        bool next2 = lp < pend;
        n0 = load_chunk(ra0, rb0);             // chunk c + 2 -> R0 (if next2)
        compute_chunk(0, nb, [&](int st) { store_part(1, ra1, rb1, st, NS); });
        if (!more) break;
        nb = n1;
        __syncthreads();
        more = next2;
        // odd step: compute buffer 1, chunk c+1 is in R0
        next2 = lp < pend;
        n1 = load_chunk(ra1, rb1);             // chunk c + 2 -> R1 (if next2)
        compute_chunk(1, nb, [&](int st) { store_part(0, ra0, rb0, st, NS); });
        if (!more) break;
        nb = n0;
        __syncthreads();
        more = next2;
      }
    }

    const bool beta1 = (t_flags & TAMX_TASK_BETA1) != 0;
    T* __restrict__ c = p.C + t_coff;
    tamx::with_result_index(p.c_scatter, p.task_tile, lo, t_flags, t_ldc, [&](auto cidx) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = wr + i * 16 + M::row(lk, v);
            const int col = wc + j * 16 + lrow;
            if (row < mrem && col < nrem) {
              const int64_t idx = cidx(r0 + row, c0 + col);
              T val = p.alpha * acc[i][j][v];
              if (beta1) val += c[idx];
              c[idx] = val;
            }
          }
    });
  }

  if (tid == 0) {
    __threadfence();
    const int done = atomicAdd(&p.queue[kRegions], 1);
    if (done == (int)gridDim.x - 1) {
      // last workgroup out: every grab has happened, reset for the next launch
      for (int r = 0; r <= kRegions; ++r) atomicExch(&p.queue[r], 0);
    }
  }
}

struct TileCfg {
  int bm, bn;
};

template <typename T, int BM, int BN, int BK, bool TA, bool TB, int GM>
int launch_cfg(tamx_context* ctx, tamx_worklist* wl, const GemmParams<T>& prm,
               hipStream_t s, int wgs_per_cu) {
  auto kern = grouped_gemm_kernel<T, BM, BN, BK, TA, TB, GM>;
  static std::mutex mu;
  static std::unordered_map<int, int> occ_cache;
  int occ = 0;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = occ_cache.find(ctx->device);
    if (it == occ_cache.end()) {
      int o = 0;
      TAMX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, kern, kThreads, 0));
      occ_cache[ctx->device] = o > 0 ? o : 1;
      occ = occ_cache[ctx->device];
    } else {
      occ = it->second;
    }
  }
  const int grid = ctx->num_cus * (wgs_per_cu > 0 ? std::min(wgs_per_cu, occ) : occ);
  if (int rc = tamx::plan_items(wl, BM, BN, s)) return rc;
  kern<<<grid, kThreads, 0, s>>>(prm);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

template <typename T, int BM, int BN, int BK, int GM>
int launch_ops(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
               const GemmParams<T>& prm, hipStream_t s, int wgs) {
  if (!op_a && !op_b) return launch_cfg<T, BM, BN, BK, false, false, GM>(ctx, wl, prm, s, wgs);
  if (!op_a && op_b) return launch_cfg<T, BM, BN, BK, false, true, GM>(ctx, wl, prm, s, wgs);
  if (op_a && !op_b) return launch_cfg<T, BM, BN, BK, true, false, GM>(ctx, wl, prm, s, wgs);
  return launch_cfg<T, BM, BN, BK, true, true, GM>(ctx, wl, prm, s, wgs);
}

// Macro-tile shapes compiled into the library (BM, BN, BK).
constexpr int kConfigs[][3] = {{64, 64, 16}, {64, 64, 32}, {128, 128, 16},
                               {128, 64, 16}, {64, 128, 16}, {32, 32, 16},
                               {32, 32, 32}, {128, 64, 32}, {64, 128, 32}};

// One launch of the grouped GEMM with the configuration `cfg` (null: the
// default 64 x 64 x 16).  GM: which operands are gathered (0: none).  The
// gathered kernels are compiled for the three macro tiles that suit large
// tiles (64 x 64, 128 x 64 and 64 x 128, BK 16).
template <typename T, int GM>
int grouped_gemm(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
                 T alpha, const T* A, const T* B, T* C, void* stream,
                 const tamx_gemm_config* cfg, const tamx::GatherTile* a_gather = nullptr,
                 const tamx::GatherTile* b_gather = nullptr,
                 const tamx::GatherTile* c_scatter = nullptr) {
  TAMX_REQUIRE(wl != nullptr, "null work list");
  TAMX_REQUIRE(op_a == 0 || op_a == 1, "op_a must be TAMX_NO_TRANS or TAMX_TRANS");
  TAMX_REQUIRE(op_b == 0 || op_b == 1, "op_b must be TAMX_NO_TRANS or TAMX_TRANS");
  if (int rc = tamx::activate(ctx)) return rc;
  GemmParams<T> prm{A, B, C, wl->tasks, wl->pairs, wl->counters,
                    wl->item_prefix, wl->item_task, wl->queue,
                    wl->task_capacity, wl->item_capacity, alpha, a_gather, b_gather,
                    c_scatter, wl->task_tile};
  hipStream_t s = tamx::as_stream(stream);
  const int bm = cfg ? cfg->bm : 64, bn = cfg ? cfg->bn : 64, bk = cfg ? cfg->bk : 16;
  const int w = cfg ? cfg->wgs_per_cu : 0;
#define TAMX_CFG(BM_, BN_, BK_) \
  if (bm == BM_ && bn == BN_ && bk == BK_) return launch_ops<T, BM_, BN_, BK_, GM>(ctx, wl, op_a, op_b, prm, s, w);
  TAMX_CFG(64, 64, 16)
  TAMX_CFG(128, 64, 16)
  TAMX_CFG(64, 128, 16)
  if constexpr (GM == 0) {
    TAMX_CFG(64, 64, 32)
    TAMX_CFG(128, 128, 16)
    TAMX_CFG(32, 32, 16)
    TAMX_CFG(32, 32, 32)
    TAMX_CFG(128, 64, 32)
    TAMX_CFG(64, 128, 32)
  }
#undef TAMX_CFG
  tamx::set_error(GM ? "unsupported macro tile for gathered operands (64x64x16, 128x64x16 "
                       "or 64x128x16)"
                     : "unsupported macro tile: see tamx_gemm_config_get");
  return TAMX_ERR_UNSUPPORTED;
}

}  // namespace

namespace tamx {
// The plain-operand launches (tamx_gemm.hip), with an optional result
// scatter plan.
int plain_gemm_f64(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b, double alpha,
                   const double* A, const double* B, double* C, const tamx_gemm_config* cfg,
                   const GatherTile* c_scatter, void* stream);
int plain_gemm_f32(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b, float alpha,
                   const float* A, const float* B, float* C, const tamx_gemm_config* cfg,
                   const GatherTile* c_scatter, void* stream);
}  // namespace tamx

namespace {

inline int plain_config(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b, double alpha,
                        const double* A, const double* B, double* C,
                        const tamx_gemm_config* cfg, const tamx::GatherTile* c_scatter,
                        void* stream) {
  return tamx::plain_gemm_f64(ctx, wl, op_a, op_b, alpha, A, B, C, cfg, c_scatter, stream);
}
inline int plain_config(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b, float alpha,
                        const float* A, const float* B, float* C,
                        const tamx_gemm_config* cfg, const tamx::GatherTile* c_scatter,
                        void* stream) {
  return tamx::plain_gemm_f32(ctx, wl, op_a, op_b, alpha, A, B, C, cfg, c_scatter, stream);
}

// One launch with gathered operands (null plans: plain operands) and an
// optional result scatter plan.
template <typename T>
int gemm_gathered(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b, T alpha,
                  const T* A, const tamx_gather_plan* a_gather, const T* B,
                  const tamx_gather_plan* b_gather, T* C, const tamx_gather_plan* c_scatter,
                  const tamx_gemm_config* cfg, void* stream) {
  TAMX_REQUIRE(a_gather == nullptr || a_gather->ctx == ctx, "A gather plan of another context");
  TAMX_REQUIRE(b_gather == nullptr || b_gather->ctx == ctx, "B gather plan of another context");
  TAMX_REQUIRE(c_scatter == nullptr || c_scatter->ctx == ctx, "scatter plan of another context");
  const tamx::GatherTile* ga = a_gather ? a_gather->tiles : nullptr;
  const tamx::GatherTile* gb = b_gather ? b_gather->tiles : nullptr;
  const tamx::GatherTile* gc = c_scatter ? c_scatter->tiles : nullptr;
  if (ga && gb)
    return grouped_gemm<T, 3>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, cfg, ga, gb, gc);
  if (ga) return grouped_gemm<T, 1>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, cfg, ga, gb, gc);
  if (gb) return grouped_gemm<T, 2>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, cfg, ga, gb, gc);
  return plain_config(ctx, wl, op_a, op_b, alpha, A, B, C, cfg, gc, stream);
}

}  // namespace
