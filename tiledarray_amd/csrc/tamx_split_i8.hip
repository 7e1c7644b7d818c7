// Grouped split GEMM on INT8 limbs, two-wave workgroups.
//
// Same arithmetic as split_gemm_kernel<S, true> (tamx_split.hip): per tile
// pair the limb products s + t < S on the INT8 MFMA 16x16x64, one exact
// INT32 accumulator per level s + t, folded into FP64 at pair ends with the
// scale 2^(sigma_A + sigma_B - 14) and the level weights 2^-7(s+t).  The
// results are identical bit for bit: every level sum is exact and the folds
// happen at the same points in the same order.
//
// What differs is the MFMA and how the limbs reach it.  The products run
// on the INT8 MFMA 32x32x32, which on this part does twice the work of the
// 16x16x64 in about the same issue time (measured: 4.07 against 2.29 INT8
// POP/s in a register-only loop).  A workgroup of two waves computes a
// 64 x 64 macro tile, each wave 64 rows x 32 columns (two 32 x 32 blocks),
// so every B fragment feeds two blocks.  The A limbs (shared by both waves)
// go from global memory straight into a three-stage LDS ring; the B limbs
// (each wave its own columns) go straight into registers, three sets in a
// ring.  Chunk c + 2 is loaded while chunk c is multiplied, one barrier per
// chunk.  One wave per SIMD: the level and FP64 accumulators alone take 224
// registers.
//
// 32x32x32 INT8 operand layout (checked on the device): lane l holds row
// (A) or column (B) l % 32 and the 16 k from 16 (l / 32) of the 32-k step;
// result register v of lane l is row 8 (v / 4) + 4 (l / 32) + v % 4,
// column l % 32.
#include <type_traits>

#include "tamx_internal.hpp"
#include "tamx_split.hpp"

namespace {

using namespace tamx::split;


constexpr int kWaves = 2;
constexpr int kT = 64 * kWaves;  // threads per workgroup
constexpr int kStages = 3;       // A chunks in the LDS ring
constexpr int IBM = 64, IBN = 64;
constexpr int kAUnits = IBM * 4;          // 16-byte units of one limb plane of an A chunk
constexpr int kALoads = kAUnits / kT;     // direct loads per thread per plane
static_assert(kAUnits % kT == 0, "A chunk split over the threads");

// vmcnt wait (the other counters left alone), gfx9 encoding
template <int X>
using st_t = std::integral_constant<int, X>;

constexpr int waitcnt_vm(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }

// Workgroup barrier for the pipelined loops below, after the caller's
// explicit wait for the loads it needs.  (__syncthreads carries a workgroup
// release fence, which makes the compiler wait for every direct-to-LDS load
// in flight, those of later chunks included; the loops keep those in flight
// on purpose.  Their LDS reads have completed before each barrier: the MFMAs
// that consume them precede it.)
__device__ inline void ring_barrier() { asm volatile("s_barrier" ::: "memory"); }

template <int S>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(1, 1)))
void split_gemm_i8_kernel(SplitParams p) {
  constexpr int NG = S;     // one level per s + t
  constexpr int CK = SBK8;  // 64 k per chunk
  // vector memory operations per thread per chunk: A into LDS, B into registers
  constexpr int kOps = S * kALoads + S * 2;
  // one variable per stage: the compiler then knows that loads landing in
  // one stage do not alias reads of another, and waits for them only where
  // that stage is read
  __shared__ i4_t sA0[S][kAUnits], sA1[S][kAUnits], sA2[S][kAUnits];
  __shared__ int s_item;
  static_assert(kStages == 3, "three stage variables");
  auto stage = [&](auto X) -> i4_t(&)[S][kAUnits] {
    constexpr int x = decltype(X)::value;
    if constexpr (x == 0) return sA0;
    else if constexpr (x == 1) return sA1;
    else return sA2;
  };

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = uni(tid >> 6);
  const int wc = wave * 32;                    // the wave's columns
  const auto* LA = reinterpret_cast<const i4_t*>(p.LA);
  const auto* LB = reinterpret_cast<const i4_t*>(p.LB);

  int64_t ntasks64 = p.counters[TAMX_WL_NTASKS];
  if (ntasks64 < 0 || ntasks64 > p.task_capacity || p.counters[TAMX_WL_STATUS] != 0)
    ntasks64 = 0;
  const int32_t ntasks = uni((int32_t)ntasks64);
  const int32_t total = uni(p.prefix[ntasks > 0 ? ntasks : 0]);
  const int home = blockIdx.x % kRegions;
  int tried = 0;
  for (;;) {
    __syncthreads();  // both waves are done with the LDS ring and s_item
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
    const auto* tp = tamx::scalar_ptr(p.tasks) + lo;
    const int64_t t_coff = uni64(tp->c_off);
    const int t_m = uni(tp->m), t_n = uni(tp->n), t_ldc = uni(tp->ldc);
    const int t_pb = uni(tp->pair_begin), t_pc = uni(tp->pair_count);
    const int t_flags = uni(tp->flags);
    const int local = item - uni(p.prefix[lo]);
    const int tiles_n = (t_n + IBN - 1) / IBN;
    const int r0 = (local / tiles_n) * IBM, c0 = (local % tiles_n) * IBN;
    const int mrem = min(IBM, t_m - r0), nrem = min(IBN, t_n - c0);
    // 32-row blocks inside the tile (both waves); a wave whose 32 columns
    // lie wholly past the tile's edge skips its MFMAs
    const int ni = uni(min(2, (mrem + 31) / 32));
    const bool wave_active = wc < nrem;

    double acc[2][16];
    i16_t lvl[NG][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        acc[i][v] = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g) lvl[g][i][v] = 0;
      }

    const int pend = t_pb + t_pc;

    // load cursor: pair lp, chunk lk0, two chunks ahead of compute.  Per
    // pair: this thread's A units (16-byte units of the limb pool) of its
    // kALoads slots in plane 0 at k = 0, the plane strides, and its B
    // fragment unit (row c0 + wc + lane % 32, k unit lane / 32; the second
    // 32-k step of a chunk is two units further).
    int lp = t_pb, lk0 = 0, lkp = 0;
    int64_t a_unit[kALoads], b_unit[2], a_plane = 0, b_plane = 0, a_blk = 0, b_blk = 0;
    auto set_pair = [&](int pi) {
      const auto& pp = tamx::scalar_ptr(p.pairs)[pi];
      const auto* A = tamx::scalar_ptr(p.ta) + uni(pp.a_tile);
      const auto* B = tamx::scalar_ptr(p.tb) + uni(pp.b_tile);
      lkp = uni(A->k_pad);
      a_plane = uni64((int64_t)uni(A->rows_pad) * lkp) >> 4;
      b_plane = uni64((int64_t)uni(B->rows_pad) * lkp) >> 4;
      const int a_rp = uni(A->rows_pad), b_rp = uni(B->rows_pad);
      a_blk = 2 * a_rp;  // units per 32-k block
      b_blk = 2 * b_rp;
      const int64_t a0 = uni64(A->limb_off) >> 4, b0 = uni64(B->limb_off) >> 4;
#pragma unroll
      for (int h = 0; h < kALoads; ++h) {
        // LDS slot L of this lane in load h holds row L / 4, unit L % 4 of
        // the swizzled layout, i.e. source unit q = (L % 4) ^ ((row / 2) % 4)
        const int L = kT * h + 64 * wave + lane;
        const int r = L >> 2, q = (L & 3) ^ ((r >> 1) & 3);
        a_unit[h] = i8_unit(a0, a_rp, r0 + r, q);
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
        b_unit[h] = i8_unit(b0, b_rp, c0 + wc + (lane & 31), (lane >> 5) + 2 * h);
    };
    if (lp < pend) set_pair(lp);
    // Loads of the cursor's chunk: A into LDS stage `st`, B into `bs`.
    // Once the cursor has passed the last pair the loads repeat a chunk of
    // the last pair, unused, so every step issues the same number of loads
    // and the wait counts stay exact.
    auto load_chunk = [&](auto st, i4_t (&bs)[S][2]) {
      // the chunk's first 32-k block in each operand's layout
      const int64_t ka = (lk0 / kI8Block) * a_blk, kb = (lk0 / kI8Block) * b_blk;
#pragma unroll
      for (int s = 0; s < S; ++s) {
#pragma unroll
        for (int h = 0; h < kALoads; ++h)
          __builtin_amdgcn_global_load_lds(
              (const __attribute__((address_space(1))) void*)(LA + a_unit[h] + s * a_plane + ka),
              (__attribute__((address_space(3))) void*)(&stage(st)[s][kT * h + 64 * wave]), 16, 0, 0);
#pragma unroll
        for (int h = 0; h < 2; ++h) bs[s][h] = *tamx::global_ptr(LB + b_unit[h] + s * b_plane + kb);
      }
      if (lp < pend) {
        lk0 += CK;
        if (lk0 >= lkp) {
          ++lp;
          lk0 = 0;
          if (lp < pend) set_pair(lp);
        }
      }
    };

    // The wave's 32-row blocks i < ni lie inside the tile (uniform per
    // wave).  Block rows and the chunk's two 32-k steps are multiplied one
    // after the other, each block behind its own uniform branch, so a ragged
    // edge costs no MFMAs; the next step's A fragments are read from LDS
    // (always: rows past the edge hold whatever lies there) while the
    // current step's MFMAs run.  bs[t][h]: limb t of the B fragment for
    // k step h.
    auto compute_chunk = [&](auto st, const i4_t (&bs)[S][2]) {
      if (!wave_active) return;
      auto read_a = [&](i4_t (&fa)[S], int i, int h) {
#pragma unroll
        for (int s = 0; s < S; ++s)
          fa[s] = stage(st)[s][lds_unit(32 * i + (lane & 31), 2 * h + (lane >> 5))];
      };
      i4_t fa[2][S];
      read_a(fa[0], 0, 0);
#pragma unroll
      for (int step = 0; step < 4; ++step) {
        const int i = step >> 1, h = step & 1, cur = step & 1;
        if (step < 3) read_a(fa[cur ^ 1], (step + 1) >> 1, (step + 1) & 1);
        if (i < ni) {
#pragma unroll
          for (int s = 0; s < S; ++s)
#pragma unroll
            for (int t = 0; t < S - s; ++t)
              lvl[s + t][i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[cur][s], bs[t][h],
                                                                     lvl[s + t][i], 0, 0, 0);
        }
      }
    };
    // fold the levels of the pair just finished into FP64 (level weights
    // 2^-7g are exact, so each fma rounds once, as a sum would); blocks past
    // the edge hold zeros and are skipped
    auto fold = [&](double scale) {
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if (i < ni) {
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            double tsum = 0.0;
#pragma unroll
            for (int g = NG - 1; g >= 0; --g) {
              tsum = fma((double)lvl[g][i][v], 1.0 / (double)(1LL << (7 * g)), tsum);
              lvl[g][i][v] = 0;
            }
            acc[i][v] = fma(tsum, scale, acc[i][v]);
          }
        }
      }
    };

    // compute cursor: pair cp, chunk ck0; chunks since the last fold
    int cp = t_pb, ck0 = 0, ckp = 0, since_fold = 0;
    double cscale = 0.0;
    auto set_compute_pair = [&](int pi) {
      const auto& pp = tamx::scalar_ptr(p.pairs)[pi];
      const auto* A = tamx::scalar_ptr(p.ta) + uni(pp.a_tile);
      const auto* B = tamx::scalar_ptr(p.tb) + uni(pp.b_tile);
      ckp = uni(A->k_pad);
      cscale = ldexp(1.0, uni(A->sigma) + uni(B->sigma) - 14);
    };
    if (cp < pend) set_compute_pair(cp);
    auto after_compute = [&]() {
      ++since_fold;
      ck0 += CK;
      const bool pair_end = ck0 >= ckp;
      if (pair_end || since_fold == kFoldChunks8) {
        if (wave_active) fold(cscale);
        since_fold = 0;
      }
      if (pair_end) {
        ++cp;
        ck0 = 0;
        if (cp < pend) set_compute_pair(cp);
      }
    };

    i4_t b0[S][2], b1[S][2], b2[S][2];
    if (lp < pend) {
      load_chunk(st_t<0>{}, b0);  // chunk 0
      load_chunk(st_t<1>{}, b1);  // chunk 1 (or a repeat)
      // one step: chunk c in stage X with B in bx; chunk c + 1 in flight;
      // chunk c + 2 goes to stage (X + 2) % 3, which held chunk c - 1
      auto step = [&](auto X, const i4_t (&bx)[S][2], i4_t (&bn)[S][2]) {
        __builtin_amdgcn_s_waitcnt(waitcnt_vm(kOps));
        ring_barrier();
        load_chunk(st_t<(decltype(X)::value + 2) % 3>{}, bn);
        compute_chunk(X, bx);
        after_compute();
        return cp < pend;
      };
      for (;;) {
        if (!step(st_t<0>{}, b0, b2)) break;
        if (!step(st_t<1>{}, b1, b0)) break;
        if (!step(st_t<2>{}, b2, b1)) break;
      }
      // the repeats still in flight land in the ring before it is reused
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }

    const bool beta1 = (t_flags & TAMX_TASK_BETA1) != 0;
    double* __restrict__ c = p.C + t_coff;
    tamx::with_result_index(p.c_scatter, p.task_tile, lo, t_flags, t_ldc, [&](auto cidx) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          // 32x32 MFMA output: row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32
          const int row = 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
          const int col = wc + (lane & 31);
          if (row < mrem && col < nrem) {
            const int64_t idx = cidx(r0 + row, c0 + col);
            double val = p.alpha * acc[i][v];
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
      for (int r = 0; r <= kRegions; ++r) atomicExch(&p.queue[r], 0);
    }
  }
}

template <int S>
int launch_i8(tamx_context* ctx, tamx_worklist* wl, const SplitParams& prm, hipStream_t s) {
  auto kern = split_gemm_i8_kernel<S>;
  int occ = 0;
  TAMX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kT, 0));
  if (occ < 1) occ = 1;
  if (int rc = tamx::plan_items(wl, IBM, IBN, s)) return rc;
  kern<<<ctx->num_cus * occ, kT, 0, s>>>(prm);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

// ------------------------------------------------------------ the ring --
//
// Four waves on a 64 x 64 macro tile, each wave's 32 x 32 quadrant one block
// of the INT8 MFMA 32x32x32 (the same products, levels and folds as above,
// so the same result bit for bit).  Chunks of 32 k stream through a ring of
// kRing LDS stages filled by direct global->LDS loads, kRing - 1 chunks in
// flight while one is multiplied: the load path of these kernels is bound by
// the latency of the limb loads (about 2 us a chunk under load), so what
// counts is the bytes in flight, and 32-k chunks buy three of them for the
// LDS of two 64-k ones.  Chunks that hold only k padding (past a pair's
// contracted extent) are neither loaded nor multiplied.  Waves 0 and 1 load
// the A limbs of a stage, waves 2 and 3 the B limbs.

constexpr int kRing = 4;                 // LDS stages
constexpr int RCK = 32;                  // k per chunk
constexpr int kRingUnits = 64 * RCK / 16;  // 16-byte units of one plane of one operand

// LDS slot of unit q (k 16q .. 16q + 15) of row r: rows of two units, the
// unit index flipped every four rows (and back every eight) so that the
// rows a fragment read touches at once fall in distinct banks
__device__ inline int ring_slot(int r, int q) { return 2 * r + (q ^ (((r >> 2) ^ (r >> 3)) & 1)); }

template <int S>
__global__ __launch_bounds__(256, S <= 5 ? 2 : 1) void split_gemm_i8_ring_kernel(SplitParams p) {
  constexpr int NG = S;
  constexpr int kFold = kFoldChunks8 * (SBK8 / RCK);  // chunks between folds (4096 k)
  // one variable per stage and operand (the compiler then waits for a
  // stage's loads only where that stage is read)
  __shared__ i4_t rA0[S][kRingUnits], rA1[S][kRingUnits], rA2[S][kRingUnits], rA3[S][kRingUnits];
  __shared__ i4_t rB0[S][kRingUnits], rB1[S][kRingUnits], rB2[S][kRingUnits], rB3[S][kRingUnits];
  // the grabbed item passes through the first word of stage 0 while the
  // ring is idle between items
  int* s_item = reinterpret_cast<int*>(&rA0[0][0]);
  static_assert(kRing == 4, "four stage variables");
  auto stA = [&](auto X) -> i4_t(&)[S][kRingUnits] {
    constexpr int x = decltype(X)::value;
    if constexpr (x == 0) return rA0;
    else if constexpr (x == 1) return rA1;
    else if constexpr (x == 2) return rA2;
    else return rA3;
  };
  auto stB = [&](auto X) -> i4_t(&)[S][kRingUnits] {
    constexpr int x = decltype(X)::value;
    if constexpr (x == 0) return rB0;
    else if constexpr (x == 1) return rB1;
    else if constexpr (x == 2) return rB2;
    else return rB3;
  };

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = uni(tid >> 6);
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const bool loads_a = wave < 2;
  // this thread's load slot: unit u of its operand's plane, row u / 2
  const int lu = 64 * (wave & 1) + lane;
  const int l_r = lu >> 1;
  const int l_q = (lu & 1) ^ (((l_r >> 2) ^ (l_r >> 3)) & 1);  // source unit of that slot
  const auto* LA = reinterpret_cast<const i4_t*>(p.LA);
  const auto* LB = reinterpret_cast<const i4_t*>(p.LB);

  int64_t ntasks64 = p.counters[TAMX_WL_NTASKS];
  if (ntasks64 < 0 || ntasks64 > p.task_capacity || p.counters[TAMX_WL_STATUS] != 0)
    ntasks64 = 0;
  const int32_t ntasks = uni((int32_t)ntasks64);
  const int32_t total = uni(p.prefix[ntasks > 0 ? ntasks : 0]);
  const int home = blockIdx.x % kRegions;
  int tried = 0;
  for (;;) {
    __syncthreads();  // every wave is done with the ring
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
      *s_item = it;
    }
    __syncthreads();
    const int item = uni(*s_item);
    __syncthreads();  // read before the ring is loaded again
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
    const auto* tp = tamx::scalar_ptr(p.tasks) + lo;
    const int64_t t_coff = uni64(tp->c_off);
    const int t_m = uni(tp->m), t_n = uni(tp->n), t_ldc = uni(tp->ldc);
    const int t_pb = uni(tp->pair_begin), t_pc = uni(tp->pair_count);
    const int t_flags = uni(tp->flags);
    const int local = item - uni(p.prefix[lo]);
    const int tiles_n = (t_n + 63) / 64;
    const int r0 = (local / tiles_n) * 64, c0 = (local % tiles_n) * 64;
    const int mrem = min(64, t_m - r0), nrem = min(64, t_n - c0);
    const bool wave_active = (wr < mrem) && (wc < nrem);

    double acc[16];
    i16_t lvl[NG];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      acc[v] = 0.0;
#pragma unroll
      for (int g = 0; g < NG; ++g) lvl[g][v] = 0;
    }
    const int pend = t_pb + t_pc;

    // load cursor: pair lp, k lk0 of its chunk, lkend its k rounded up to
    // the chunk; this thread's source unit (plane 0, k 0) and plane stride
    int lp = t_pb, lk0 = 0, lkend = 0;
    int64_t l_unit = 0, l_plane = 0, l_blk = 0;
    auto set_pair = [&](int pi) {
      const auto& pp = tamx::scalar_ptr(p.pairs)[pi];
      const auto* A = tamx::scalar_ptr(p.ta) + uni(pp.a_tile);
      const auto* B = tamx::scalar_ptr(p.tb) + uni(pp.b_tile);
      lkend = (uni(A->out_k) + RCK - 1) / RCK * RCK;
      const auto* X = loads_a ? A : B;  // (uniform per wave)
      const int kp = uni(X->k_pad), rp = uni(X->rows_pad);
      l_plane = uni64((int64_t)rp * kp) >> 4;
      l_blk = 2 * rp;  // units per 32-k block: a chunk is one block
      l_unit = i8_unit(uni64(X->limb_off) >> 4, rp, (loads_a ? r0 : c0) + l_r, l_q);
    };
    if (lp < pend) set_pair(lp);
    // loads of the cursor's chunk into stage X (S per thread); once the
    // cursor has passed the last pair they repeat a chunk of it, unused, so
    // every step issues the same number of loads and the waits stay exact
    static_assert(RCK == kI8Block, "a ring chunk is one block of the limb layout");
    auto load_chunk = [&](auto X) {
      const int64_t k16 = (lk0 / kI8Block) * l_blk;
      if (loads_a) {
#pragma unroll
        for (int s = 0; s < S; ++s)
          __builtin_amdgcn_global_load_lds(
              (const __attribute__((address_space(1))) void*)(LA + l_unit + s * l_plane + k16),
              (__attribute__((address_space(3))) void*)(&stA(X)[s][64 * (wave & 1)]), 16, 0, 0);
      } else {
#pragma unroll
        for (int s = 0; s < S; ++s)
          __builtin_amdgcn_global_load_lds(
              (const __attribute__((address_space(1))) void*)(LB + l_unit + s * l_plane + k16),
              (__attribute__((address_space(3))) void*)(&stB(X)[s][64 * (wave & 1)]), 16, 0, 0);
      }
      if (lp < pend) {
        lk0 += RCK;
        if (lk0 >= lkend) {
          ++lp;
          lk0 = 0;
          if (lp < pend) set_pair(lp);
        }
      }
    };
    auto compute_chunk = [&](auto X) {
      if (!wave_active) return;
      const int fr = lane & 31, fq = lane >> 5;
      i4_t bf[S];
#pragma unroll
      for (int t = 0; t < S; ++t) bf[t] = stB(X)[t][ring_slot(wc + fr, fq)];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const i4_t af = stA(X)[s][ring_slot(wr + fr, fq)];
#pragma unroll
        for (int t = 0; t < S - s; ++t)
          lvl[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf[t], lvl[s + t], 0, 0, 0);
      }
    };
    auto fold = [&](double scale) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        double tsum = 0.0;
#pragma unroll
        for (int g = NG - 1; g >= 0; --g) {
          tsum = fma((double)lvl[g][v], 1.0 / (double)(1LL << (7 * g)), tsum);
          lvl[g][v] = 0;
        }
        acc[v] = fma(tsum, scale, acc[v]);
      }
    };
    int cp = t_pb, ck0 = 0, ckend = 0, since_fold = 0;
    double cscale = 0.0;
    auto set_compute_pair = [&](int pi) {
      const auto& pp = tamx::scalar_ptr(p.pairs)[pi];
      const auto* A = tamx::scalar_ptr(p.ta) + uni(pp.a_tile);
      const auto* B = tamx::scalar_ptr(p.tb) + uni(pp.b_tile);
      ckend = (uni(A->out_k) + RCK - 1) / RCK * RCK;
      cscale = ldexp(1.0, uni(A->sigma) + uni(B->sigma) - 14);
    };
    if (cp < pend) set_compute_pair(cp);
    auto after_compute = [&]() {
      ++since_fold;
      ck0 += RCK;
      const bool pair_end = ck0 >= ckend;
      if (pair_end || since_fold == kFold) {
        if (wave_active) fold(cscale);
        since_fold = 0;
      }
      if (pair_end) {
        ++cp;
        ck0 = 0;
        if (cp < pend) set_compute_pair(cp);
      }
    };

    if (lp < pend) {
      load_chunk(st_t<0>{});
      load_chunk(st_t<1>{});
      load_chunk(st_t<2>{});
      // chunk c in stage X; chunks c + 1 and c + 2 in flight; chunk c + 3
      // goes to the stage of chunk c - 1, which every wave has finished
      auto step = [&](auto X) {
        __builtin_amdgcn_s_waitcnt(waitcnt_vm(2 * S));
        ring_barrier();
        load_chunk(st_t<(decltype(X)::value + 3) % kRing>{});
        compute_chunk(X);
        after_compute();
        return cp < pend;
      };
      for (;;) {
        if (!step(st_t<0>{})) break;
        if (!step(st_t<1>{})) break;
        if (!step(st_t<2>{})) break;
        if (!step(st_t<3>{})) break;
      }
      // the repeats still in flight land before the ring is reused
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }

    const bool beta1 = (t_flags & TAMX_TASK_BETA1) != 0;
    double* __restrict__ c = p.C + t_coff;
    tamx::with_result_index(p.c_scatter, p.task_tile, lo, t_flags, t_ldc, [&](auto cidx) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        // 32x32 MFMA output: row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32
        const int row = wr + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int col = wc + (lane & 31);
        if (row < mrem && col < nrem) {
          const int64_t idx = cidx(r0 + row, c0 + col);
          double val = p.alpha * acc[v];
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
      for (int r = 0; r <= kRegions; ++r) atomicExch(&p.queue[r], 0);
    }
  }
}

template <int S>
int launch_ring(tamx_context* ctx, tamx_worklist* wl, const SplitParams& prm, hipStream_t s) {
  auto kern = split_gemm_i8_ring_kernel<S>;
  int occ = 0;
  TAMX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 256, 0));
  if (occ < 1) occ = 1;
  if (int rc = tamx::plan_items(wl, 64, 64, s)) return rc;
  kern<<<ctx->num_cus * occ, 256, 0, s>>>(prm);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

}  // namespace

int tamx::split::launch_split_gemm_i8_ring(int nslices, tamx_context* ctx, tamx_worklist* wl,
                                           const SplitParams& prm, hipStream_t s) {
  switch (nslices) {
    case 2: return launch_ring<2>(ctx, wl, prm, s);
    case 3: return launch_ring<3>(ctx, wl, prm, s);
    case 4: return launch_ring<4>(ctx, wl, prm, s);
    case 5: return launch_ring<5>(ctx, wl, prm, s);
    default: return launch_ring<6>(ctx, wl, prm, s);
  }
}

int tamx::split::launch_split_gemm_i8_2w(int nslices, tamx_context* ctx, tamx_worklist* wl,
                                         const SplitParams& prm, hipStream_t s) {
  switch (nslices) {
    case 4: return launch_i8<4>(ctx, wl, prm, s);
    case 5: return launch_i8<5>(ctx, wl, prm, s);
    default: break;
  }
  tamx::set_error("the two-wave INT8 split GEMM takes 4 or 5 limbs");
  return TAMX_ERR_UNSUPPORTED;
}
