// Grouped split-precision GEMM with wide macro tiles: 128 x 64 (or 64 x
// 128) results per workgroup, 64 x 32 (32 x 64) per wave, for 2 to 4 BF16
// limbs.  Same arithmetic as split_gemm_kernel (three FP32 accumulators per
// block, P_0 and P_1 exact over 128 k, FP64 folds at pair ends and every
// 128 k), with twice the results per wave, so every limb fragment read from
// LDS feeds more MFMA work and each workgroup loads a quarter fewer limb
// bytes per result.  The accumulators need more registers than two waves
// per SIMD allow, so one wave per SIMD runs: the loads for the next chunk
// are in flight while the current one is multiplied.
#include <type_traits>

#include "tamx_internal.hpp"
#include "tamx_split.hpp"

namespace {

using namespace tamx::split;

// I8: INT8 limbs on the INT8 MFMA with one exact INT32 accumulator per level
// (as in split_gemm_kernel); otherwise BF16 limbs.
template <int S, int WI, int WJ, bool I8>
__global__ __launch_bounds__(kThreads, 1) void split_gemm_wide_kernel(SplitParams p) {
  constexpr int NG = I8 ? S : (S >= 3 ? 3 : S);  // BF16: P_0, P_1, P_T; INT8: levels
  constexpr int CK = I8 ? SBK8 : SBK;            // k per chunk
  constexpr int USH = I8 ? 4 : 3;                // log2 limbs per 16-byte unit
  constexpr int FOLD = I8 ? kFoldChunks8 : kFoldChunks;
  using lvl_t = std::conditional_t<I8, i4_t, f4_t>;
  constexpr int BM = 32 * WI, BN = 32 * WJ;   // macro tile (2 x 2 waves)
  constexpr int RA = BM / 64, RB = BN / 64;   // staged rows per thread per plane
  static_assert(RA >= 1 && RB >= 1, "macro tile below 64 rows");
  __shared__ bf8_t sA[2][S][BM * 4];
  __shared__ bf8_t sB[2][S][BN * 4];
  int* s_item = reinterpret_cast<int*>(&sA[0][0][0]);

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = uni(tid >> 6);
  const int wr = (wave >> 1) * (16 * WI), wc = (wave & 1) * (16 * WJ);
  const int lrow = lane & 15, lq = lane >> 4;  // MFMA: row lrow, k unit lq
  // staging: thread loads rows tid / 4 + 64 u, unit tid % 4, of every plane
  const int st_r = tid >> 2, st_q = tid & 3;

  int64_t ntasks64 = p.counters[TAMX_WL_NTASKS];
  if (ntasks64 < 0 || ntasks64 > p.task_capacity || p.counters[TAMX_WL_STATUS] != 0)
    ntasks64 = 0;
  const int32_t ntasks = uni((int32_t)ntasks64);
  const int32_t total = uni(p.prefix[ntasks > 0 ? ntasks : 0]);
  const int home = blockIdx.x % kRegions;
  int tried = 0;
  for (;;) {
    __syncthreads();  // every wave is done with the LDS buffers
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
    const auto* tp = tamx::scalar_ptr(p.tasks) + lo;
    const int64_t t_coff = uni64(tp->c_off);
    const int t_m = uni(tp->m), t_n = uni(tp->n), t_ldc = uni(tp->ldc);
    const int t_pb = uni(tp->pair_begin), t_pc = uni(tp->pair_count);
    const int t_flags = uni(tp->flags);
    const int local = item - uni(p.prefix[lo]);
    const int tiles_n = (t_n + BN - 1) / BN;
    const int r0 = (local / tiles_n) * BM, c0 = (local % tiles_n) * BN;
    const int mrem = min(BM, t_m - r0), nrem = min(BN, t_n - c0);
    // 16 x 16 blocks of the wave inside the tile: rows [0, ni), columns
    // [0, nj) (uniform per wave; zero for a wave wholly outside)
    const int ni = uni(max(0, min(WI, (mrem - wr + 15) / 16)));
    const int nj = uni(max(0, min(WJ, (nrem - wc + 15) / 16)));
    const bool wave_active = ni > 0 && nj > 0;

    d4_t acc[WI][WJ];
    lvl_t lvl[NG][WI][WJ];
#pragma unroll
    for (int i = 0; i < WI; ++i)
#pragma unroll
      for (int j = 0; j < WJ; ++j) {
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[i][j][v] = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
          for (int v = 0; v < 4; ++v) lvl[g][i][j][v] = 0;
      }

    const int pend = t_pb + t_pc;

    // load cursor: pair lp, chunk lk0; runs one chunk ahead of compute
    int lp = t_pb, lk0 = 0, lkp = 0;
    int64_t la_row = 0, lb_row = 0, la_plane = 0, lb_plane = 0, la_step = 0, lb_step = 0;
    int64_t la_blk = 0, lb_blk = 0;
    auto set_pair = [&](int pi) {
      const auto& pp = tamx::scalar_ptr(p.pairs)[pi];
      const auto* A = tamx::scalar_ptr(p.ta) + uni(pp.a_tile);
      const auto* B = tamx::scalar_ptr(p.tb) + uni(pp.b_tile);
      lkp = uni(A->k_pad);
      // limb offsets in 16-byte units: plane stride, this thread's first
      // staging row, and the 64-row step between its rows (INT8: in the
      // 32-k block layout, with the units per block)
      la_plane = uni64((int64_t)uni(A->rows_pad) * lkp) >> USH;
      lb_plane = uni64((int64_t)uni(B->rows_pad) * lkp) >> USH;
      if constexpr (I8) {
        la_blk = 2 * uni(A->rows_pad);
        lb_blk = 2 * uni(B->rows_pad);
        la_row = i8_unit(uni64(A->limb_off) >> 4, 0, r0 + st_r, 0);
        lb_row = i8_unit(uni64(B->limb_off) >> 4, 0, c0 + st_r, 0);
        la_step = lb_step = 2 * 64;
      } else {
        la_row = (uni64(A->limb_off) + (int64_t)(r0 + st_r) * lkp) >> USH;
        lb_row = (uni64(B->limb_off) + (int64_t)(c0 + st_r) * lkp) >> USH;
        la_step = lb_step = (int64_t)(64 >> USH) * lkp;
      }
    };
    if (lp < pend) set_pair(lp);
    // Once the cursor has passed the last pair the loads repeat its first
    // chunk, unused: every step issues the same loads, which keeps the wait
    // counts before the LDS stores exact.
    auto load_chunk = [&](bf8_t (&ra)[S][RA], bf8_t (&rb)[S][RB]) {
      int64_t qa, qb;  // unit st_q of the chunk at lk0 in the first staging row
      if constexpr (I8) {
        qa = la_row + ((lk0 / kI8Block) + (st_q >> 1)) * la_blk + (st_q & 1);
        qb = lb_row + ((lk0 / kI8Block) + (st_q >> 1)) * lb_blk + (st_q & 1);
      } else {
        qa = la_row + (lk0 >> USH) + st_q;
        qb = lb_row + (lk0 >> USH) + st_q;
      }
#pragma unroll
      for (int s = 0; s < S; ++s) {
#pragma unroll
        for (int u = 0; u < RA; ++u) ra[s][u] = p.LA[qa + u * la_step + s * la_plane];
#pragma unroll
        for (int u = 0; u < RB; ++u) rb[s][u] = p.LB[qb + u * lb_step + s * lb_plane];
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
    auto store_chunk = [&](int buf, const bf8_t (&ra)[S][RA], const bf8_t (&rb)[S][RB]) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
#pragma unroll
        for (int u = 0; u < RA; ++u) sA[buf][s][lds_unit(st_r + 64 * u, st_q)] = ra[s][u];
#pragma unroll
        for (int u = 0; u < RB; ++u) sB[buf][s][lds_unit(st_r + 64 * u, st_q)] = rb[s][u];
      }
    };
    // the wave's rows of blocks i < ni, with NJ column blocks
    auto compute_blocks = [&](int buf, auto NJ_) {
      constexpr int NJ = decltype(NJ_)::value;
      bf8_t bf[NJ][S];
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int t = 0; t < S; ++t) bf[j][t] = sB[buf][t][lds_unit(wc + j * 16 + lrow, lq)];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        bf8_t af[WI];
#pragma unroll
        for (int i = 0; i < WI; ++i)
          if (i < ni) af[i] = sA[buf][s][lds_unit(wr + i * 16 + lrow, lq)];
#pragma unroll
        for (int t = 0; t < S - s; ++t) {
          const int g = I8 ? s + t : (s + t < 2 ? s + t : 2);
#pragma unroll
          for (int i = 0; i < WI; ++i) {
            if (i < ni) {
#pragma unroll
              for (int j = 0; j < NJ; ++j) {
                if constexpr (I8)
                  lvl[g][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                      __builtin_bit_cast(i4_t, af[i]), __builtin_bit_cast(i4_t, bf[j][t]),
                      lvl[g][i][j], 0, 0, 0);
                else
                  lvl[g][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j][t],
                                                                         lvl[g][i][j], 0, 0, 0);
              }
            }
          }
        }
      }
    };
    auto compute_chunk = [&](int buf) {
      if (!wave_active) return;
      if constexpr (WJ == 4) {
        if (nj == 4) compute_blocks(buf, std::integral_constant<int, 4>{});
        else if (nj == 3) compute_blocks(buf, std::integral_constant<int, 3>{});
        else if (nj == 2) compute_blocks(buf, std::integral_constant<int, 2>{});
        else compute_blocks(buf, std::integral_constant<int, 1>{});
      } else {
        if (nj == 2) compute_blocks(buf, std::integral_constant<int, 2>{});
        else compute_blocks(buf, std::integral_constant<int, 1>{});
      }
    };
    // fold the FP32 accumulators of the pair being computed (scale
    // 2^(sigma_A + sigma_B - 16)) for the blocks inside the tile
    auto fold = [&](double scale) {
#pragma unroll
      for (int i = 0; i < WI; ++i)
#pragma unroll
        for (int j = 0; j < WJ; ++j) {
          if (i < ni && j < nj) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              double tsum = 0.0;
#pragma unroll
              for (int g = NG - 1; g >= 0; --g) {
                if constexpr (I8)
                  tsum = fma((double)lvl[g][i][j][v], 1.0 / (double)(1LL << (7 * g)), tsum);
                else
                  tsum += (double)lvl[g][i][j][v];
                lvl[g][i][j][v] = 0;
              }
              acc[i][j][v] = fma(tsum, scale, acc[i][j][v]);
            }
          }
        }
    };

    // compute cursor: pair cp, chunks done since the last fold
    int cp = t_pb, ck0 = 0, ckp = 0, since_fold = 0;
    double cscale = 0.0;
    auto set_compute_pair = [&](int pi) {
      const auto& pp = tamx::scalar_ptr(p.pairs)[pi];
      const auto* A = tamx::scalar_ptr(p.ta) + uni(pp.a_tile);
      const auto* B = tamx::scalar_ptr(p.tb) + uni(pp.b_tile);
      ckp = uni(A->k_pad);
      cscale = ldexp(1.0, uni(A->sigma) + uni(B->sigma) - (I8 ? 14 : 16));
    };
    if (cp < pend) set_compute_pair(cp);
    auto after_compute = [&]() {
      ++since_fold;
      ck0 += CK;
      const bool pair_end = ck0 >= ckp;
      if (pair_end || since_fold == FOLD) {
        if (wave_active) fold(cscale);
        since_fold = 0;
      }
      if (pair_end) {
        ++cp;
        ck0 = 0;
        if (cp < pend) set_compute_pair(cp);
      }
    };

    // one register set in flight: chunk c + 1 loads while chunk c is
    // multiplied out of LDS buffer c % 2; one barrier per chunk
    bf8_t ra[S][RA], rb[S][RB];
    if (lp < pend) {
      load_chunk(ra, rb);
      store_chunk(0, ra, rb);
      __syncthreads();
      int buf = 0;
      for (;;) {
        const bool has_next = lp < pend;
        load_chunk(ra, rb);
        compute_chunk(buf);
        after_compute();
        if (!has_next) break;
        store_chunk(buf ^ 1, ra, rb);
        __syncthreads();
        buf ^= 1;
      }
    }

    const bool beta1 = (t_flags & TAMX_TASK_BETA1) != 0;
    double* __restrict__ c = p.C + t_coff;
    tamx::with_result_index(p.c_scatter, p.task_tile, lo, t_flags, t_ldc, [&](auto cidx) {
#pragma unroll
      for (int i = 0; i < WI; ++i)
#pragma unroll
        for (int j = 0; j < WJ; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            // f32-output 16x16 MFMA layout: row 4 * (lane / 16) + v, column lane % 16
            const int row = wr + i * 16 + lq * 4 + v;
            const int col = wc + j * 16 + lrow;
            if (row < mrem && col < nrem) {
              const int64_t idx = cidx(r0 + row, c0 + col);
              double val = p.alpha * acc[i][j][v];
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

template <int S, int WI, int WJ, bool I8>
int launch_wide(tamx_context* ctx, tamx_worklist* wl, const SplitParams& prm, hipStream_t s) {
  auto kern = split_gemm_wide_kernel<S, WI, WJ, I8>;
  int occ = 0;
  TAMX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kThreads, 0));
  if (occ < 1) occ = 1;
  if (int rc = tamx::plan_items(wl, 32 * WI, 32 * WJ, s)) return rc;
  kern<<<ctx->num_cus * occ, kThreads, 0, s>>>(prm);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

}  // namespace

int tamx::split::launch_split_gemm_wide(int limb_type, int nslices, int bm, int bn,
                                        tamx_context* ctx, tamx_worklist* wl,
                                        const SplitParams& prm, hipStream_t s) {
#define TAMX_WIDE(I8_, S_, BM_, BN_)                            \
  if ((limb_type == TAMX_LIMB_I8) == I8_ && nslices == S_ && bm == BM_ && bn == BN_) \
    return launch_wide<S_, BM_ / 32, BN_ / 32, I8_>(ctx, wl, prm, s);
  TAMX_WIDE(false, 2, 128, 64)
  TAMX_WIDE(false, 3, 128, 64)
  TAMX_WIDE(false, 4, 128, 64)
  TAMX_WIDE(false, 2, 64, 128)
  TAMX_WIDE(false, 3, 64, 128)
  TAMX_WIDE(false, 4, 64, 128)
  // (five INT8 limbs at these tiles would need more than 512 registers per
  // lane; this compiler also fails on them)
  TAMX_WIDE(true, 3, 128, 64)
  TAMX_WIDE(true, 4, 128, 64)
  TAMX_WIDE(true, 3, 64, 128)
  TAMX_WIDE(true, 4, 64, 128)
#undef TAMX_WIDE
  tamx::set_error("split GEMM macro tile: 64x64 (any limb count), 128x64 or 64x128 (2 to 4 "
                  "BF16 or 3 to 4 INT8 limbs), 128x64 (5 or 6 INT8 limbs)");
  return TAMX_ERR_UNSUPPORTED;
}
