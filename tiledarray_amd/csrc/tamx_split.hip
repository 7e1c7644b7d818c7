// Split-precision contraction: FP64 tiles are split into BF16 limbs with one
// power-of-two scale per tile, the limb products run on BF16 MFMA with FP32
// accumulation, and the FP32 partial sums are folded into FP64 before the
// ones that matter can round.
//
// Splitting (one pass per operand, tamx_split_f64):
//   sigma = exponent of the tile's max |a| plus one, so x = a * 2^-sigma is in
//   (-1, 1).  The first limb takes 8 bits, every later limb 9:
//     l_0 = rint(256 x),          x_1 = 256 x - l_0          (|x_1| <= 1/2)
//     l_s = rint(512 x_s),        x_{s+1} = 512 x_s - l_s    (s >= 1)
//   Every step is exact in FP64 and |l_s| <= 256, so each l_s is an integer
//   that BF16 holds exactly (8 significant bits).  Limb s is stored as
//   L_s = l_s 2^-9s (exact: only the exponent moves), so that
//     a = 2^(sigma-8) * sum_s L_s + r,   |r| <= 2^(sigma - 9S)
//   and a limb product L_s L_t carries its own level weight 2^-9(s+t).
//   Limbs are written in the layout the MFMA wants: the logical op(A) tile as
//   rows x K and op(B) as columns x K, K contiguous, rows padded to 16 and K
//   to 32 with zeros, so the GEMM needs no bounds checks and 16-byte loads.
//
// Multiplying (tamx_grouped_gemm_split): for a tile pair the limb products
// with s + t < S (the triangular set, S(S+1)/2 MFMA products) go to three
// FP32 accumulators:
//   P_0 = sum L_0 L_0       integers, |.| <= 2^16 per k
//   P_1 = sum L_0 L_1 + L_1 L_0   multiples of 2^-9, |.| <= 2^8 per k
//   P_T = sum of the levels s + t >= 2
// Over 128 consecutive k, P_0 and P_1 stay within 24 significant bits, so
// both are exact; P_T rounds, but at 2^-24 of a level that is itself 2^-18
// below P_0.  Every 128 k and at each pair boundary the three are folded in
// FP64:  C += 2^(sigma_A + sigma_B - 16) * (P_0 + P_1 + P_T).
// The result error is the truncation of the limb expansion (about 2^-9S
// relative to the tiles' maxima, from the dropped remainders and the dropped
// levels s + t >= S) plus FP64 rounding.
#include <algorithm>
#include <new>
#include <type_traits>
#include <vector>

#include "tamx_internal.hpp"
#include "tamx_scan.hpp"
#include "tamx_split.hpp"

namespace {

using namespace tamx::split;
constexpr int SBM = 64, SBN = 64;  // macro tile of split_gemm_kernel

// The split passes walk 16-row x 128-k blocks of every tile's limb layout;
// thread t of a 256-thread workgroup owns row t / 16 and the 8 consecutive
// k from 8 (t % 16) (one 8-byte INT8 or 16-byte BF16 store per limb
// plane), so the 16 threads of a row write whole 128-byte lines of every
// plane.  k_pad (a multiple of 32 or 64) need not be a multiple of 128: the
// threads past it load zeros and store nothing.
constexpr int kSplitBR = 16, kSplitBK = 128, kSplitTQ = kSplitBK / 8;
// LDS staging: element (row r, k) at r * kSplitRS + k + k / 8; the pad
// double after every 8 k spreads a row's 16 thread groups over all banks
constexpr int kSplitRS = kSplitBK + kSplitBK / 8 + 1;
static_assert(kSplitBR * kSplitBK == 256 * 8 && 256 % kSplitBK == 0 && 256 % kSplitBR == 0,
              "one 8-element group per thread");

__host__ __device__ inline int split_blocks_k(int k_pad) { return (k_pad + kSplitBK - 1) / kSplitBK; }

__device__ inline int64_t find_tile(const int64_t* prefix, int64_t n, int64_t item) {
  int64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The contiguous range [b, e) of 16 x 128 blocks of this workgroup and the
// tile of its first block: the tile table is searched once per workgroup,
// then followed forward (tile t covers blocks [prefix[t], prefix[t+1])).
__device__ inline void block_range(const int64_t* prefix, int64_t ntiles, int64_t& b,
                                   int64_t& e, int64_t& ti) {
  const int64_t n = prefix[ntiles];
  b = n * blockIdx.x / gridDim.x;
  e = n * (blockIdx.x + 1) / gridDim.x;
  ti = b < e ? find_tile(prefix, ntiles, b) : 0;
}

__global__ void zero_u64_kernel(unsigned long long* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Offset of index x over n modes (extents ext, strides str, the last fastest).
__device__ inline int64_t mixed_radix_offset(int x, int n, const int32_t* ext,
                                             const int64_t* str) {
  int64_t off = 0;
  for (int d = n - 1; d >= 0; --d) {
    const int e = ext[d];
    const int q = x / e;
    off += (int64_t)(x - q * e) * str[d];
    x = q;
  }
  return off;
}

// A block is read with the lanes along the stored rows (a transposed tile,
// or a general one whose last row mode is the source's stride-1 mode:
// element j of thread tid is row tid % 16, k tid / 16 + 16 j) or along k
// (element j: k tid % 128, row tid / 128 + 2 j), so a wave's global reads
// are contiguous, and turned into the thread layout (row tid / 16, k
// 8 (tid % 16) .. + 8) through LDS.
__device__ inline bool lanes_along_rows(const SplitTile& t) {
  return t.gen_nr > 0 ? t.gr_str[t.gen_nr - 1] == 1 : t.transpose != 0;
}

// Global loads of thread tid's 8 elements of block (r0, k0), zero outside
// the tile.
__device__ inline void fetch_block8(const double* __restrict__ pool, const SplitTile& t, int r0,
                                    int k0, double (&x)[8]) {
  const int tid = threadIdx.x;
  if (t.gen_nr > 0) {
    if (lanes_along_rows(t)) {
      const int gr = r0 + tid % kSplitBR;
      const bool row_ok = gr < t.out_rows;
      const int64_t base =
          t.src_off + (row_ok ? mixed_radix_offset(gr, t.gen_nr, t.gr_ext, t.gr_str) : 0);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int gk = k0 + tid / kSplitBR + (256 / kSplitBR) * j;
        x[j] = row_ok && gk < t.out_k
                   ? pool[base + mixed_radix_offset(gk, t.gen_nk, t.gk_ext, t.gk_str)]
                   : 0.0;
      }
    } else {
      const int gk = k0 + tid % kSplitBK;
      const bool k_ok = gk < t.out_k;
      const int64_t base =
          t.src_off + (k_ok ? mixed_radix_offset(gk, t.gen_nk, t.gk_ext, t.gk_str) : 0);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int gr = r0 + tid / kSplitBK + (256 / kSplitBK) * j;
        x[j] = k_ok && gr < t.out_rows
                   ? pool[base + mixed_radix_offset(gr, t.gen_nr, t.gr_ext, t.gr_str)]
                   : 0.0;
      }
    }
  } else if (t.transpose) {
    const int gr = r0 + tid % kSplitBR;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int gk = k0 + tid / kSplitBR + (256 / kSplitBR) * j;
      x[j] = gk < t.out_k && gr < t.out_rows ? pool[t.src_off + (int64_t)gk * t.cols + gr] : 0.0;
    }
  } else {
    const int gk = k0 + tid % kSplitBK;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int gr = r0 + tid / kSplitBK + (256 / kSplitBK) * j;
      x[j] = gk < t.out_k && gr < t.out_rows ? pool[t.src_off + (int64_t)gr * t.cols + gk] : 0.0;
    }
  }
}

// The fetched elements through LDS into the thread layout: v[i] is row
// tid / 16, k 8 (tid % 16) + i of the block.  Every thread of the block
// calls it (two barriers).
__device__ inline int stg_index(int r, int k) { return r * kSplitRS + k + k / 8; }

__device__ inline void stage_block8(const SplitTile& t, double* stg, const double (&x)[8],
                                    double (&v)[8]) {
  const int tid = threadIdx.x;
  if (lanes_along_rows(t)) {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      stg[stg_index(tid % kSplitBR, tid / kSplitBR + (256 / kSplitBR) * j)] = x[j];
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      stg[stg_index(tid / kSplitBK + (256 / kSplitBK) * j, tid % kSplitBK)] = x[j];
  }
  __syncthreads();
  const int r = tid / kSplitTQ, q = tid % kSplitTQ;
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = stg[stg_index(r, 8 * q + i)];
  __syncthreads();
}

__device__ inline void block_of(const SplitTile& t, int64_t local, int& r0, int& k0) {
  const int nbk = split_blocks_k(t.k_pad);
  r0 = (int)(local / nbk) * kSplitBR;
  k0 = (int)(local % nbk) * kSplitBK;
}

// Runs body(tile, tile index, r0, k0, v) over the workgroup's contiguous
// range of blocks, v being thread tid's 8 elements (row tid >> 2, k
// 8 (tid & 3) + i).  The loads of the next block are issued before the body
// of the current one runs, so they are in flight during its arithmetic and
// stores.  Every thread runs every body (the bodies may hold barriers).
template <class Body>
__device__ inline void for_each_block8(const double* __restrict__ pool,
                                       const SplitTile* __restrict__ tiles,
                                       const int64_t* __restrict__ prefix, int64_t ntiles,
                                       double* stg, Body&& body) {
  int64_t b, e, ti;
  block_range(prefix, ntiles, b, e, ti);
  if (b >= e) return;
  while (prefix[ti + 1] <= b) ++ti;
  int r0, k0;
  block_of(tiles[ti], b - prefix[ti], r0, k0);
  double x[8];
  fetch_block8(pool, tiles[ti], r0, k0, x);
  for (int64_t item = b; item < e; ++item) {
    const int64_t cti = ti;
    const int cr0 = r0, ck0 = k0;
    double v[8];
    stage_block8(tiles[cti], stg, x, v);
    if (item + 1 < e) {
      while (prefix[ti + 1] <= item + 1) ++ti;
      block_of(tiles[ti], item + 1 - prefix[ti], r0, k0);
      fetch_block8(pool, tiles[ti], r0, k0, x);
    }
    body(tiles[cti], cti, cr0, ck0, v);
  }
}

// max |a| per tile; positive doubles order like their bits
__global__ __launch_bounds__(256) void split_max_kernel(
    const double* __restrict__ pool, const SplitTile* __restrict__ tiles,
    const int64_t* __restrict__ prefix, int64_t ntiles,
    unsigned long long* __restrict__ maxbits) {
  __shared__ double stg[kSplitBR * kSplitRS];
  __shared__ double red[256];
  for_each_block8(pool, tiles, prefix, ntiles, stg,
                  [&](const SplitTile&, int64_t ti, int, int, const double (&v)[8]) {
    double m = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) m = fmax(m, fabs(v[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] > 0.0)
      atomicMax(&maxbits[ti], (unsigned long long)__double_as_longlong(red[0]));
    __syncthreads();
  });
}

// maxabs (when given) holds each tile's max |a|, else maxbits its bits
__global__ void split_sigma_kernel(SplitTile* tiles, const unsigned long long* maxbits,
                                   const double* maxabs, int64_t ntiles, int i8) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const double m = maxabs ? maxabs[t] : __longlong_as_double((long long)maxbits[t]);
  // x = a * 2^-sigma lies in (-1, 1): m < 2^(ilogb(m) + 1); INT8 limbs
  // also need |x| <= 127/128 so that rint(128 x) fits a signed byte
  int sigma = m > 0.0 ? ilogb(m) + 1 : 0;
  if (i8 && m > 0.0 && ldexp(m, 7 - sigma) > 127.0) ++sigma;
  tiles[t].sigma = sigma;
}

// Writes the S limb planes of every 16 x 128 block of the limb layout.
template <int S>
__global__ __launch_bounds__(256) void split_limbs_kernel(
    const double* __restrict__ pool, const SplitTile* __restrict__ tiles,
    const int64_t* __restrict__ prefix, int64_t ntiles, bf8_t* __restrict__ limbs) {
#pragma clang fp contract(off)
  __shared__ double stg[kSplitBR * kSplitRS];
  const int r = threadIdx.x / kSplitTQ, q = threadIdx.x % kSplitTQ;
  for_each_block8(pool, tiles, prefix, ntiles, stg,
                  [&](const SplitTile& t, int64_t, int r0, int k0, const double (&v)[8]) {
    if (r0 + r >= t.rows_pad || k0 + 8 * q >= t.k_pad) return;
    const int sigma = t.sigma;
    bf8_t out[S];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double xs0 = ldexp(v[i], 8 - sigma);  // 256 * a * 2^-sigma
      double l = rint(xs0);
      out[0][i] = (__bf16)(float)l;
      double rem = xs0 - l;
#pragma unroll
      for (int s = 1; s < S; ++s) {
        const double xs = rem * 512.0;
        l = rint(xs);
        out[s][i] = (__bf16)(float)ldexp(l, -9 * s);
        rem = xs - l;
      }
    }
    // in units of 8 elements
    const int64_t plane = ((int64_t)t.rows_pad * t.k_pad) >> 3;
    const int64_t o = (t.limb_off + (int64_t)(r0 + r) * t.k_pad + k0 + 8 * q) >> 3;
#pragma unroll
    for (int s = 0; s < S; ++s) limbs[o + s * plane] = out[s];
  });
}

// INT8 limbs: x = a 2^-sigma with |x| <= 127/128; l_0 = rint(128 x) (|l_0|
// <= 127), x_1 = 128 x - l_0 (|x_1| <= 1/2); l_s = rint(128 x_s) (|l_s| <=
// 64), x_{s+1} = 128 x_s - l_s.  Every step is exact in FP64, so
// a = 2^(sigma-7) sum_s l_s 2^-7s + r with |r| <= 2^(sigma - 7 S - 1).  A
// thread writes its 8 consecutive k of each limb as one 8-byte store.
//
// Rounding: for |y| < 2^51, y + 1.5 * 2^52 lies in [2^52, 2^53), where the
// spacing of doubles is 1, so the sum is y rounded to an integer (ties to
// even, as rint), and because 1.5 * 2^52 is even and the exponent does not
// change, the low word of the sum's bits is that integer in two's
// complement.  So each limb costs one FMA (round), one subtraction (the
// limb as a double) and one FMA (the exact remainder), with no conversion.
constexpr double kRoundShift = 6755399441055744.0;  // 1.5 * 2^52

__device__ inline uint32_t low_bytes4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  return (a & 0xffu) | ((b & 0xffu) << 8) | ((c & 0xffu) << 16) | (d << 24);
}

template <int S>
__global__ __launch_bounds__(256) void split_limbs_i8_kernel(
    const double* __restrict__ pool, const SplitTile* __restrict__ tiles,
    const int64_t* __restrict__ prefix, int64_t ntiles, uint64_t* __restrict__ limbs) {
#pragma clang fp contract(off)
  __shared__ double stg[kSplitBR * kSplitRS];
  const int r = threadIdx.x / kSplitTQ, q = threadIdx.x % kSplitTQ;
  for_each_block8(pool, tiles, prefix, ntiles, stg,
                  [&](const SplitTile& t, int64_t, int r0, int k0, const double (&v)[8]) {
    if (r0 + r >= t.rows_pad || k0 + 8 * q >= t.k_pad) return;
    const int sigma = t.sigma;
    uint32_t w[S][8];  // limb s of element i in the low byte of w[s][i]
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double xs0 = ldexp(v[i], 7 - sigma);  // 128 * a * 2^-sigma
      double tr = xs0 + kRoundShift;
      double l = tr - kRoundShift;  // rint(xs0)
      w[0][i] = (uint32_t)__double2loint(tr);
      double rem = xs0 - l;
#pragma unroll
      for (int s = 1; s < S; ++s) {
        tr = __fma_rn(rem, 128.0, kRoundShift);  // 128 rem is exact
        l = tr - kRoundShift;                     // rint(128 rem)
        w[s][i] = (uint32_t)__double2loint(tr);
        rem = __fma_rn(rem, 128.0, -l);           // exact: |128 rem - l| <= 1/2
      }
    }
    uint64_t out[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
      out[s] = (uint64_t)low_bytes4(w[s][0], w[s][1], w[s][2], w[s][3]) |
               (uint64_t)low_bytes4(w[s][4], w[s][5], w[s][6], w[s][7]) << 32;
    // in units of 8 limbs, in 32-k blocks (i8_unit)
    const int64_t plane = ((int64_t)t.rows_pad * t.k_pad) >> 3;
    const int kk = k0 + 8 * q;
    const int64_t o = (t.limb_off + ((int64_t)(kk / kI8Block) * t.rows_pad + r0 + r) * kI8Block +
                       kk % kI8Block) >> 3;
#pragma unroll
    for (int s = 0; s < S; ++s) limbs[o + s * plane] = out[s];
  });
}

template <int S>
void launch_split(int grid, hipStream_t s, const double* pool, const tamx_split_plan* p,
                  void* limbs) {
  if (p->limb_type == TAMX_LIMB_I8)
    split_limbs_i8_kernel<S><<<grid, 256, 0, s>>>(pool, p->tiles, p->block_prefix, p->ntiles,
                                                  static_cast<uint64_t*>(limbs));
  else
    split_limbs_kernel<S><<<grid, 256, 0, s>>>(pool, p->tiles, p->block_prefix, p->ntiles,
                                               static_cast<bf8_t*>(limbs));
}

// ------------------------------------------------------------ the GEMM --

// I8: INT8 limbs (7 bits each) on the INT8 MFMA 16x16x64 with one exact
// INT32 accumulator per level s + t; otherwise BF16 limbs as above.  Both
// stage 16-byte units (8 BF16 or 16 INT8 limbs of one row), 4 units of a
// row per chunk, so the staging and the LDS layout are shared.
// M32 (INT8 only): each wave's 32 x 32 quadrant is one block of the INT8
// MFMA 32x32x32 instead of four of the 16x16x64 (the same products and
// levels; on this part the 32x32x32 does twice the work in about the same
// issue time).
// WM: waves along the rows of the macro tile, 2 (64 x 64, four waves) or 4
// (128 x 64, eight waves in one workgroup per CU: every B limb staged into
// LDS feeds twice the rows, a quarter fewer limb bytes per product; INT8
// limbs from three on, whose chunks are loaded straight into LDS).
template <int S, bool I8, bool M32 = false, int WM = 2>
__global__ __launch_bounds__(128 * WM, I8 ? 2 : 1) void split_gemm_kernel(SplitParams p) {
  static_assert(!M32 || I8, "the 32x32x32 blocks are INT8 only");
  static_assert(WM == 2 || (WM == 4 && I8 && !M32 && S > 2), "eight waves: INT8 16x16 blocks, 3+ limbs");
  constexpr int BM = 32 * WM;  // macro tile rows
  constexpr int NG = I8 ? S : (S >= 3 ? 3 : S);  // BF16: P_0, P_1, P_T; INT8: levels
  constexpr int CK = I8 ? SBK8 : SBK;            // k per chunk
  constexpr int USH = I8 ? 4 : 3;                // log2 limbs per 16-byte unit
  constexpr int FOLD = I8 ? kFoldChunks8 : kFoldChunks;
  using lvl_t = std::conditional_t<I8, i4_t, f4_t>;
  // LDS holds only the limb chunks (S * 16 KB); the grabbed item is passed
  // through the first word of sA while the buffers are idle between items
  // (one variable per buffer: the compiler then knows that loads landing in
  // one buffer do not alias reads of the other, and waits for them only
  // where the buffer they fill is read)
  __shared__ bf8_t sA0[S][BM * 4], sA1[S][BM * 4];
  __shared__ bf8_t sB0[S][SBN * 4], sB1[S][SBN * 4];
  int* s_item = reinterpret_cast<int*>(&sA0[0][0]);
  using zero = std::integral_constant<int, 0>;
  using one = std::integral_constant<int, 1>;
  using two = std::integral_constant<int, 2>;
  auto bufA = [&](auto B) -> bf8_t(&)[S][BM * 4] {
    if constexpr (decltype(B)::value == 0) return sA0; else return sA1;
  };
  auto bufB = [&](auto B) -> bf8_t(&)[S][SBN * 4] {
    if constexpr (decltype(B)::value == 0) return sB0; else return sB1;
  };

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = uni(tid >> 6);
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lrow = lane & 15, lq = lane >> 4;  // MFMA: row lrow, k unit lq
  // staging: thread loads row (tid / 4), unit (tid % 4) of every plane
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
    const tamx_gemm_task* tp = p.tasks + lo;
    const int64_t t_coff = uni64(tp->c_off);
    const int t_m = uni(tp->m), t_n = uni(tp->n), t_ldc = uni(tp->ldc);
    const int t_pb = uni(tp->pair_begin), t_pc = uni(tp->pair_count);
    const int t_flags = uni(tp->flags);
    const int local = item - uni(p.prefix[lo]);
    const int tiles_n = (t_n + SBN - 1) / SBN;
    const int r0 = (local / tiles_n) * BM, c0 = (local % tiles_n) * SBN;
    const int mrem = min(BM, t_m - r0), nrem = min(SBN, t_n - c0);
    const bool wave_active = (wr < mrem) && (wc < nrem);

    // 16 x 16 blocks (acc, lvl) or one 32 x 32 block (acc32, lvl32)
    constexpr int NB = M32 ? 1 : 2;
    d4_t acc[NB][NB];
    lvl_t lvl[M32 ? 1 : NG][NB][NB];
    double acc32[M32 ? 16 : 1];
    i16_t lvl32[M32 ? NG : 1];
    if constexpr (M32) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        acc32[v] = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g) lvl32[g][v] = 0;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[i][j][v] = 0.0;
#pragma unroll
          for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int v = 0; v < 4; ++v) lvl[g][i][j][v] = 0;
        }
    }

    const int pend = t_pb + t_pc;

    // load cursor: pair lp, chunk lk0; runs two chunks ahead of compute
    int lp = t_pb, lk0 = 0, lkp = 0;
    int64_t la_row = 0, lb_row = 0, la_plane = 0, lb_plane = 0, la_blk = 0, lb_blk = 0;
    auto set_pair = [&](int pi) {
      const auto& pp = tamx::scalar_ptr(p.pairs)[pi];
      const int at = uni(pp.a_tile), bt = uni(pp.b_tile);
      const auto* A = tamx::scalar_ptr(p.ta) + at;
      const auto* B = tamx::scalar_ptr(p.tb) + bt;
      lkp = uni(A->k_pad);
      // limb offsets of this thread's staging row in plane 0, in 16-byte
      // units (INT8: at k 0 of the 32-k block layout, and the units per block)
      la_plane = uni64((int64_t)uni(A->rows_pad) * lkp) >> USH;
      lb_plane = uni64((int64_t)uni(B->rows_pad) * lkp) >> USH;
      if constexpr (I8) {
        la_blk = 2 * uni(A->rows_pad);
        lb_blk = 2 * uni(B->rows_pad);
        la_row = i8_unit(uni64(A->limb_off) >> 4, 0, r0 + st_r, 0);
        lb_row = i8_unit(uni64(B->limb_off) >> 4, 0, c0 + st_r, 0);
      } else {
        la_row = (uni64(A->limb_off) + (int64_t)(r0 + st_r) * lkp) >> USH;
        lb_row = (uni64(B->limb_off) + (int64_t)(c0 + st_r) * lkp) >> USH;
      }
    };
    if (lp < pend) set_pair(lp);
    // unit q of the chunk at lk0 in this thread's staging row of plane 0
    auto a_unit = [&](int q) -> int64_t {
      if constexpr (I8) return la_row + ((lk0 / kI8Block) + (q >> 1)) * la_blk + (q & 1);
      else return la_row + (lk0 >> USH) + q;
    };
    auto b_unit = [&](int q) -> int64_t {
      if constexpr (I8) return lb_row + ((lk0 / kI8Block) + (q >> 1)) * lb_blk + (q & 1);
      else return lb_row + (lk0 >> USH) + q;
    };
    // Once the cursor has passed the last pair the loads repeat its first
    // chunk, unused: every step issues the same loads, which keeps the wait
    // counts before the LDS stores exact (a conditional load would make the
    // compiler wait for all loads in flight).
    auto load_chunk = [&](bf8_t (&ra)[S], bf8_t (&rb)[S]) {
      const int64_t ua = a_unit(st_q), ub = b_unit(st_q);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        ra[s] = p.LA[ua + s * la_plane];
        rb[s] = p.LB[ub + s * lb_plane];
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
    // The same chunk loaded straight into LDS buffer `buf` (no registers):
    // lane l of wave w writes LDS unit 64 w + l of each plane, which is the
    // swizzled slot lds_unit(st_r, q) of row st_r = tid / 4 for the unit
    // q = (tid % 4) ^ ((st_r / 2) % 4) of that row.  The caller waits for
    // the loads (vmcnt) before the barrier that precedes reading them.
    const int st_qs = st_q ^ ((st_r >> 1) & 3);
    auto load_chunk_lds = [&](auto buf) {
      const int64_t ua = a_unit(st_qs), ub = b_unit(st_qs);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        __builtin_amdgcn_global_load_lds(
            (const __attribute__((address_space(1))) void*)(p.LA + ua + s * la_plane),
            (__attribute__((address_space(3))) void*)(&bufA(buf)[s][wave * 64]), 16, 0, 0);
        // (eight waves: the first four stage B's 64 rows)
        if (WM == 2 || wave < 4)
          __builtin_amdgcn_global_load_lds(
              (const __attribute__((address_space(1))) void*)(p.LB + ub + s * lb_plane),
              (__attribute__((address_space(3))) void*)(&bufB(buf)[s][wave * 64]), 16, 0, 0);
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
    auto store_chunk = [&](auto buf, const bf8_t (&ra)[S], const bf8_t (&rb)[S]) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        bufA(buf)[s][lds_unit(st_r, st_q)] = ra[s];
        bufB(buf)[s][lds_unit(st_r, st_q)] = rb[s];
      }
    };
    // NI x NJ of the wave's 2 x 2 blocks hold rows and columns of the tile
    // (a ragged edge leaves the rest empty; the split is uniform per wave)
    const int ni = wr + 16 < mrem ? 2 : 1, nj = wc + 16 < nrem ? 2 : 1;
    auto compute_blocks = [&](auto buf, auto NI_, auto NJ_) {
      constexpr int NI = decltype(NI_)::value, NJ = decltype(NJ_)::value;
      bf8_t bf[NJ][S];
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int t = 0; t < S; ++t) bf[j][t] = bufB(buf)[t][lds_unit(wc + j * 16 + lrow, lq)];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        bf8_t af[NI];
#pragma unroll
        for (int i = 0; i < NI; ++i) af[i] = bufA(buf)[s][lds_unit(wr + i * 16 + lrow, lq)];
#pragma unroll
        for (int t = 0; t < S - s; ++t) {
          const int g = I8 ? s + t : (s + t < 2 ? s + t : 2);
#pragma unroll
          for (int i = 0; i < NI; ++i)
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
    };
    auto compute_chunk = [&](auto buf) {
      if (!wave_active) return;
      if constexpr (M32) {
        // the quadrant as one 32 x 32 block; two 32-k steps per chunk,
        // lane l: row (col) l % 32, k unit 2 h + l / 32
        (void)compute_blocks;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          i4_t bf[S];
#pragma unroll
          for (int t = 0; t < S; ++t)
            bf[t] = __builtin_bit_cast(
                i4_t, bufB(buf)[t][lds_unit(wc + (lane & 31), 2 * h + (lane >> 5))]);
#pragma unroll
          for (int s = 0; s < S; ++s) {
            const i4_t af = __builtin_bit_cast(
                i4_t, bufA(buf)[s][lds_unit(wr + (lane & 31), 2 * h + (lane >> 5))]);
#pragma unroll
            for (int t = 0; t < S - s; ++t)
              lvl32[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf[t], lvl32[s + t], 0, 0, 0);
          }
        }
        return;
      }
      if (ni == 2 && nj == 2) compute_blocks(buf, two{}, two{});
      else if (ni == 2) compute_blocks(buf, two{}, one{});
      else if (nj == 2) compute_blocks(buf, one{}, two{});
      else compute_blocks(buf, one{}, one{});
    };
    // fold the level accumulators of the pair being computed (scale
    // 2^(sigma_A + sigma_B - 16), BF16; 2^(sigma_A + sigma_B - 14) with level
    // weights 2^-7g, INT8)
    auto fold = [&](double scale) {
      if constexpr (M32) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          double tsum = 0.0;
#pragma unroll
          for (int g = NG - 1; g >= 0; --g) {
            tsum = fma((double)lvl32[g][v], 1.0 / (double)(1LL << (7 * g)), tsum);
            lvl32[g][v] = 0;
          }
          acc32[v] = fma(tsum, scale, acc32[v]);
        }
        return;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            double tsum = 0.0;
#pragma unroll
            for (int g = NG - 1; g >= 0; --g) {
              // (INT8: the level weight 2^-7g is exact, so the fma rounds
              // once, as the sum would)
              if constexpr (I8) tsum = fma((double)lvl[g][i][j][v], 1.0 / (double)(1LL << (7 * g)), tsum);
              else tsum += (double)lvl[g][i][j][v];
              lvl[g][i][j][v] = 0;
            }
            acc[i][j][v] = fma(tsum, scale, acc[i][j][v]);
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
    // after computing one chunk: fold at a pair end or every kFoldChunks
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

    // Pipeline: registers R0/R1 hold chunks c+1 and c+2 in flight while
    // chunk c is multiplied out of LDS buffer c % 2.  One barrier per chunk.
    // Above four limbs the second register set would cost occupancy, so the
    // loads run one chunk ahead instead.
    bf8_t ra0[S], rb0[S], ra1[S], rb1[S];
    // Above four BF16 limbs, and from three INT8 limbs on, a second
    // register set would cost occupancy: the loads run one chunk ahead,
    // straight into LDS (measured: two chunks ahead through registers is
    // faster where it fits, 49.7 against 47.7 TFLOP/s for four BF16 limbs).
    // (The 32x32x32 blocks need fewer accumulator registers: up to four
    // INT8 limbs they keep two register sets at two waves per SIMD.)
    constexpr bool ONE_AHEAD = I8 ? (M32 ? S > 4 : S > 2) : S > 4;
    if constexpr (ONE_AHEAD) {
      if (lp < pend) {
        load_chunk_lds(zero{});
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (;;) {
          bool has_next = lp < pend;
          if (has_next) load_chunk_lds(one{});  // chunk c+1 -> the idle buffer
          compute_chunk(zero{});
          after_compute();
          if (!has_next) break;
          __builtin_amdgcn_s_waitcnt(0);
          __syncthreads();
          has_next = lp < pend;
          if (has_next) load_chunk_lds(zero{});
          compute_chunk(one{});
          after_compute();
          if (!has_next) break;
          __builtin_amdgcn_s_waitcnt(0);
          __syncthreads();
        }
      }
    } else if (lp < pend) {
      load_chunk(ra0, rb0);                // chunk 0
      const bool have1 = lp < pend;
      load_chunk(ra1, rb1);                // chunk 1 (if have1)
      store_chunk(zero{}, ra0, rb0);
      __syncthreads();
      bool more = have1;  // chunk c + 1 exists
      for (;;) {
        bool next2 = lp < pend;
        load_chunk(ra0, rb0);              // chunk c + 2 -> R0 (if next2)
        compute_chunk(zero{});
        after_compute();
        if (!more) break;
        store_chunk(one{}, ra1, rb1);
        __syncthreads();
        more = next2;
        next2 = lp < pend;
        load_chunk(ra1, rb1);              // chunk c + 2 -> R1 (if next2)
        compute_chunk(one{});
        after_compute();
        if (!more) break;
        store_chunk(zero{}, ra0, rb0);
        __syncthreads();
        more = next2;
      }
    }

    const bool beta1 = (t_flags & TAMX_TASK_BETA1) != 0;
    double* __restrict__ c = p.C + t_coff;
    tamx::with_result_index(p.c_scatter, p.task_tile, lo, t_flags, t_ldc, [&](auto cidx) {
      if constexpr (M32) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          // 32x32 MFMA output: row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32
          const int row = wr + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
          const int col = wc + (lane & 31);
          if (row < mrem && col < nrem) {
            const int64_t idx = cidx(r0 + row, c0 + col);
            double val = p.alpha * acc32[v];
            if (beta1) val += c[idx];
            c[idx] = val;
          }
        }
        return;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
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

template <int S, bool I8>
int launch_split_gemm(tamx_context* ctx, tamx_worklist* wl, const SplitParams& prm,
                      hipStream_t s, bool m32 = false, bool rows128 = false) {
  auto kern = split_gemm_kernel<S, I8>;
  int threads = kThreads, bm = SBM;
  if constexpr (I8) {
    if (m32) kern = split_gemm_kernel<S, I8, true>;
    if constexpr (S > 2)
      if (rows128) {
        kern = split_gemm_kernel<S, I8, false, 4>;
        threads = 2 * kThreads;
        bm = 2 * SBM;
      }
  }
  int occ = 0;
  TAMX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, threads, 0));
  if (occ < 1) occ = 1;
  if (int rc = tamx::plan_items(wl, bm, SBN, s)) return rc;
  kern<<<ctx->num_cus * occ, threads, 0, s>>>(prm);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}


// ---------------------------------------------------- FP32 two limbs --
//
// FP64 tiles are split while they are staged: hi = fp32(a), lo = fp32(a - hi)
// (a = hi + lo + O(2^-48 a)).  Per 16-k chunk the products hi*hi and
// hi*lo + lo*hi run on FP32 MFMA (16x16x4) with FP32 accumulation, then are
// folded into FP64.  FP32 accumulation rounds, so this variant carries about
// 24 bits; it is reported beside the BF16-limb path, not used for accuracy.

typedef float f4_t2 __attribute__((ext_vector_type(4)));

template <bool TA, bool TB>
__global__ __launch_bounds__(kThreads) void f32x2_gemm_kernel(
    const double* __restrict__ A, const double* __restrict__ B, SplitParams p) {
  constexpr int BM = 64, BN = 64, BK = 16;
  constexpr int A_LD = BM * BK / kThreads, B_LD = BK * BN / kThreads;
  constexpr int A_STEP = TA ? kThreads / BM : kThreads / BK;
  constexpr int B_STEP = TB ? kThreads / BK : kThreads / BN;
  __shared__ float sAh[2][BK][BM + 1], sAl[2][BK][BM + 1];
  __shared__ float sBh[2][BK][BN + 1], sBl[2][BK][BN + 1];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = uni(tid >> 6);
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lrow = lane & 15, lk = lane >> 4;
  const int a_fast = TA ? tid % BM : tid % BK, a_slow = TA ? tid / BM : tid / BK;
  const int b_fast = TB ? tid % BK : tid % BN, b_slow = TB ? tid / BK : tid / BN;

  int64_t ntasks64 = p.counters[TAMX_WL_NTASKS];
  if (ntasks64 < 0 || ntasks64 > p.task_capacity || p.counters[TAMX_WL_STATUS] != 0)
    ntasks64 = 0;
  const int32_t ntasks = uni((int32_t)ntasks64);
  const int32_t total = uni(p.prefix[ntasks > 0 ? ntasks : 0]);
  const int home = blockIdx.x % kRegions;
  int tried = 0;
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
    const bool wave_active = (wr < mrem) && (wc < nrem);

    d4_t acc[2][2];
    f4_t2 phh[2][2], px[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          acc[i][j][v] = 0.0;
          phh[i][j][v] = 0.0f;
          px[i][j][v] = 0.0f;
        }

    int64_t a_base = 0, b_base = 0;
    int a_lda = 0, b_ldb = 0, pk = 0;
    int64_t a_off[A_LD], b_off[B_LD];
    bool a_ok[A_LD], b_ok[B_LD];
    auto set_pair = [&](int pi) {
      const tamx_gemm_pair& pp = p.pairs[pi];
      a_base = uni64(pp.a_off);
      b_base = uni64(pp.b_off);
      pk = uni(pp.k);
      a_lda = uni(pp.lda);
      b_ldb = uni(pp.ldb);
#pragma unroll
      for (int e = 0; e < A_LD; ++e) {
        const int slow = a_slow + e * A_STEP;
        a_ok[e] = TA ? a_fast < mrem : slow < mrem;
        a_off[e] = TA ? (int64_t)slow * a_lda + r0 + a_fast : (int64_t)(r0 + slow) * a_lda + a_fast;
      }
#pragma unroll
      for (int e = 0; e < B_LD; ++e) {
        const int slow = b_slow + e * B_STEP;
        b_ok[e] = TB ? slow < nrem : b_fast < nrem;
        b_off[e] = TB ? (int64_t)(c0 + slow) * b_ldb + b_fast : (int64_t)slow * b_ldb + c0 + b_fast;
      }
    };
    double ra[A_LD], rb[B_LD];
    auto load_chunk = [&](int k0) {
      const int kc = min(BK, pk - k0);
      const int64_t a_k = TA ? (int64_t)k0 * a_lda : k0;
      const int64_t b_k = TB ? k0 : (int64_t)k0 * b_ldb;
#pragma unroll
      for (int e = 0; e < A_LD; ++e) {
        const int kk = TA ? a_slow + e * A_STEP : a_fast;
        const bool ok = a_ok[e] && kk < kc;
        const double v = A[a_base + (ok ? a_off[e] + a_k : 0)];
        ra[e] = ok ? v : 0.0;
      }
#pragma unroll
      for (int e = 0; e < B_LD; ++e) {
        const int kk = TB ? b_fast : b_slow + e * B_STEP;
        const bool ok = b_ok[e] && kk < kc;
        const double v = B[b_base + (ok ? b_off[e] + b_k : 0)];
        rb[e] = ok ? v : 0.0;
      }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
      for (int e = 0; e < A_LD; ++e) {
        const float h = (float)ra[e];
        const float l = (float)(ra[e] - (double)h);
        const int kk = TA ? a_slow + e * A_STEP : a_fast;
        const int m = TA ? a_fast : a_slow + e * A_STEP;
        sAh[buf][kk][m] = h;
        sAl[buf][kk][m] = l;
      }
#pragma unroll
      for (int e = 0; e < B_LD; ++e) {
        const float h = (float)rb[e];
        const float l = (float)(rb[e] - (double)h);
        const int kk = TB ? b_fast : b_slow + e * B_STEP;
        const int n = TB ? b_slow + e * B_STEP : b_fast;
        sBh[buf][kk][n] = h;
        sBl[buf][kk][n] = l;
      }
    };
    auto compute_chunk = [&](int buf) {
      if (!wave_active) return;
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        float ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          ah[i] = sAh[buf][ks + lk][wr + i * 16 + lrow];
          al[i] = sAl[buf][ks + lk][wr + i * 16 + lrow];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          bh[j] = sBh[buf][ks + lk][wc + j * 16 + lrow];
          bl[j] = sBl[buf][ks + lk][wc + j * 16 + lrow];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            phh[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[i], bh[j], phh[i][j], 0, 0, 0);
            px[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[i], bl[j], px[i][j], 0, 0, 0);
            px[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[i], bh[j], px[i][j], 0, 0, 0);
          }
      }
      // fold this chunk into FP64
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            acc[i][j][v] += (double)phh[i][j][v] + (double)px[i][j][v];
            phh[i][j][v] = 0.0f;
            px[i][j][v] = 0.0f;
          }
    };

    int pi = t_pb;
    const int pend = t_pb + t_pc;
    if (pi < pend) {
      set_pair(pi);
      int k0 = 0;
      load_chunk(k0);
      store_chunk(0);
      __syncthreads();
      int buf = 0;
      k0 += BK;
      if (k0 >= pk) {
        ++pi;
        k0 = 0;
        if (pi < pend) set_pair(pi);
      }
      for (;;) {
        const bool has_next = pi < pend;
        if (has_next) load_chunk(k0);
        compute_chunk(buf);
        if (!has_next) break;
        store_chunk(buf ^ 1);
        __syncthreads();
        buf ^= 1;
        k0 += BK;
        if (k0 >= pk) {
          ++pi;
          k0 = 0;
          if (pi < pend) set_pair(pi);
        }
      }
    }
    const bool beta1 = (t_flags & TAMX_TASK_BETA1) != 0;
    double* __restrict__ c = p.C + t_coff;
    tamx::with_result_index(p.c_scatter, p.task_tile, lo, t_flags, t_ldc, [&](auto cidx) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = wr + i * 16 + lk * 4 + v;  // f32 16x16 MFMA layout
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

template <bool TA, bool TB>
int launch_f32x2(tamx_context* ctx, tamx_worklist* wl, const double* A, const double* B,
                 const SplitParams& prm, hipStream_t s) {
  auto kern = f32x2_gemm_kernel<TA, TB>;
  int occ = 0;
  TAMX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kThreads, 0));
  if (occ < 1) occ = 1;
  if (int rc = tamx::plan_items(wl, SBM, SBN, s)) return rc;
  kern<<<ctx->num_cus * occ, kThreads, 0, s>>>(A, B, prm);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

}  // namespace

namespace {

int split_f64(tamx_context* ctx, const tamx_split_plan* plan, const double* pool,
              const double* tile_maxabs, void* limbs, void* stream) {
  TAMX_REQUIRE(plan != nullptr && limbs != nullptr, "null argument");
  if (plan->ntiles == 0) return TAMX_OK;
  TAMX_REQUIRE(pool != nullptr || plan->nblocks == 0, "null pool");
  if (int rc = tamx::activate(ctx)) return rc;
  hipStream_t s = tamx::as_stream(stream);
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(plan->nblocks, ctx->num_cus * 16));
  const unsigned tgrid = (unsigned)((plan->ntiles + 255) / 256);
  if (!tile_maxabs) {
    zero_u64_kernel<<<tgrid, 256, 0, s>>>(plan->maxbits, plan->ntiles);
    TAMX_LAUNCH_CHECK();
    if (plan->nblocks > 0) {
      split_max_kernel<<<grid, 256, 0, s>>>(pool, plan->tiles, plan->block_prefix, plan->ntiles,
                                            plan->maxbits);
      TAMX_LAUNCH_CHECK();
    }
  }
  split_sigma_kernel<<<tgrid, 256, 0, s>>>(plan->tiles, plan->maxbits, tile_maxabs, plan->ntiles,
                                           plan->limb_type == TAMX_LIMB_I8 ? 1 : 0);
  TAMX_LAUNCH_CHECK();
  if (plan->nblocks > 0) {
    switch (plan->nslices) {
      case 2: launch_split<2>(grid, s, pool, plan, limbs); break;
      case 3: launch_split<3>(grid, s, pool, plan, limbs); break;
      case 4: launch_split<4>(grid, s, pool, plan, limbs); break;
      case 5: launch_split<5>(grid, s, pool, plan, limbs); break;
      default: launch_split<6>(grid, s, pool, plan, limbs); break;
    }
    TAMX_LAUNCH_CHECK();
  }
  return TAMX_OK;
}

}  // namespace

extern "C" {

int tamx_split_plan_create(tamx_context* ctx, int nslices, int is_right, int op,
                           int64_t ntiles, const int64_t* host_offsets,
                           const int32_t* host_rows, const int32_t* host_cols,
                           tamx_split_plan** out) {
  return tamx_split_plan_create_typed(ctx, TAMX_LIMB_BF16, nslices, is_right, op, ntiles,
                                      host_offsets, host_rows, host_cols, out);
}

int tamx_split_plan_create_typed(tamx_context* ctx, int limb_type, int nslices, int is_right,
                                 int op, int64_t ntiles, const int64_t* host_offsets,
                                 const int32_t* host_rows, const int32_t* host_cols,
                                 tamx_split_plan** out) {
  TAMX_REQUIRE(out != nullptr, "null output pointer");
  *out = nullptr;
  TAMX_REQUIRE(limb_type == TAMX_LIMB_BF16 || limb_type == TAMX_LIMB_I8,
               "limb_type must be TAMX_LIMB_BF16 or TAMX_LIMB_I8");
  TAMX_REQUIRE(nslices >= 2 && nslices <= 6, "nslices must be in [2, 6]");
  const int kq = limb_type == TAMX_LIMB_I8 ? SBK8 : SBK;  // k padding quantum
  TAMX_REQUIRE(op == 0 || op == 1, "op must be TAMX_NO_TRANS or TAMX_TRANS");
  TAMX_REQUIRE(ntiles >= 0, "negative tile count");
  TAMX_REQUIRE(ntiles == 0 || (host_offsets && host_rows && host_cols), "null tile table");
  if (int rc = tamx::activate(ctx)) return rc;
  // A (left): limb rows = m.  NoTrans stores m x k (no transpose); Trans
  // stores k x m.  B (right): limb rows = n.  NoTrans stores k x n
  // (transpose); Trans stores n x k.
  const int transpose = is_right ? (op == 0) : (op == 1);
  std::vector<SplitTile> tiles(ntiles);
  std::vector<int64_t> prefix(ntiles + 1, 0);
  int64_t off = 0, blocks = 0, max_kpad = 32;
  for (int64_t t = 0; t < ntiles; ++t) {
    SplitTile& st = tiles[t];
    st = SplitTile{};
    st.src_off = host_offsets[t];
    st.rows = host_rows[t];
    st.cols = host_cols[t];
    st.transpose = transpose;
    prefix[t] = blocks;
    if (st.src_off < 0) {  // absent tile: no limbs
      st.limb_off = 0;
      continue;
    }
    TAMX_REQUIRE(st.rows >= 0 && st.cols >= 0, "negative tile extent");
    st.out_rows = transpose ? st.cols : st.rows;
    st.out_k = transpose ? st.rows : st.cols;
    st.rows_pad = (st.out_rows + 15) / 16 * 16;
    st.k_pad = std::max(kq, (st.out_k + kq - 1) / kq * kq);
    st.limb_off = off;
    off += (int64_t)nslices * st.rows_pad * st.k_pad;
    max_kpad = std::max<int64_t>(max_kpad, st.k_pad);
    blocks += (int64_t)((st.rows_pad + kSplitBR - 1) / kSplitBR) * split_blocks_k(st.k_pad);
  }
  prefix[ntiles] = blocks;
  auto* plan = new (std::nothrow) tamx_split_plan();
  if (!plan) return TAMX_ERR_OOM;
  plan->ctx = ctx;
  plan->ntiles = ntiles;
  plan->nslices = nslices;
  plan->limb_type = limb_type;
  plan->limb_elems = off + (int64_t)kLimbSlackRows * max_kpad * nslices;
  plan->nblocks = blocks;
  hipError_t e = hipMalloc(&plan->tiles, std::max<size_t>(1, tiles.size()) * sizeof(SplitTile));
  if (e == hipSuccess) e = hipMalloc(&plan->block_prefix, prefix.size() * sizeof(int64_t));
  if (e == hipSuccess)
    e = hipMalloc(&plan->maxbits, std::max<size_t>(1, tiles.size()) * sizeof(unsigned long long));
  if (e == hipSuccess && !tiles.empty())
    e = hipMemcpy(plan->tiles, tiles.data(), tiles.size() * sizeof(SplitTile), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(plan->block_prefix, prefix.data(), prefix.size() * sizeof(int64_t),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    tamx::set_error(std::string("split plan: ") + hipGetErrorString(e));
    if (plan->tiles) (void)hipFree(plan->tiles);
    if (plan->block_prefix) (void)hipFree(plan->block_prefix);
    if (plan->maxbits) (void)hipFree(plan->maxbits);
    delete plan;
    return e == hipErrorOutOfMemory ? TAMX_ERR_OOM : TAMX_ERR_DEVICE;
  }
  *out = plan;
  return TAMX_OK;
}

int tamx_split_plan_create_general(tamx_context* ctx, int nslices, int64_t ntiles,
                                   const int64_t* host_offsets, int nrow_modes, int nk_modes,
                                   const int32_t* host_row_ext, const int64_t* host_row_str,
                                   const int32_t* host_k_ext, const int64_t* host_k_str,
                                   tamx_split_plan** out) {
  return tamx_split_plan_create_general_typed(ctx, TAMX_LIMB_BF16, nslices, ntiles, host_offsets,
                                              nrow_modes, nk_modes, host_row_ext, host_row_str,
                                              host_k_ext, host_k_str, out);
}

int tamx_split_plan_create_general_typed(tamx_context* ctx, int limb_type, int nslices,
                                         int64_t ntiles, const int64_t* host_offsets,
                                         int nrow_modes, int nk_modes,
                                         const int32_t* host_row_ext,
                                         const int64_t* host_row_str,
                                         const int32_t* host_k_ext, const int64_t* host_k_str,
                                         tamx_split_plan** out) {
  TAMX_REQUIRE(out != nullptr, "null output pointer");
  *out = nullptr;
  TAMX_REQUIRE(limb_type == TAMX_LIMB_BF16 || limb_type == TAMX_LIMB_I8,
               "limb_type must be TAMX_LIMB_BF16 or TAMX_LIMB_I8");
  TAMX_REQUIRE(nslices >= 2 && nslices <= 6, "nslices must be in [2, 6]");
  const int kq = limb_type == TAMX_LIMB_I8 ? SBK8 : SBK;  // k padding quantum
  TAMX_REQUIRE(ntiles >= 0, "negative tile count");
  TAMX_REQUIRE(nrow_modes >= 1 && nrow_modes <= 4 && nk_modes >= 1 && nk_modes <= 4,
               "row and k modes must number 1 to 4 each");
  TAMX_REQUIRE(ntiles == 0 || (host_offsets && host_row_ext && host_row_str && host_k_ext &&
                               host_k_str),
               "null tile table");
  if (int rc = tamx::activate(ctx)) return rc;
  std::vector<SplitTile> tiles(ntiles);
  std::vector<int64_t> prefix(ntiles + 1, 0);
  int64_t off = 0, blocks = 0, max_kpad = 32;
  for (int64_t t = 0; t < ntiles; ++t) {
    SplitTile& st = tiles[t];
    st = SplitTile{};
    st.src_off = host_offsets[t];
    prefix[t] = blocks;
    if (st.src_off < 0) continue;  // absent tile: no limbs
    int64_t nr = 1, nk = 1;
    st.gen_nr = nrow_modes;
    st.gen_nk = nk_modes;
    for (int d = 0; d < nrow_modes; ++d) {
      st.gr_ext[d] = host_row_ext[t * nrow_modes + d];
      st.gr_str[d] = host_row_str[t * nrow_modes + d];
      TAMX_REQUIRE(st.gr_ext[d] >= 0 && st.gr_str[d] >= 0, "negative extent or stride");
      nr *= st.gr_ext[d];
    }
    for (int d = 0; d < nk_modes; ++d) {
      st.gk_ext[d] = host_k_ext[t * nk_modes + d];
      st.gk_str[d] = host_k_str[t * nk_modes + d];
      TAMX_REQUIRE(st.gk_ext[d] >= 0 && st.gk_str[d] >= 0, "negative extent or stride");
      nk *= st.gk_ext[d];
    }
    TAMX_REQUIRE(nr < (int64_t(1) << 31) && nk < (int64_t(1) << 31), "tile too large");
    st.out_rows = (int32_t)nr;
    st.out_k = (int32_t)nk;
    st.rows = st.out_rows;
    st.cols = st.out_k;
    st.rows_pad = (st.out_rows + 15) / 16 * 16;
    st.k_pad = std::max(kq, (st.out_k + kq - 1) / kq * kq);
    st.limb_off = off;
    off += (int64_t)nslices * st.rows_pad * st.k_pad;
    max_kpad = std::max<int64_t>(max_kpad, st.k_pad);
    blocks += (int64_t)((st.rows_pad + kSplitBR - 1) / kSplitBR) * split_blocks_k(st.k_pad);
  }
  prefix[ntiles] = blocks;
  auto* plan = new (std::nothrow) tamx_split_plan();
  if (!plan) return TAMX_ERR_OOM;
  plan->ctx = ctx;
  plan->ntiles = ntiles;
  plan->nslices = nslices;
  plan->limb_type = limb_type;
  plan->limb_elems = off + (int64_t)kLimbSlackRows * max_kpad * nslices;
  plan->nblocks = blocks;
  hipError_t e = hipMalloc(&plan->tiles, std::max<size_t>(1, tiles.size()) * sizeof(SplitTile));
  if (e == hipSuccess) e = hipMalloc(&plan->block_prefix, prefix.size() * sizeof(int64_t));
  if (e == hipSuccess)
    e = hipMalloc(&plan->maxbits, std::max<size_t>(1, tiles.size()) * sizeof(unsigned long long));
  if (e == hipSuccess && !tiles.empty())
    e = hipMemcpy(plan->tiles, tiles.data(), tiles.size() * sizeof(SplitTile), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(plan->block_prefix, prefix.data(), prefix.size() * sizeof(int64_t),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    tamx::set_error(std::string("split plan: ") + hipGetErrorString(e));
    if (plan->tiles) (void)hipFree(plan->tiles);
    if (plan->block_prefix) (void)hipFree(plan->block_prefix);
    if (plan->maxbits) (void)hipFree(plan->maxbits);
    delete plan;
    return e == hipErrorOutOfMemory ? TAMX_ERR_OOM : TAMX_ERR_DEVICE;
  }
  *out = plan;
  return TAMX_OK;
}

int tamx_split_plan_destroy(tamx_split_plan* plan) {
  if (!plan) return TAMX_OK;
  if (plan->tiles) (void)hipFree(plan->tiles);
  if (plan->block_prefix) (void)hipFree(plan->block_prefix);
  if (plan->maxbits) (void)hipFree(plan->maxbits);
  delete plan;
  return TAMX_OK;
}

int tamx_split_plan_limb_type(const tamx_split_plan* plan) {
  return plan ? plan->limb_type : TAMX_LIMB_BF16;
}

int tamx_split_plan_nslices(const tamx_split_plan* plan) {
  return plan ? plan->nslices : 0;
}

int64_t tamx_split_plan_limb_elems(const tamx_split_plan* plan) {
  return plan ? plan->limb_elems : 0;
}

int tamx_split_plan_sigmas(const tamx_split_plan* plan, int32_t* host_sigma) {
  TAMX_REQUIRE(plan != nullptr && host_sigma != nullptr, "null argument");
  std::vector<SplitTile> tiles(plan->ntiles);
  if (!tiles.empty())
    TAMX_CHECK_HIP(hipMemcpy(tiles.data(), plan->tiles, tiles.size() * sizeof(SplitTile),
                             hipMemcpyDeviceToHost));
  for (size_t t = 0; t < tiles.size(); ++t) host_sigma[t] = tiles[t].sigma;
  return TAMX_OK;
}

int tamx_split_f64(tamx_context* ctx, const tamx_split_plan* plan, const double* pool,
                   void* limbs, void* stream) {
  return split_f64(ctx, plan, pool, nullptr, limbs, stream);
}

int tamx_split_f64_premax(tamx_context* ctx, const tamx_split_plan* plan, const double* pool,
                          const double* tile_maxabs, void* limbs, void* stream) {
  TAMX_REQUIRE(tile_maxabs != nullptr, "null tile_maxabs");
  return split_f64(ctx, plan, pool, tile_maxabs, limbs, stream);
}

int tamx_grouped_gemm_split(tamx_context* ctx, tamx_worklist* wl, double alpha,
                            const tamx_split_plan* a_plan, const void* a_limbs,
                            const tamx_split_plan* b_plan, const void* b_limbs,
                            double* C, void* stream) {
  return tamx_grouped_gemm_split_config(ctx, wl, alpha, a_plan, a_limbs, b_plan, b_limbs, C,
                                        nullptr, stream);
}

int tamx_grouped_gemm_split_config(tamx_context* ctx, tamx_worklist* wl, double alpha,
                                   const tamx_split_plan* a_plan, const void* a_limbs,
                                   const tamx_split_plan* b_plan, const void* b_limbs,
                                   double* C, const tamx_gemm_config* cfg, void* stream) {
  return tamx_grouped_gemm_split_scatter(ctx, wl, alpha, a_plan, a_limbs, b_plan, b_limbs, C,
                                         nullptr, cfg, stream);
}

int tamx_grouped_gemm_split_scatter(tamx_context* ctx, tamx_worklist* wl, double alpha,
                                    const tamx_split_plan* a_plan, const void* a_limbs,
                                    const tamx_split_plan* b_plan, const void* b_limbs,
                                    double* C, const tamx_gather_plan* c_scatter,
                                    const tamx_gemm_config* cfg, void* stream) {
  TAMX_REQUIRE(wl && a_plan && b_plan && a_limbs && b_limbs, "null argument");
  TAMX_REQUIRE(c_scatter == nullptr || c_scatter->ctx == ctx, "scatter plan of another context");
  TAMX_REQUIRE(a_plan->nslices == b_plan->nslices, "operands split into different limb counts");
  TAMX_REQUIRE(a_plan->limb_type == b_plan->limb_type, "operands split into different limb types");
  if (int rc = tamx::activate(ctx)) return rc;
  SplitParams prm{static_cast<const bf8_t*>(a_limbs),
                  static_cast<const bf8_t*>(b_limbs),
                  a_plan->tiles,
                  b_plan->tiles,
                  C,
                  wl->tasks,
                  wl->pairs,
                  wl->counters,
                  wl->item_prefix,
                  wl->item_task,
                  wl->queue,
                  wl->task_capacity,
                  wl->item_capacity,
                  alpha,
                  c_scatter ? c_scatter->tiles : nullptr,
                  wl->task_tile};
  hipStream_t s = tamx::as_stream(stream);
  if (a_plan->limb_type == TAMX_LIMB_I8) {
    if (cfg && cfg->bm == SBM && cfg->bn == SBN && cfg->bk == 2 * SBK8)
      return launch_split_gemm_i8_2w(a_plan->nslices, ctx, wl, prm, s);
    if (cfg && cfg->bm == SBM && cfg->bn == SBN && cfg->bk == SBK8 / 2)
      return launch_split_gemm_i8_ring(a_plan->nslices, ctx, wl, prm, s);
    // 128 x 64 from five limbs on: eight waves of 32 x 32 (the wide
    // kernel's 64 x 32 waves need too many accumulator registers there)
    if (cfg && cfg->bm == 2 * SBM && cfg->bn == SBN && a_plan->nslices >= 5)
      return a_plan->nslices == 5 ? launch_split_gemm<5, true>(ctx, wl, prm, s, false, true)
                                  : launch_split_gemm<6, true>(ctx, wl, prm, s, false, true);
    if (cfg && !(cfg->bm == SBM && cfg->bn == SBN))
      return launch_split_gemm_wide(TAMX_LIMB_I8, a_plan->nslices, cfg->bm, cfg->bn, ctx, wl,
                                    prm, s);
    // bk 64: the 32x32x32 blocks
    const bool m32 = cfg && cfg->bk == SBK8;
    switch (a_plan->nslices) {
      case 2: return launch_split_gemm<2, true>(ctx, wl, prm, s, m32);
      case 3: return launch_split_gemm<3, true>(ctx, wl, prm, s, m32);
      case 4: return launch_split_gemm<4, true>(ctx, wl, prm, s, m32);
      case 5: return launch_split_gemm<5, true>(ctx, wl, prm, s, m32);
      default: return launch_split_gemm<6, true>(ctx, wl, prm, s, m32);
    }
  }
  if (cfg && !(cfg->bm == SBM && cfg->bn == SBN))
    return launch_split_gemm_wide(TAMX_LIMB_BF16, a_plan->nslices, cfg->bm, cfg->bn, ctx, wl,
                                  prm, s);
  switch (a_plan->nslices) {
    case 2: return launch_split_gemm<2, false>(ctx, wl, prm, s);
    case 3: return launch_split_gemm<3, false>(ctx, wl, prm, s);
    case 4: return launch_split_gemm<4, false>(ctx, wl, prm, s);
    case 5: return launch_split_gemm<5, false>(ctx, wl, prm, s);
    default: return launch_split_gemm<6, false>(ctx, wl, prm, s);
  }
}

int tamx_grouped_gemm_f32x2(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
                            double alpha, const double* A, const double* B, double* C,
                            void* stream) {
  return tamx_grouped_gemm_f32x2_scatter(ctx, wl, op_a, op_b, alpha, A, B, C, nullptr, stream);
}

int tamx_grouped_gemm_f32x2_scatter(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
                                    double alpha, const double* A, const double* B, double* C,
                                    const tamx_gather_plan* c_scatter, void* stream) {
  TAMX_REQUIRE(wl != nullptr, "null work list");
  TAMX_REQUIRE(c_scatter == nullptr || c_scatter->ctx == ctx, "scatter plan of another context");
  TAMX_REQUIRE(op_a == 0 || op_a == 1, "op_a must be TAMX_NO_TRANS or TAMX_TRANS");
  TAMX_REQUIRE(op_b == 0 || op_b == 1, "op_b must be TAMX_NO_TRANS or TAMX_TRANS");
  if (int rc = tamx::activate(ctx)) return rc;
  SplitParams prm{nullptr, nullptr, nullptr, nullptr, C, wl->tasks, wl->pairs, wl->counters,
                  wl->item_prefix, wl->item_task, wl->queue, wl->task_capacity,
                  wl->item_capacity, alpha, c_scatter ? c_scatter->tiles : nullptr,
                  wl->task_tile};
  hipStream_t s = tamx::as_stream(stream);
  if (!op_a && !op_b) return launch_f32x2<false, false>(ctx, wl, A, B, prm, s);
  if (!op_a && op_b) return launch_f32x2<false, true>(ctx, wl, A, B, prm, s);
  if (op_a && !op_b) return launch_f32x2<true, false>(ctx, wl, A, B, prm, s);
  return launch_f32x2<true, true>(ctx, wl, A, B, prm, s);
}

}  // extern "C"
