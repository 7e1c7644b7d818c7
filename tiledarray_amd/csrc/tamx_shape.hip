// Sparse-shape screening on the device: tile norms, inverse-volume scaling,
// the float norm-GEMM result estimate, and the work-list builder that turns
// the screened shapes into the tasks/pairs of a grouped GEMM.
//
// Float arithmetic that decides a sparsity pattern is written with plain
// operators and compiled with FP contraction off (this file is built with
// -ffp-contract=off, and each such kernel also says so with
// #pragma clang fp contract(off)), so every product and sum rounds exactly
// where the reference's expression does and the pattern is reproducible bit
// for bit.  (HIP's __fmul_rn/__fadd_rn are header functions whose operations
// keep the contractable flag, so they are not used here.)
#include <algorithm>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "tamx_internal.hpp"
#include "tamx_scan.hpp"

struct tamx_norm_plan {
  tamx_context* ctx = nullptr;
  int64_t ntiles = 0;
  int64_t nchunks = 0;
  // chunk c covers elements [cbeg[c], cend[c]) of pool offsets of tile ctile[c]
  int64_t* cbeg = nullptr;
  int64_t* cend = nullptr;
  int32_t* ctile = nullptr;
  // tile t owns chunks [tchunk[t], tchunk[t+1]); no chunk for absent tiles
  int64_t* tchunk = nullptr;
  double* partial = nullptr;  // one partial sum per chunk
  double* partial_max = nullptr;  // one partial max |a| per chunk
};

namespace {

constexpr int64_t kNormChunk = 1 << 16;  // elements per workgroup
constexpr int kNormThreads = 256;

// Where chunk c lies: a table built on the host (tamx_norm_plan) ...
struct TableChunks {
  const int64_t* __restrict__ cbeg;
  const int64_t* __restrict__ cend;
  __device__ bool operator()(int64_t c, int64_t& b, int64_t& e) const {
    b = cbeg[c];
    e = cend[c];
    return true;
  }
};
// ... or the per-tile chunk prefix tchunk (tchunk[ntiles] = the chunk count,
// negative for an invalid table) and the tile table itself: chunk c is
// chunk j = c - tchunk[t] of the tile t with tchunk[t] <= c < tchunk[t+1],
// elements [j, j + 1) * kNormChunk of the tile clipped to its volume, the
// same ranges the host table lists.
struct PrefixChunks {
  const int64_t* __restrict__ offsets;
  const int64_t* __restrict__ volumes;
  const int64_t* __restrict__ tchunk;
  int64_t ntiles;
  __device__ bool operator()(int64_t c, int64_t& b, int64_t& e) const {
    if (!(c < tchunk[ntiles])) return false;
    int64_t lo = 0, hi = ntiles - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (tchunk[mid] <= c) lo = mid; else hi = mid - 1;
    }
    const int64_t j = c - tchunk[lo], off = offsets[lo], vol = volumes[lo];
    b = off + j * kNormChunk;
    e = off + min(vol, (j + 1) * kNormChunk);
    return true;
  }
};

// Partial sums of squares, one workgroup per chunk.  Accumulation in the
// tile's scalar type: double for FP64 tiles, float for FP32 tiles, as
// Tensor<T>::squared_norm does (tensor/tensor.h:3768-3776).  With MAX, the
// same pass also takes each chunk's max |a| (for the limb split's scale).
template <typename T, bool MAX, class Chunks>
__global__ __launch_bounds__(kNormThreads) void norm_partial_kernel(
    const T* __restrict__ pool, Chunks chunks, double* __restrict__ partial,
    double* __restrict__ partial_max) {
  __shared__ T red[kNormThreads];
  __shared__ T redm[MAX ? kNormThreads : 1];
  const int64_t c = blockIdx.x;
  int64_t b, e;
  if (!chunks(c, b, e)) return;  // uniform over the workgroup
  // eight independent loads in flight per thread, eight partial sums
  constexpr int U = 8;
  T s[U];
  T m = T(0);
#pragma unroll
  for (int u = 0; u < U; ++u) s[u] = T(0);
  int64_t i = b + threadIdx.x;
  for (; i + (U - 1) * kNormThreads < e; i += U * kNormThreads) {
    T x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = pool[i + u * kNormThreads];
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] += x[u] * x[u];
    if (MAX) {
#pragma unroll
      for (int u = 0; u < U; ++u) m = fmax(m, fabs(x[u]));
    }
  }
  for (; i < e; i += kNormThreads) {
    const T x = pool[i];
    s[0] += x * x;
    if (MAX) m = fmax(m, fabs(x));
  }
  T t = T(0);
#pragma unroll
  for (int u = 0; u < U; ++u) t += s[u];
  red[threadIdx.x] = t;
  if (MAX) redm[threadIdx.x] = m;
  __syncthreads();
  for (int w = kNormThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x] += red[threadIdx.x + w];
      if (MAX) redm[threadIdx.x] = fmax(redm[threadIdx.x], redm[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[c] = (double)red[0];
    if (MAX) partial_max[c] = (double)redm[0];
  }
}

// A tile's norm from its chunks [b, e) in chunk order: sqrt(sum), cast to
// the float shape type (make_array stores into Tensor<float>,
// conversions/make_array.h:170-172); 0 for a tile with no chunk.
template <typename T>
__device__ inline float finish_norm(const double* __restrict__ partial, int64_t b, int64_t e) {
  if (b == e) return 0.0f;
  // the chunks are added one after another, in order; sixteen are loaded
  // before they are added, so a tile of many chunks waits for a load
  // sixteen times less often
  constexpr int U = 16;
  using acc_t = std::conditional_t<sizeof(T) == sizeof(double), double, float>;
  acc_t s = 0;
  int64_t c = b;
  for (; c + U <= e; c += U) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = partial[c + u];
#pragma unroll
    for (int u = 0; u < U; ++u) s += (acc_t)v[u];
  }
  for (; c < e; ++c) s += (acc_t)partial[c];
  if constexpr (sizeof(T) == sizeof(double)) return (float)sqrt(s);
  else return sqrtf(s);
}

// Combines the chunks of each tile (finish_norm); with maxabs, also the
// tile's max |a|.
template <typename T>
__global__ void norm_finish_kernel(const int64_t* __restrict__ tchunk,
                                   const double* __restrict__ partial,
                                   const double* __restrict__ partial_max,
                                   int64_t ntiles, float* __restrict__ norms,
                                   double* __restrict__ maxabs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const int64_t b = tchunk[t], e = tchunk[t + 1];
  if (maxabs) {
    double m = 0.0;
    int64_t c = b;
    for (; c + 16 <= e; c += 16) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = partial_max[c + u];
#pragma unroll
      for (int u = 0; u < 16; ++u) m = fmax(m, v[u]);
    }
    for (; c < e; ++c) m = fmax(m, partial_max[c]);
    maxabs[t] = m;
  }
  norms[t] = finish_norm<T>(partial, b, e);
}

template <typename T>
int tile_norms(tamx_context* ctx, const tamx_norm_plan* plan, const T* pool,
               float* norms, double* maxabs, void* stream) {
  TAMX_REQUIRE(plan != nullptr && norms != nullptr, "null argument");
  if (plan->ntiles == 0) return TAMX_OK;
  TAMX_REQUIRE(pool != nullptr || plan->nchunks == 0, "null pool");
  if (int rc = tamx::activate(ctx)) return rc;
  hipStream_t s = tamx::as_stream(stream);
  if (plan->nchunks > 0) {
    if (maxabs)
      norm_partial_kernel<T, true><<<(unsigned)plan->nchunks, kNormThreads, 0, s>>>(
          pool, TableChunks{plan->cbeg, plan->cend}, plan->partial, plan->partial_max);
    else
      norm_partial_kernel<T, false><<<(unsigned)plan->nchunks, kNormThreads, 0, s>>>(
          pool, TableChunks{plan->cbeg, plan->cend}, plan->partial, nullptr);
    TAMX_LAUNCH_CHECK();
  }
  const unsigned grid = (unsigned)((plan->ntiles + 255) / 256);
  norm_finish_kernel<T><<<grid, 256, 0, s>>>(plan->tchunk, plan->partial, plan->partial_max,
                                              plan->ntiles, norms, maxabs);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

// ------------------------------------------ norms without allocation --
// tamx_tile_norms_f64_noalloc finds the chunks on the device (the same
// ranges tamx_norm_plan_create lists on the host) and keeps the chunk prefix
// and the partial sums in the file-scope `scratch` array below.

// tchunk[t] = exclusive prefix over tiles of ceil(volume / kNormChunk)
// (absent or empty tiles have none); tchunk[ntiles] = the total, or -1 when
// the total exceeds max_chunks or a present tile has a negative volume.
__global__ __launch_bounds__(1024) void norm_chunk_scan_kernel(
    const int64_t* __restrict__ offsets, const int64_t* __restrict__ volumes,
    int64_t ntiles, int64_t max_chunks, int64_t* __restrict__ tchunk) {
  __shared__ int64_t part[1024];
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const int64_t total = tamx::block_exclusive_scan<1024, int64_t>(
      ntiles, part,
      [&](int64_t t) -> int64_t {
        const int64_t o = offsets[t], v = volumes[t];
        if (o < 0) return 0;
        if (v < 0) {
          bad = 1;
          return 0;
        }
        return (v + kNormChunk - 1) / kNormChunk;
      },
      [&](int64_t t, const int64_t& x) { tchunk[t] = x; });
  if (threadIdx.x == 0) tchunk[ntiles] = (bad || total > max_chunks) ? -1 : total;
}

// finish_norm per tile, NaN for every tile of an invalid table.
template <typename T>
__global__ void norm_finish_prefix_kernel(const int64_t* __restrict__ tchunk,
                                          const double* __restrict__ partial, int64_t ntiles,
                                          float* __restrict__ norms) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  norms[t] = tchunk[ntiles] < 0 ? __builtin_nanf("") : finish_norm<T>(partial, tchunk[t], tchunk[t + 1]);
}

constexpr int64_t kNoallocPrefixBytes = ((TAMX_NOALLOC_MAX_TILES + 1) * 8 + 255) / 256 * 256;
constexpr int64_t kNoallocScratchBytes = kNoallocPrefixBytes + TAMX_NOALLOC_MAX_CHUNKS * 8;

// ------------------------------------------------------- shape scaling --

// Adds the number of the wave's active lanes with `zero` set to *count: one
// atomic per wave (the lanes that returned early take no part).
__device__ inline void count_zeros(bool zero, int64_t* count) {
  const unsigned long long zmask = __ballot(zero);
  const unsigned long long amask = __ballot(1);
  if (zmask && __lane_id() == (unsigned)(__ffsll((long long)amask) - 1))
    atomicAdd((unsigned long long*)count, (unsigned long long)__popcll(zmask));
}

constexpr int kMaxRank = 8;

struct ScaleParams {
  float* norms;
  const float* sizes;
  int64_t* zero_count;
  int64_t ntiles[kMaxRank];
  int64_t mode_off[kMaxRank];  // start of each mode in `sizes`
  int64_t per_batch;           // tiles of one grid
  int64_t sizes_batch_stride;  // 0 when sizes are shared by all batches
  int64_t nbatch;
  int rank;
  float threshold;
};

// recursive_outer_product with the inverse-volume op (sparse_shape.h:106-131,
// :177-181) evaluated for one multi-index: modes [lo, hi).
__device__ float inv_rop(const ScaleParams& p, const float* sizes,
                         const int64_t* idx, int lo, int hi) {
#pragma clang fp contract(off)
  if (hi - lo == 1) return 1.0f / sizes[p.mode_off[lo] + idx[lo]];
  const int n = hi - lo;
  const int middle = (n >> 1) + (n & 1);
  const float l = inv_rop(p, sizes, idx, lo, lo + middle);
  const float r = inv_rop(p, sizes, idx, lo + middle, hi);
  return l * r;
}

__global__ void shape_scale_kernel(ScaleParams p) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= p.per_batch * p.nbatch) return;
  const int64_t b = g / p.per_batch;
  int64_t ord = g % p.per_batch;
  const float* sizes = p.sizes + b * p.sizes_batch_stride;
  int64_t idx[kMaxRank];
  for (int d = p.rank - 1; d >= 0; --d) {
    idx[d] = ord % p.ntiles[d];
    ord /= p.ntiles[d];
  }
  float v = p.norms[g];
  if (p.rank == 1) {
    v = v / sizes[idx[0]];  // rank-1 divides (sparse_shape.h:160-173)
  } else {
    const int middle = (p.rank >> 1) + (p.rank & 1);
    const float x = inv_rop(p, sizes, idx, 0, middle);
    const float y = inv_rop(p, sizes, idx, middle, p.rank);
    v = v * (x * y);  // norm *= x * y (:205-213)
  }
  const bool zero = v < p.threshold;
  if (zero) v = 0.0f;
  if (p.zero_count) count_zeros(zero, p.zero_count);
  p.norms[g] = v;
}

// ---------------------------------------------------------- shape GEMM --

struct ShapeGemmParams {
  const float* left;
  const float* right;
  const float* ksz;
  float* result;
  int64_t* zero_count;
  int64_t M, N, K, nbatch;
  int op_a, op_b;
  float abs_factor, threshold;
};

// est_ij = abs_factor * sum_k l_ik * r_kj with the k_sizes scaling applied
// along contiguous storage runs as SparseShape::gemm does
// (sparse_shape.h:1632-1647): left[s] *= k_sizes[s % K], right[s] *=
// k_sizes[s / N].  One thread per result tile, k ascending.
__global__ void shape_gemm_kernel(ShapeGemmParams p) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= p.M * p.N * p.nbatch) return;
  const int64_t b = g / (p.M * p.N);
  const int64_t ij = g % (p.M * p.N);
  const int64_t i = ij / p.N, j = ij % p.N;
  const int64_t kk = p.K > 0 ? p.K : 1;  // an outer product's operands hold M and N values
  const float* L = p.left + b * p.M * kk;
  const float* R = p.right + b * kk * p.N;
  const float* ks = p.ksz + b * p.K;
  float v;
  if (p.K == 0) {
    v = (L[i] * R[j]) * p.abs_factor;  // outer product (:1666-1678)
  } else {
    // storage indices sl of (i, k) in left and sr of (k, j) in right, with
    // sl % K and (sr / N, sr % N) carried along as k steps sl by dl and sr
    // by dr (no division in the loop)
    const int64_t K = p.K, N = p.N;
    const int64_t dl = p.op_a ? p.M : 1, dr = p.op_b ? 1 : N;
    int64_t sl = p.op_a ? i : i * K, sr = p.op_b ? j * K : j;
    int64_t lk = sl % K, rq = sr / N, rr = sr % N;
    const int64_t dlk = dl % K, drq = dr / N, drr = dr % N;
    float acc = 0.0f;
    // unrolled so that several steps' loads are in flight; the sum still
    // runs k ascending
#pragma unroll 8
    for (int64_t k = 0; k < K; ++k) {
      const float l = L[sl] * ks[lk];
      const float r = R[sr] * ks[rq];
      const float prod = l * r;
      acc = acc + prod;
      sl += dl;
      lk += dlk;
      if (lk >= K) lk -= K;
      sr += dr;
      rq += drq;
      rr += drr;
      if (rr >= N) {
        rr -= N;
        ++rq;
      }
    }
    v = p.abs_factor * acc;
  }
  const bool zero = v < p.threshold;
  if (zero) v = 0.0f;
  if (p.zero_count) count_zeros(zero, p.zero_count);
  p.result[g] = v;
}

// ---------------------------------------------------- work-list builder --

// Sums scanned to place every non-zero result tile: (task flag, pair count,
// element count, flops).
struct Quad {
  int64_t t, p, c, f;
  __device__ Quad(int v = 0) : t(v), p(v), c(v), f(v) {}
  __device__ Quad(int64_t t_, int64_t p_, int64_t c_, int64_t f_) : t(t_), p(p_), c(c_), f(f_) {}
  __device__ Quad& operator+=(const Quad& o) {
    t += o.t;
    p += o.p;
    c += o.c;
    f += o.f;
    return *this;
  }
};

constexpr int kWlThreads = 256;  // threads per workgroup of count / fill
// lanes that share one result tile's k walk (each takes a contiguous k
// range), and so result tiles per workgroup
constexpr int kWlSub = 8;
constexpr int kWlTiles = kWlThreads / kWlSub;

struct WlParams {
  tamx_contraction_desc d;
  tamx_gemm_task* tasks;
  int32_t* task_tile;
  tamx_gemm_pair* pairs;
  int64_t* counters;
  int32_t* tile_npairs;
  int64_t* tile_volume;
  int64_t* tile_flops;
  Quad* partial;  // one sum per count workgroup, then the exclusive offsets
  int64_t task_capacity, pair_capacity;
};

// The work-list walks load the operands' offsets and norms for kWlK
// consecutive k at once (every load in flight before any test; a k past the
// end reads the last one again and is not tested), then test the pairs in k
// order: (A_ik, B_kj) contributes when both are stored and not screened out
// and (optional per-pair rule) |A_ik| |B_kj| >= thr_k.
constexpr int kWlK = 8;
struct KSides {
  int64_t a_off[kWlK], b_off[kWlK];
  float a_norm[kWlK], b_norm[kWlK];
};
template <class AiOf, class BiOf>
__device__ inline void load_k_sides(const tamx_contraction_desc& d, int64_t k0, AiOf ai_of,
                                    BiOf bi_of, KSides& s) {
#pragma unroll
  for (int u = 0; u < kWlK; ++u) {
    const int64_t k = min<int64_t>(k0 + u, d.Kt - 1);
    const int64_t ai = ai_of(k), bi = bi_of(k);
    s.a_off[u] = d.a_off[ai];
    s.b_off[u] = d.b_off[bi];
    s.a_norm[u] = d.a_norms ? d.a_norms[ai] : 0.0f;
    s.b_norm[u] = d.b_norms ? d.b_norms[bi] : 0.0f;
  }
}
__device__ inline bool pair_kept_k(const tamx_contraction_desc& d, const KSides& s, int u,
                                   float thr_k) {
#pragma clang fp contract(off)
  if (!(s.a_off[u] >= 0 && (!d.a_norms || !(s.a_norm[u] < d.a_threshold)))) return false;
  if (!(s.b_off[u] >= 0 && (!d.b_norms || !(s.b_norm[u] < d.b_threshold)))) return false;
  if (thr_k > 0.0f && d.a_norms && d.b_norms && s.a_norm[u] * s.b_norm[u] < thr_k) return false;
  return true;
}

__device__ inline Quad tile_quad(const WlParams& p, int64_t g) {
  const int64_t v = p.tile_volume[g];
  return Quad(v > 0 ? 1 : 0, p.tile_npairs[g], v, p.tile_flops[g]);
}

// k range [kb, ke) of sub-walker `sub` of a tile's walk
__device__ inline void sub_range(int64_t Kt, int sub, int64_t& kb, int64_t& ke) {
  const int64_t chunk = (Kt + kWlSub - 1) / kWlSub;
  kb = min<int64_t>(Kt, sub * chunk);
  ke = min<int64_t>(Kt, kb + chunk);
}

// Sum over the kWlSub consecutive lanes that walk one tile (the lanes of a
// group are aligned inside the wave).
template <typename V>
__device__ inline V sub_sum(V v) {
#pragma unroll
  for (int o = 1; o < kWlSub; o <<= 1) v += __shfl_xor(v, o, kWlSub);
  return v;
}

// Pass 1: per result tile, the number of contributing pairs, the element
// count and the flops (zero for result tiles screened out by the shape);
// per workgroup, the sum of its tiles' Quads.  kWlSub lanes share a tile,
// each walking a contiguous range of k.
__global__ __launch_bounds__(kWlThreads) void wl_count_kernel(WlParams p) {
#pragma clang fp contract(off)
  const tamx_contraction_desc& d = p.d;
  const int sub = threadIdx.x % kWlSub;
  const int64_t g = (int64_t)blockIdx.x * kWlTiles + threadIdx.x / kWlSub;
  const int64_t per = d.Mt * d.Nt;
  int32_t np = 0;
  int64_t vol = 0, flops = 0;
  if (g < per * d.nbatch) {
    const int64_t b = g / per, ij = g % per, i = ij / d.Nt, j = ij % d.Nt;
    if (!(d.c_norms[g] < d.c_threshold)) {
      const int64_t m = d.mext[b * d.Mt + i], n = d.next[b * d.Nt + j];
      vol = m * n;
      const float thr_k = d.pair_threshold > 0.0f
                              ? d.pair_threshold / (float)d.Kt
                              : 0.0f;
      auto ai_of = [&](int64_t k) {
        return b * d.Mt * d.Kt + (d.op_a ? k * d.Mt + i : i * d.Kt + k);
      };
      auto bi_of = [&](int64_t k) {
        return b * d.Kt * d.Nt + (d.op_b ? j * d.Kt + k : k * d.Nt + j);
      };
      int64_t kb, ke;
      sub_range(d.Kt, sub, kb, ke);
      for (int64_t k0 = kb; k0 < ke; k0 += kWlK) {
        KSides sd;
        load_k_sides(d, k0, ai_of, bi_of, sd);
#pragma unroll
        for (int u = 0; u < kWlK; ++u)
          if (k0 + u < ke && pair_kept_k(d, sd, u, thr_k)) {
            ++np;
            flops += 2 * m * n * (int64_t)d.kext[b * d.Kt + k0 + u];
          }
      }
    }
  }
  np = sub_sum(np);
  flops = sub_sum(flops);
  if (sub != 0) vol = 0;  // the tile counts once in the workgroup's sum
  if (sub == 0 && g < per * d.nbatch) {
    p.tile_npairs[g] = np;
    p.tile_volume[g] = vol;
    p.tile_flops[g] = flops;
  }
  // raw storage: __shared__ objects cannot have constructors
  __shared__ __align__(16) unsigned char red_raw[sizeof(Quad) * kWlThreads];
  Quad* red = reinterpret_cast<Quad*>(red_raw);
  red[threadIdx.x] = sub == 0 ? Quad(vol > 0 ? 1 : 0, np, vol, flops) : Quad(0);
  __syncthreads();
  for (int w = kWlThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(1024) void wl_scan_kernel(WlParams p, int nparts) {
  __shared__ __align__(16) unsigned char part_raw[sizeof(Quad) * 1024];
  Quad* part = reinterpret_cast<Quad*>(part_raw);
  const tamx_contraction_desc& d = p.d;
  const Quad tot = tamx::block_exclusive_scan<1024, Quad>(
      nparts, part, [&](int64_t i) -> Quad { return p.partial[i]; },
      [&](int64_t i, const Quad& v) { p.partial[i] = v; });
  if (threadIdx.x == 0) {
    int64_t status = 0;
    if (tot.t > p.task_capacity || tot.p > p.pair_capacity || tot.c > d.c_capacity)
      status = TAMX_ERR_CAPACITY;
    p.counters[TAMX_WL_NTASKS] = status ? 0 : tot.t;
    p.counters[TAMX_WL_NPAIRS] = status ? 0 : tot.p;
    p.counters[TAMX_WL_CELEMS] = tot.c;
    p.counters[TAMX_WL_STATUS] = status;
    p.counters[TAMX_WL_FLOPS] = tot.f;
    // the totals that were asked for, for the host's capacity message
    p.counters[5] = tot.t;
    p.counters[6] = tot.p;
    p.counters[7] = 0;
  }
}

// Pass 3: every workgroup rescans its tiles from its offset and writes the
// tasks (row-major tile order), the C offsets and the pairs (k ascending):
// the kWlSub lanes of a tile count their k ranges, scan the counts, and
// write their pairs in k order.
__global__ __launch_bounds__(kWlThreads) void wl_fill_kernel(WlParams p) {
#pragma clang fp contract(off)
  const tamx_contraction_desc& d = p.d;
  const int64_t per = d.Mt * d.Nt;
  const int64_t first = (int64_t)blockIdx.x * kWlTiles;
  const int sub = threadIdx.x % kWlSub;
  const int lt = threadIdx.x / kWlSub;  // tile of the workgroup
  const int64_t g = first + lt;
  const int64_t nloc = min<int64_t>(kWlTiles, per * d.nbatch - first);
  __shared__ __align__(16) unsigned char part_raw[sizeof(Quad) * kWlThreads];
  __shared__ __align__(16) unsigned char mine_raw[sizeof(Quad) * kWlTiles];
  Quad* part = reinterpret_cast<Quad*>(part_raw);
  Quad* mine_l = reinterpret_cast<Quad*>(mine_raw);
  // thread t < nloc receives the prefix of tile first + t
  tamx::block_exclusive_scan<kWlThreads, Quad>(
      nloc, part, [&](int64_t i) -> Quad { return tile_quad(p, first + i); },
      [&](int64_t i, const Quad& v) {
        Quad x = p.partial[blockIdx.x];
        x += v;
        mine_l[i] = x;
      });
  __syncthreads();
  if (g >= per * d.nbatch) return;
  const bool ok = p.counters[TAMX_WL_STATUS] == 0;
  if (!ok || p.tile_volume[g] == 0) {
    if (sub == 0) d.c_off[g] = -1;
    return;
  }
  const Quad mine = mine_l[lt];
  const int64_t b = g / per, ij = g % per, i = ij / d.Nt, j = ij % d.Nt;
  const int32_t m = d.mext[b * d.Mt + i], n = d.next[b * d.Nt + j];
  const int32_t t = (int32_t)mine.t;
  const int32_t pb = (int32_t)mine.p;
  if (sub == 0) {
    const int64_t coff = mine.c;
    d.c_off[g] = coff;
    tamx_gemm_task task;
    task.c_off = coff;
    task.m = m;
    task.n = n;
    task.ldc = d.store_trans ? m : n;
    task.pair_begin = pb;
    task.pair_count = p.tile_npairs[g];
    task.flags = d.store_trans ? TAMX_TASK_STORE_TRANS : 0;
    p.tasks[t] = task;
    p.task_tile[t] = (int32_t)g;
  }
  const float thr_k =
      d.pair_threshold > 0.0f ? d.pair_threshold / (float)d.Kt : 0.0f;
  auto ai_of = [&](int64_t k) {
    return b * d.Mt * d.Kt + (d.op_a ? k * d.Mt + i : i * d.Kt + k);
  };
  auto bi_of = [&](int64_t k) {
    return b * d.Kt * d.Nt + (d.op_b ? j * d.Kt + k : k * d.Nt + j);
  };
  int64_t kb, ke;
  sub_range(d.Kt, sub, kb, ke);
  // this lane's pairs, then its start among the tile's pairs
  int32_t cnt = 0;
  for (int64_t k0 = kb; k0 < ke; k0 += kWlK) {
    KSides sd;
    load_k_sides(d, k0, ai_of, bi_of, sd);
#pragma unroll
    for (int u = 0; u < kWlK; ++u)
      if (k0 + u < ke && pair_kept_k(d, sd, u, thr_k)) ++cnt;
  }
  int32_t incl = cnt;
#pragma unroll
  for (int o = 1; o < kWlSub; o <<= 1) {
    const int32_t y = __shfl_up(incl, o, kWlSub);
    if (sub >= o) incl += y;
  }
  int32_t q = pb + incl - cnt;
  for (int64_t k0 = kb; k0 < ke; k0 += kWlK) {
    KSides sd;
    load_k_sides(d, k0, ai_of, bi_of, sd);
#pragma unroll
    for (int u = 0; u < kWlK; ++u) {
      const int64_t k = k0 + u;
      if (!(k < ke && pair_kept_k(d, sd, u, thr_k))) continue;
      const int32_t kk = d.kext[b * d.Kt + k];
      tamx_gemm_pair pr;
      pr.a_off = sd.a_off[u];
      pr.b_off = sd.b_off[u];
      pr.k = kk;
      pr.lda = d.op_a ? m : kk;
      pr.ldb = d.op_b ? kk : n;
      pr.a_tile = (int32_t)ai_of(k);
      pr.b_tile = (int32_t)bi_of(k);
      pr.reserved = 0;
      p.pairs[q++] = pr;
    }
  }
}

}  // namespace

// Scratch of tamx_tile_norms_f64_noalloc: the chunk prefix, then one partial
// sum per chunk (about 10 MiB).  Every GPU the library runs on holds its own
// copy until the process ends.  In host code this name is a stand-in, not a
// device address: the device address comes from hipGetSymbolAddress
// (norm_scratch_address), once per device, and only that address is used.
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace tamx {
// The device address of `scratch` on ctx's device (current when called):
// resolved once per device, at context creation.
int norm_scratch_address(tamx_context* ctx) {
  static std::mutex mu;
  static std::unordered_map<int, void*> addr;
  std::lock_guard<std::mutex> g(mu);
  auto it = addr.find(ctx->device);
  if (it == addr.end()) {
    void* ptr = nullptr;
    TAMX_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    it = addr.emplace(ctx->device, ptr).first;
  }
  ctx->norm_scratch = it->second;
  return TAMX_OK;
}
}  // namespace tamx

extern "C" {

int tamx_norm_plan_create(tamx_context* ctx, const int64_t* host_offsets,
                          const int64_t* host_volumes, int64_t ntiles,
                          tamx_norm_plan** out) {
  TAMX_REQUIRE(out != nullptr, "null output pointer");
  *out = nullptr;
  TAMX_REQUIRE(ntiles >= 0, "negative tile count");
  TAMX_REQUIRE(ntiles == 0 || (host_offsets && host_volumes), "null tile table");
  if (int rc = tamx::activate(ctx)) return rc;
  std::vector<int64_t> cbeg, cend, tchunk(ntiles + 1);
  std::vector<int32_t> ctile;
  for (int64_t t = 0; t < ntiles; ++t) {
    tchunk[t] = (int64_t)cbeg.size();
    if (host_offsets[t] < 0) continue;
    TAMX_REQUIRE(host_volumes[t] >= 0, "negative tile volume");
    for (int64_t b = 0; b < host_volumes[t]; b += kNormChunk) {
      cbeg.push_back(host_offsets[t] + b);
      cend.push_back(host_offsets[t] + std::min(host_volumes[t], b + kNormChunk));
      ctile.push_back((int32_t)t);
    }
  }
  tchunk[ntiles] = (int64_t)cbeg.size();
  auto* plan = new (std::nothrow) tamx_norm_plan();
  if (!plan) return TAMX_ERR_OOM;
  plan->ctx = ctx;
  plan->ntiles = ntiles;
  plan->nchunks = (int64_t)cbeg.size();
  const size_t nc = std::max<size_t>(1, cbeg.size());
  hipError_t e = hipSuccess;
  auto alloc = [&](auto** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  };
  alloc(&plan->cbeg, nc * sizeof(int64_t));
  alloc(&plan->cend, nc * sizeof(int64_t));
  alloc(&plan->ctile, nc * sizeof(int32_t));
  alloc(&plan->partial, nc * sizeof(double));
  alloc(&plan->partial_max, nc * sizeof(double));
  alloc(&plan->tchunk, (ntiles + 1) * sizeof(int64_t));
  if (e == hipSuccess && !cbeg.empty()) {
    e = hipMemcpy(plan->cbeg, cbeg.data(), cbeg.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(plan->cend, cend.data(), cend.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(plan->ctile, ctile.data(), ctile.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess)
    e = hipMemcpy(plan->tchunk, tchunk.data(), tchunk.size() * sizeof(int64_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    tamx::set_error(std::string("norm plan: ") + hipGetErrorString(e));
    tamx_norm_plan_destroy(plan);
    return e == hipErrorOutOfMemory ? TAMX_ERR_OOM : TAMX_ERR_DEVICE;
  }
  *out = plan;
  return TAMX_OK;
}

int tamx_norm_plan_create_cells(tamx_context* ctx, int64_t ntiles, const int64_t* host_tile_cells,
                                const int64_t* host_cell_off, const int64_t* host_cell_vol,
                                tamx_norm_plan** out) {
  TAMX_REQUIRE(out != nullptr, "null output pointer");
  *out = nullptr;
  TAMX_REQUIRE(ntiles >= 0, "negative tile count");
  TAMX_REQUIRE(ntiles == 0 || host_tile_cells, "null tile table");
  if (int rc = tamx::activate(ctx)) return rc;
  // the chunks of every cell of tile t, cell by cell: the finish sums them
  // in this order
  std::vector<int64_t> cbeg, cend, tchunk(ntiles + 1);
  std::vector<int32_t> ctile;
  for (int64_t t = 0; t < ntiles; ++t) {
    tchunk[t] = (int64_t)cbeg.size();
    const int64_t c0 = host_tile_cells[t], c1 = host_tile_cells[t + 1];
    TAMX_REQUIRE(c0 >= 0 && c1 >= c0, "tile cell ranges must be a non-decreasing prefix");
    TAMX_REQUIRE(c1 == c0 || (host_cell_off && host_cell_vol), "null cell table");
    for (int64_t c = c0; c < c1; ++c) {
      if (host_cell_off[c] < 0) continue;
      TAMX_REQUIRE(host_cell_vol[c] >= 0, "negative cell volume");
      for (int64_t b = 0; b < host_cell_vol[c]; b += kNormChunk) {
        cbeg.push_back(host_cell_off[c] + b);
        cend.push_back(host_cell_off[c] + std::min(host_cell_vol[c], b + kNormChunk));
        ctile.push_back((int32_t)t);
      }
    }
  }
  tchunk[ntiles] = (int64_t)cbeg.size();
  auto* plan = new (std::nothrow) tamx_norm_plan();
  if (!plan) return TAMX_ERR_OOM;
  plan->ctx = ctx;
  plan->ntiles = ntiles;
  plan->nchunks = (int64_t)cbeg.size();
  const size_t nc = std::max<size_t>(1, cbeg.size());
  hipError_t e = hipSuccess;
  auto alloc = [&](auto** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  };
  alloc(&plan->cbeg, nc * sizeof(int64_t));
  alloc(&plan->cend, nc * sizeof(int64_t));
  alloc(&plan->ctile, nc * sizeof(int32_t));
  alloc(&plan->partial, nc * sizeof(double));
  alloc(&plan->partial_max, nc * sizeof(double));
  alloc(&plan->tchunk, (ntiles + 1) * sizeof(int64_t));
  if (e == hipSuccess && !cbeg.empty()) {
    e = hipMemcpy(plan->cbeg, cbeg.data(), cbeg.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(plan->cend, cend.data(), cend.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(plan->ctile, ctile.data(), ctile.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess)
    e = hipMemcpy(plan->tchunk, tchunk.data(), tchunk.size() * sizeof(int64_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    tamx::set_error(std::string("norm plan: ") + hipGetErrorString(e));
    tamx_norm_plan_destroy(plan);
    return e == hipErrorOutOfMemory ? TAMX_ERR_OOM : TAMX_ERR_DEVICE;
  }
  *out = plan;
  return TAMX_OK;
}

int tamx_norm_plan_destroy(tamx_norm_plan* plan) {
  if (!plan) return TAMX_OK;
  void* ptrs[] = {plan->cbeg, plan->cend, plan->ctile, plan->partial, plan->partial_max,
                  plan->tchunk};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete plan;
  return TAMX_OK;
}

int tamx_tile_norms_f64(tamx_context* ctx, const tamx_norm_plan* plan,
                        const double* pool, float* norms, void* stream) {
  return tile_norms<double>(ctx, plan, pool, norms, nullptr, stream);
}

int tamx_tile_norms_max_f64(tamx_context* ctx, const tamx_norm_plan* plan,
                            const double* pool, float* norms, double* maxabs,
                            void* stream) {
  TAMX_REQUIRE(maxabs != nullptr, "null maxabs");
  return tile_norms<double>(ctx, plan, pool, norms, maxabs, stream);
}

int tamx_tile_norms_f32(tamx_context* ctx, const tamx_norm_plan* plan,
                        const float* pool, float* norms, void* stream) {
  return tile_norms<float>(ctx, plan, pool, norms, nullptr, stream);
}

int tamx_tile_norms_f64_noalloc(tamx_context* ctx, const int64_t* offsets,
                                const int64_t* volumes, int64_t ntiles, const double* pool,
                                int64_t pool_elems, float* norms, void* stream) {
  TAMX_REQUIRE(ctx != nullptr, "null context");
  TAMX_REQUIRE(ntiles >= 0 && pool_elems >= 0, "negative size");
  if (ntiles == 0) return TAMX_OK;
  TAMX_REQUIRE(offsets != nullptr && volumes != nullptr && norms != nullptr, "null argument");
  TAMX_REQUIRE(pool != nullptr || pool_elems == 0, "null pool");
  if (ntiles > TAMX_NOALLOC_MAX_TILES) {
    tamx::set_error("tamx_tile_norms_f64_noalloc: " + std::to_string(ntiles) +
                    " tiles exceed the limit of " + std::to_string(TAMX_NOALLOC_MAX_TILES));
    return TAMX_ERR_CAPACITY;
  }
  // tiles lying inside the pool have at most this many chunks in all
  const int64_t max_chunks = pool_elems / kNormChunk + ntiles;
  if (max_chunks > TAMX_NOALLOC_MAX_CHUNKS) {
    tamx::set_error("tamx_tile_norms_f64_noalloc: a pool of " + std::to_string(pool_elems) +
                    " elements in " + std::to_string(ntiles) + " tiles may need " +
                    std::to_string(max_chunks) + " chunks of 65536, beyond the limit of " +
                    std::to_string(TAMX_NOALLOC_MAX_CHUNKS));
    return TAMX_ERR_CAPACITY;
  }
  TAMX_REQUIRE(ctx->norm_scratch != nullptr, "context without a scratch address");
  if (int rc = tamx::activate(ctx)) return rc;
  hipStream_t s = tamx::as_stream(stream);
  auto* tchunk = static_cast<int64_t*>(ctx->norm_scratch);
  auto* partial = reinterpret_cast<double*>(static_cast<char*>(ctx->norm_scratch) +
                                            kNoallocPrefixBytes);
  norm_chunk_scan_kernel<<<1, 1024, 0, s>>>(offsets, volumes, ntiles, max_chunks, tchunk);
  TAMX_LAUNCH_CHECK();
  if (max_chunks > 0) {
    norm_partial_kernel<double, false><<<(unsigned)max_chunks, kNormThreads, 0, s>>>(
        pool, PrefixChunks{offsets, volumes, tchunk, ntiles}, partial, nullptr);
    TAMX_LAUNCH_CHECK();
  }
  norm_finish_prefix_kernel<double><<<(unsigned)((ntiles + 255) / 256), 256, 0, s>>>(
      tchunk, partial, ntiles, norms);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

int tamx_shape_scale(tamx_context* ctx, float* norms, int rank,
                     const int64_t* host_ntiles, int64_t nbatch,
                     const float* sizes, int sizes_per_batch, float threshold,
                     int64_t* zero_count, void* stream) {
  TAMX_REQUIRE(rank >= 1 && rank <= kMaxRank, "rank must be in [1, 8]");
  TAMX_REQUIRE(host_ntiles != nullptr && norms != nullptr && sizes != nullptr,
               "null argument");
  TAMX_REQUIRE(nbatch >= 1, "nbatch must be >= 1");
  if (int rc = tamx::activate(ctx)) return rc;
  ScaleParams p{};
  p.norms = norms;
  p.sizes = sizes;
  p.zero_count = zero_count;
  p.rank = rank;
  p.threshold = threshold;
  p.nbatch = nbatch;
  int64_t per = 1, off = 0;
  for (int d = 0; d < rank; ++d) {
    TAMX_REQUIRE(host_ntiles[d] >= 0, "negative tile count");
    p.ntiles[d] = host_ntiles[d];
    p.mode_off[d] = off;
    off += host_ntiles[d];
    per *= host_ntiles[d];
  }
  p.per_batch = per;
  p.sizes_batch_stride = sizes_per_batch ? off : 0;
  const int64_t n = per * nbatch;
  if (n == 0) return TAMX_OK;
  shape_scale_kernel<<<(unsigned)((n + 255) / 256), 256, 0, tamx::as_stream(stream)>>>(p);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

int tamx_shape_gemm_f32(tamx_context* ctx, const float* left,
                        const float* right, int op_a, int op_b, int64_t M,
                        int64_t N, int64_t K, int64_t nbatch,
                        const float* k_sizes, float abs_factor,
                        float threshold, float* result, int64_t* zero_count,
                        void* stream) {
  TAMX_REQUIRE(M >= 0 && N >= 0 && K >= 0 && nbatch >= 1, "bad sizes");
  TAMX_REQUIRE(left && right && result && (K == 0 || k_sizes), "null argument");
  TAMX_REQUIRE(op_a == 0 || op_a == 1, "op_a must be 0 or 1");
  TAMX_REQUIRE(op_b == 0 || op_b == 1, "op_b must be 0 or 1");
  if (int rc = tamx::activate(ctx)) return rc;
  ShapeGemmParams p{left, right, k_sizes, result, zero_count, M, N, K, nbatch,
                    op_a, op_b, abs_factor, threshold};
  const int64_t n = M * N * nbatch;
  if (n == 0) return TAMX_OK;
  shape_gemm_kernel<<<(unsigned)((n + 255) / 256), 256, 0, tamx::as_stream(stream)>>>(p);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

int tamx_contraction_worklist(tamx_context* ctx,
                              const tamx_contraction_desc* desc,
                              tamx_worklist* wl, void* stream) {
  TAMX_REQUIRE(desc != nullptr && wl != nullptr, "null argument");
  const tamx_contraction_desc& d = *desc;
  TAMX_REQUIRE(d.op_a == 0 || d.op_a == 1, "op_a must be 0 or 1");
  TAMX_REQUIRE(d.op_b == 0 || d.op_b == 1, "op_b must be 0 or 1");
  TAMX_REQUIRE(d.Mt >= 0 && d.Nt >= 0 && d.Kt >= 0 && d.nbatch >= 1, "bad grid");
  TAMX_REQUIRE(d.mext && d.next && d.kext && d.a_off && d.b_off && d.c_norms &&
                   d.c_off,
               "null descriptor array");
  const int64_t grid = d.Mt * d.Nt * d.nbatch;
  if (grid > wl->grid_capacity) {
    tamx::set_error("result tile grid exceeds the work list's grid capacity");
    return TAMX_ERR_CAPACITY;
  }
  if (int rc = tamx::activate(ctx)) return rc;
  WlParams p{d,
             wl->tasks,
             wl->task_tile,
             wl->pairs,
             wl->counters,
             wl->tile_npairs,
             wl->tile_volume,
             wl->tile_flops,
             static_cast<Quad*>(wl->scan_partials),
             wl->task_capacity,
             wl->pair_capacity};
  hipStream_t s = tamx::as_stream(stream);
  const unsigned g = (unsigned)std::max<int64_t>(1, (grid + kWlTiles - 1) / kWlTiles);
  wl_count_kernel<<<g, kWlThreads, 0, s>>>(p);
  TAMX_LAUNCH_CHECK();
  wl_scan_kernel<<<1, 1024, 0, s>>>(p, (int)g);
  TAMX_LAUNCH_CHECK();
  wl_fill_kernel<<<g, kWlThreads, 0, s>>>(p);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

}  // extern "C"
