// Persistent grouped GEMM over a device work list (FP64 and FP32 MFMA).
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
#include <algorithm>
#include <mutex>
#include <unordered_map>

#include "tamx_internal.hpp"
#include "tamx_scan.hpp"

namespace {

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
};

constexpr int kThreads = 256;

// prefix[t] = number of BM x BN macro tiles of tasks 0..t-1, and the
// item -> task map when it fits in item_capacity entries.
__global__ __launch_bounds__(1024) void plan_items_kernel(
    const tamx_gemm_task* __restrict__ tasks, const int64_t* __restrict__ counters,
    int bm, int bn, int32_t* __restrict__ prefix, int32_t* __restrict__ item_task,
    int64_t capacity, int64_t item_capacity) {
  __shared__ int32_t part[1024];
  int64_t n = counters[TAMX_WL_NTASKS];
  if (n < 0 || n > capacity || counters[TAMX_WL_STATUS] != 0) n = 0;
  auto count = [&](int64_t i) -> int32_t {
    const tamx_gemm_task t = tasks[i];
    return ((t.m + bm - 1) / bm) * ((t.n + bn - 1) / bn);
  };
  const int32_t total = tamx::block_exclusive_scan<1024, int32_t>(
      n, part, count, [&](int64_t i, int32_t v) { prefix[i] = v; });
  if (threadIdx.x == 0) prefix[n] = total;
  if (total <= item_capacity) {
    // each thread fills the items of a contiguous range of tasks
    const int64_t seg = (n + 1023) / 1024;
    const int64_t b = min<int64_t>(n, threadIdx.x * seg), e = min<int64_t>(n, b + seg);
    for (int64_t i = b; i < e; ++i) {
      const int32_t q0 = prefix[i], q1 = q0 + count(i);
      for (int32_t q = q0; q < q1; ++q) item_task[q] = (int32_t)i;
    }
  }
}

template <typename T, int BM, int BN, int BK, bool TA, bool TB>
__global__ __launch_bounds__(kThreads) void grouped_gemm_kernel(GemmParams<T> p) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr int FM = BM / 32, FN = BN / 32;  // 16x16 blocks per wave (2x2 waves)
  constexpr int A_LD = BM * BK / kThreads, B_LD = BK * BN / kThreads;
  constexpr int SA = BM + 1, SB = BN + 1;  // padded LDS rows
  static_assert(BM % 32 == 0 && BN % 32 == 0 && BK % 4 == 0, "tile shape");
  static_assert((BM * BK) % kThreads == 0 && (BK * BN) % kThreads == 0, "load split");

  __shared__ T sA[2][BK][SA];
  __shared__ T sB[2][BK][SB];
  __shared__ int s_item;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * (BM / 2), wc = (wave & 1) * (BN / 2);
  const int lrow = lane & 15, lk = lane >> 4;

  int64_t ntasks64 = p.counters[TAMX_WL_NTASKS];
  // same validation as the planner: an invalid list has no items
  if (ntasks64 < 0 || ntasks64 > p.task_capacity || p.counters[TAMX_WL_STATUS] != 0)
    ntasks64 = 0;
  const int32_t ntasks = (int32_t)ntasks64;
  const int32_t total = p.prefix[ntasks > 0 ? ntasks : 0];

  for (;;) {
    if (tid == 0) s_item = atomicAdd(&p.queue[0], 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= total) break;

    int lo;
    if (total <= p.item_capacity) {
      lo = p.item_task[item];
    } else {
      lo = 0;
      int hi = ntasks - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p.prefix[mid] <= item) lo = mid; else hi = mid - 1;
      }
    }
    const tamx_gemm_task task = p.tasks[lo];
    const int local = item - p.prefix[lo];
    const int tiles_n = (task.n + BN - 1) / BN;
    const int r0 = (local / tiles_n) * BM, c0 = (local % tiles_n) * BN;
    const int mrem = min(BM, task.m - r0), nrem = min(BN, task.n - c0);

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[i][j][v] = T(0);

    T ra[A_LD], rb[B_LD];
    auto load_chunk = [&](const tamx_gemm_pair& pr, int k0) {
      const int kc = min(BK, pr.k - k0);
      const T* __restrict__ a = p.A + pr.a_off;
      const T* __restrict__ b = p.B + pr.b_off;
#pragma unroll
      for (int e = 0; e < A_LD; ++e) {
        const int idx = tid + e * kThreads;
        int r, kk;
        if (!TA) { kk = idx % BK; r = idx / BK; }
        else { r = idx % BM; kk = idx / BM; }
        T v = T(0);
        if (r < mrem && kk < kc)
          v = TA ? a[(int64_t)(k0 + kk) * pr.lda + r0 + r]
                 : a[(int64_t)(r0 + r) * pr.lda + k0 + kk];
        ra[e] = v;
      }
#pragma unroll
      for (int e = 0; e < B_LD; ++e) {
        const int idx = tid + e * kThreads;
        int c, kk;
        if (!TB) { c = idx % BN; kk = idx / BN; }
        else { kk = idx % BK; c = idx / BK; }
        T v = T(0);
        if (c < nrem && kk < kc)
          v = TB ? b[(int64_t)(c0 + c) * pr.ldb + k0 + kk]
                 : b[(int64_t)(k0 + kk) * pr.ldb + c0 + c];
        rb[e] = v;
      }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
      for (int e = 0; e < A_LD; ++e) {
        const int idx = tid + e * kThreads;
        int r, kk;
        if (!TA) { kk = idx % BK; r = idx / BK; }
        else { r = idx % BM; kk = idx / BM; }
        sA[buf][kk][r] = ra[e];
      }
#pragma unroll
      for (int e = 0; e < B_LD; ++e) {
        const int idx = tid + e * kThreads;
        int c, kk;
        if (!TB) { c = idx % BN; kk = idx / BN; }
        else { kk = idx % BK; c = idx / BK; }
        sB[buf][kk][c] = rb[e];
      }
    };
    // 16x16 blocks that lie wholly outside a ragged edge are skipped (the
    // branch is uniform across the wave)
    int vi = 0, vj = 0;
#pragma unroll
    for (int i = 0; i < FM; ++i) vi += (wr + i * 16 < mrem) ? 1 : 0;
#pragma unroll
    for (int j = 0; j < FN; ++j) vj += (wc + j * 16 < nrem) ? 1 : 0;
    auto compute_chunk = [&](int buf) {
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        T af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
          if (i < vi) af[i] = sA[buf][ks + lk][wr + i * 16 + lrow];
#pragma unroll
        for (int j = 0; j < FN; ++j)
          if (j < vj) bf[j] = sB[buf][ks + lk][wc + j * 16 + lrow];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            if (i < vi && j < vj) acc[i][j] = M::op(af[i], bf[j], acc[i][j]);
      }
    };

    int pi = task.pair_begin;
    const int pend = task.pair_begin + task.pair_count;
    if (pi < pend) {
      tamx_gemm_pair cur = p.pairs[pi];
      int k0 = 0;
      load_chunk(cur, k0);
      store_chunk(0);
      __syncthreads();
      int buf = 0;
      // advance to the chunk after (pi, k0)
      k0 += BK;
      if (k0 >= cur.k) {
        ++pi;
        k0 = 0;
        if (pi < pend) cur = p.pairs[pi];
      }
      for (;;) {
        const bool has_next = pi < pend;
        if (has_next) load_chunk(cur, k0);
        compute_chunk(buf);
        if (!has_next) break;
        store_chunk(buf ^ 1);
        __syncthreads();
        buf ^= 1;
        k0 += BK;
        if (k0 >= cur.k) {
          ++pi;
          k0 = 0;
          if (pi < pend) cur = p.pairs[pi];
        }
      }
    }

    const bool beta1 = (task.flags & TAMX_TASK_BETA1) != 0;
    const bool trans = (task.flags & TAMX_TASK_STORE_TRANS) != 0;
    T* __restrict__ c = p.C + task.c_off;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = wr + i * 16 + M::row(lk, v);
          const int col = wc + j * 16 + lrow;
          if (row < mrem && col < nrem) {
            const int64_t idx = trans ? (int64_t)(c0 + col) * task.ldc + r0 + row
                                      : (int64_t)(r0 + row) * task.ldc + c0 + col;
            T val = p.alpha * acc[i][j][v];
            if (beta1) val += c[idx];
            c[idx] = val;
          }
        }
  }

  if (tid == 0) {
    __threadfence();
    const int done = atomicAdd(&p.queue[1], 1);
    if (done == (int)gridDim.x - 1) {
      atomicExch(&p.queue[0], 0);
      atomicExch(&p.queue[1], 0);
    }
  }
}

struct TileCfg {
  int bm, bn;
};

template <typename T, int BM, int BN, int BK, bool TA, bool TB>
int launch_cfg(tamx_context* ctx, tamx_worklist* wl, const GemmParams<T>& prm,
               hipStream_t s, int wgs_per_cu) {
  auto kern = grouped_gemm_kernel<T, BM, BN, BK, TA, TB>;
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
  plan_items_kernel<<<1, 1024, 0, s>>>(wl->tasks, wl->counters, BM, BN,
                                       wl->item_prefix, wl->item_task,
                                       wl->task_capacity, wl->item_capacity);
  TAMX_LAUNCH_CHECK();
  kern<<<grid, kThreads, 0, s>>>(prm);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

template <typename T, int BM, int BN, int BK>
int launch_ops(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
               const GemmParams<T>& prm, hipStream_t s, int wgs) {
  if (!op_a && !op_b) return launch_cfg<T, BM, BN, BK, false, false>(ctx, wl, prm, s, wgs);
  if (!op_a && op_b) return launch_cfg<T, BM, BN, BK, false, true>(ctx, wl, prm, s, wgs);
  if (op_a && !op_b) return launch_cfg<T, BM, BN, BK, true, false>(ctx, wl, prm, s, wgs);
  return launch_cfg<T, BM, BN, BK, true, true>(ctx, wl, prm, s, wgs);
}

// Macro-tile shapes compiled into the library (BM, BN, BK).
constexpr int kConfigs[][3] = {{64, 64, 16}, {64, 64, 32}, {128, 128, 16},
                               {128, 64, 16}, {64, 128, 16}, {32, 32, 16},
                               {32, 32, 32}, {128, 64, 32}, {64, 128, 32}};

template <typename T>
int grouped_gemm(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
                 T alpha, const T* A, const T* B, T* C, void* stream,
                 const tamx_gemm_config* cfg) {
  TAMX_REQUIRE(wl != nullptr, "null work list");
  TAMX_REQUIRE(op_a == 0 || op_a == 1, "op_a must be TAMX_NO_TRANS or TAMX_TRANS");
  TAMX_REQUIRE(op_b == 0 || op_b == 1, "op_b must be TAMX_NO_TRANS or TAMX_TRANS");
  if (int rc = tamx::activate(ctx)) return rc;
  GemmParams<T> prm{A, B, C, wl->tasks, wl->pairs, wl->counters,
                    wl->item_prefix, wl->item_task, wl->queue,
                    wl->task_capacity, wl->item_capacity, alpha};
  hipStream_t s = tamx::as_stream(stream);
  const int bm = cfg ? cfg->bm : 64, bn = cfg ? cfg->bn : 64, bk = cfg ? cfg->bk : 16;
  const int w = cfg ? cfg->wgs_per_cu : 0;
#define TAMX_CFG(BM_, BN_, BK_) \
  if (bm == BM_ && bn == BN_ && bk == BK_) return launch_ops<T, BM_, BN_, BK_>(ctx, wl, op_a, op_b, prm, s, w);
  TAMX_CFG(64, 64, 16)
  TAMX_CFG(64, 64, 32)
  TAMX_CFG(128, 128, 16)
  TAMX_CFG(128, 64, 16)
  TAMX_CFG(64, 128, 16)
  TAMX_CFG(32, 32, 16)
  TAMX_CFG(32, 32, 32)
  TAMX_CFG(128, 64, 32)
  TAMX_CFG(64, 128, 32)
#undef TAMX_CFG
  tamx::set_error("unsupported macro tile: see tamx_gemm_config_list");
  return TAMX_ERR_UNSUPPORTED;
}

}  // namespace

extern "C" {

int tamx_gemm_config_count(void) { return (int)(sizeof(kConfigs) / sizeof(kConfigs[0])); }

int tamx_gemm_config_get(int index, tamx_gemm_config* cfg) {
  TAMX_REQUIRE(cfg != nullptr, "null config");
  TAMX_REQUIRE(index >= 0 && index < tamx_gemm_config_count(), "config index out of range");
  cfg->bm = kConfigs[index][0];
  cfg->bn = kConfigs[index][1];
  cfg->bk = kConfigs[index][2];
  cfg->wgs_per_cu = 0;
  return TAMX_OK;
}

int tamx_grouped_gemm_f64_config(tamx_context* ctx, tamx_worklist* wl, int op_a,
                                 int op_b, double alpha, const double* A,
                                 const double* B, double* C,
                                 const tamx_gemm_config* cfg, void* stream) {
  return grouped_gemm<double>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, cfg);
}

int tamx_grouped_gemm_f32_config(tamx_context* ctx, tamx_worklist* wl, int op_a,
                                 int op_b, float alpha, const float* A,
                                 const float* B, float* C,
                                 const tamx_gemm_config* cfg, void* stream) {
  return grouped_gemm<float>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, cfg);
}

int tamx_grouped_gemm_f64(tamx_context* ctx, tamx_worklist* wl, int op_a,
                          int op_b, double alpha, const double* A,
                          const double* B, double* C, void* stream) {
  return grouped_gemm<double>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, nullptr);
}

int tamx_grouped_gemm_f32(tamx_context* ctx, tamx_worklist* wl, int op_a,
                          int op_b, float alpha, const float* A,
                          const float* B, float* C, void* stream) {
  return grouped_gemm<float>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, nullptr);
}

}  // extern "C"
