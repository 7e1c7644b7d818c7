// Host-side runtime services around the kernels: the tile-configuration
// autotuner of the grouped GEMM and stream capture into graphs.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "tamx_internal.hpp"

namespace {

// Shape class of a work list: everything the best tile shape depends on to
// first order.
using TuneKey = std::tuple<int, int, int, int, int, int, int, int>;  // dev, prec, ops, m, n, k, tasks
struct TuneEntry {
  tamx_gemm_config cfg;
  float ms;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tune_cache;

// half-octave bucket of a positive mean
int bucket(double x) { return x > 0 ? (int)std::lround(2.0 * std::log2(x)) : -1; }

struct WorkListSummary {
  int64_t ntasks = 0;
  double mean_m = 0, mean_n = 0, mean_k = 0;
  bool accumulates = false;  // some task has TAMX_TASK_BETA1
  int64_t c_extent = 0;      // elements of C the tasks touch, from offset 0
};

int summarize(tamx_worklist* wl, hipStream_t s, WorkListSummary& out) {
  int64_t cnt[TAMX_WL_NCOUNTERS];
  TAMX_CHECK_HIP(hipStreamSynchronize(s));
  TAMX_CHECK_HIP(hipMemcpy(cnt, wl->counters, sizeof cnt, hipMemcpyDeviceToHost));
  if (cnt[TAMX_WL_STATUS] != 0) {
    tamx::set_error("work list was not built (status " + std::to_string(cnt[TAMX_WL_STATUS]) + ")");
    return TAMX_ERR_BAD_ARG;
  }
  const int64_t nt = cnt[TAMX_WL_NTASKS], np = cnt[TAMX_WL_NPAIRS];
  if (nt < 0 || nt > wl->task_capacity || np < 0 || np > wl->pair_capacity) {
    tamx::set_error("work list counters out of range");
    return TAMX_ERR_BAD_ARG;
  }
  out.ntasks = nt;
  if (nt == 0) return TAMX_OK;
  std::vector<tamx_gemm_task> tasks(nt);
  std::vector<tamx_gemm_pair> pairs(np);
  TAMX_CHECK_HIP(hipMemcpy(tasks.data(), wl->tasks, sizeof(tamx_gemm_task) * nt,
                           hipMemcpyDeviceToHost));
  if (np > 0)
    TAMX_CHECK_HIP(hipMemcpy(pairs.data(), wl->pairs, sizeof(tamx_gemm_pair) * np,
                             hipMemcpyDeviceToHost));
  double sm = 0, sn = 0, sk = 0;
  for (const auto& t : tasks) {
    sm += t.m;
    sn += t.n;
    if (t.flags & TAMX_TASK_BETA1) out.accumulates = true;
    if (t.m > 0 && t.n > 0) {
      const bool tr = (t.flags & TAMX_TASK_STORE_TRANS) != 0;
      const int64_t rows = tr ? t.n : t.m, cols = tr ? t.m : t.n;
      out.c_extent = std::max<int64_t>(out.c_extent, t.c_off + (rows - 1) * t.ldc + cols);
    }
  }
  for (const auto& p : pairs) sk += p.k;
  out.mean_m = sm / nt;
  out.mean_n = sn / nt;
  out.mean_k = np ? sk / np : 0;
  return TAMX_OK;
}

struct HipEvent {
  hipEvent_t e = nullptr;
  ~HipEvent() {
    if (e) (void)hipEventDestroy(e);
  }
};
struct HipBuffer {
  void* p = nullptr;
  ~HipBuffer() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

struct tamx_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
};

extern "C" {

int tamx_autotune_grouped_gemm(tamx_context* ctx, tamx_worklist* wl, int precision,
                               int op_a, int op_b, const void* A, const void* B,
                               const void* C, int reps, int flags,
                               tamx_gemm_config* best, float* best_ms, void* stream) {
  return tamx_autotune_grouped_gemm_gather(ctx, wl, precision, op_a, op_b, A, nullptr, B,
                                           nullptr, C, reps, flags, best, best_ms, stream);
}

int tamx_autotune_grouped_gemm_gather(tamx_context* ctx, tamx_worklist* wl, int precision,
                                      int op_a, int op_b, const void* A,
                                      const tamx_gather_plan* a_gather, const void* B,
                                      const tamx_gather_plan* b_gather, const void* C,
                                      int reps, int flags, tamx_gemm_config* best,
                                      float* best_ms, void* stream) {
  const bool gathered = a_gather != nullptr || b_gather != nullptr;
  TAMX_REQUIRE(wl != nullptr && best != nullptr, "null argument");
  TAMX_REQUIRE(precision == TAMX_PREC_F64 || precision == TAMX_PREC_F32,
               "autotuning covers TAMX_PREC_F64 and TAMX_PREC_F32");
  TAMX_REQUIRE(op_a == 0 || op_a == 1, "op_a must be TAMX_NO_TRANS or TAMX_TRANS");
  TAMX_REQUIRE(op_b == 0 || op_b == 1, "op_b must be TAMX_NO_TRANS or TAMX_TRANS");
  TAMX_REQUIRE(reps >= 1 && reps <= 1000, "reps must be in [1, 1000]");
  if (int rc = tamx::activate(ctx)) return rc;
  hipStream_t s = tamx::as_stream(stream);
  WorkListSummary sum;
  if (int rc = summarize(wl, s, sum)) return rc;
  const int ncfg = tamx_gemm_config_count();
  if (sum.ntasks == 0) {
    tamx_gemm_config_get(0, best);
    if (best_ms) *best_ms = 0.0f;
    return TAMX_OK;
  }
  TAMX_REQUIRE(A != nullptr && B != nullptr && (C != nullptr || !sum.accumulates),
               "null operand");
  const TuneKey key{ctx->device, precision, op_a * 2 + op_b, bucket(sum.mean_m),
                    bucket(sum.mean_n), bucket(sum.mean_k), bucket((double)sum.ntasks),
                    (a_gather ? 1 : 0) + (b_gather ? 2 : 0)};
  if (!(flags & TAMX_TUNE_FORCE)) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      *best = it->second.cfg;
      if (best_ms) *best_ms = it->second.ms;
      return TAMX_OK;
    }
  }
  const size_t esz = precision == TAMX_PREC_F64 ? sizeof(double) : sizeof(float);
  const size_t cbytes = (size_t)std::max<int64_t>(sum.c_extent, 1) * esz;
  HipBuffer tmp;
  HipEvent e0, e1;
  TAMX_CHECK_HIP(hipMalloc(&tmp.p, cbytes));
  TAMX_CHECK_HIP(hipEventCreate(&e0.e));
  TAMX_CHECK_HIP(hipEventCreate(&e1.e));
  // C seeded into the scratch (when tasks accumulate), then one launch
  auto seed = [&]() -> int {
    if (sum.accumulates)
      TAMX_CHECK_HIP(hipMemcpyAsync(tmp.p, C, cbytes, hipMemcpyDeviceToDevice, s));
    return TAMX_OK;
  };
  auto launch = [&](const tamx_gemm_config& cfg) -> int {
    if (gathered && precision == TAMX_PREC_F64)
      return tamx_grouped_gemm_gather_f64(ctx, wl, op_a, op_b, 1.0,
                                          static_cast<const double*>(A), a_gather,
                                          static_cast<const double*>(B), b_gather,
                                          static_cast<double*>(tmp.p), &cfg, s);
    if (gathered)
      return tamx_grouped_gemm_gather_f32(ctx, wl, op_a, op_b, 1.0f,
                                          static_cast<const float*>(A), a_gather,
                                          static_cast<const float*>(B), b_gather,
                                          static_cast<float*>(tmp.p), &cfg, s);
    if (precision == TAMX_PREC_F64)
      return tamx_grouped_gemm_f64_config(ctx, wl, op_a, op_b, 1.0,
                                          static_cast<const double*>(A),
                                          static_cast<const double*>(B),
                                          static_cast<double*>(tmp.p), &cfg, s);
    return tamx_grouped_gemm_f32_config(ctx, wl, op_a, op_b, 1.0f,
                                        static_cast<const float*>(A),
                                        static_cast<const float*>(B),
                                        static_cast<float*>(tmp.p), &cfg, s);
  };
  float best_t = -1.0f;
  tamx_gemm_config win{};
  for (int c = 0; c < ncfg; ++c) {
    tamx_gemm_config cfg;
    tamx_gemm_config_get(c, &cfg);
    if (!(flags & TAMX_TUNE_NO_WARMUP)) {
      if (int rc = seed()) return rc;
      const int rc = launch(cfg);
      if (rc == TAMX_ERR_UNSUPPORTED && gathered) continue;  // not compiled gathered
      if (rc) return rc;
    }
    float total = 0.0f;
    bool skipped = false;
    for (int r = 0; r < reps; ++r) {
      if (int rc = seed()) return rc;
      TAMX_CHECK_HIP(hipEventRecord(e0.e, s));
      const int rc = launch(cfg);
      if (rc == TAMX_ERR_UNSUPPORTED && gathered) {
        skipped = true;
        break;
      }
      if (rc) return rc;
      TAMX_CHECK_HIP(hipEventRecord(e1.e, s));
      TAMX_CHECK_HIP(hipEventSynchronize(e1.e));
      float ms = 0.0f;
      TAMX_CHECK_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
      total += ms;
    }
    if (skipped) continue;
    const float mean = total / reps;
    if (best_t < 0.0f || mean < best_t) {
      best_t = mean;
      win = cfg;
    }
  }
  TAMX_CHECK_HIP(hipStreamSynchronize(s));
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_cache[key] = TuneEntry{win, best_t};
  }
  *best = win;
  if (best_ms) *best_ms = best_t;
  return TAMX_OK;
}

int tamx_autotune_grouped_gemm_split(tamx_context* ctx, tamx_worklist* wl,
                                     const tamx_split_plan* a_plan, const void* a_limbs,
                                     const tamx_split_plan* b_plan, const void* b_limbs,
                                     const double* C, int reps, int flags,
                                     tamx_gemm_config* best, float* best_ms, void* stream) {
  TAMX_REQUIRE(wl && a_plan && b_plan && a_limbs && b_limbs && best, "null argument");
  TAMX_REQUIRE(reps >= 1 && reps <= 1000, "reps must be in [1, 1000]");
  if (int rc = tamx::activate(ctx)) return rc;
  hipStream_t s = tamx::as_stream(stream);
  WorkListSummary sum;
  if (int rc = summarize(wl, s, sum)) return rc;
  const int nslices = tamx_split_plan_nslices(a_plan);
  const int limb_type = tamx_split_plan_limb_type(a_plan);
  // the split GEMM's macro tiles: 64 x 64 for any limb count, 128 x 64 and
  // 64 x 128 for 2 to 4 BF16 or 3 to 4 INT8 limbs; for INT8 limbs at 64 x 64
  // bk picks the kernel: 16, the 16x16x64 MFMA; 32, the 32x32x32 blocks with
  // a four-stage LDS ring of 32-k chunks; 64, the 32x32x32 blocks with two
  // 64-k buffers; 128, the two-wave kernel (4 or 5 limbs); from five INT8
  // limbs on, 128 x 64 is the eight-wave 16x16x64 kernel
  std::vector<tamx_gemm_config> cands{{64, 64, 32, 0}};
  if (limb_type == TAMX_LIMB_BF16 ? nslices <= 4 : (nslices >= 3 && nslices <= 4)) {
    cands.push_back({128, 64, 32, 0});
    cands.push_back({64, 128, 32, 0});
  }
  if (limb_type == TAMX_LIMB_I8) {
    cands.push_back({64, 64, 16, 0});
    cands.push_back({64, 64, 64, 0});
  }
  if (limb_type == TAMX_LIMB_I8 && (nslices == 4 || nslices == 5))
    cands.push_back({64, 64, 128, 0});
  if (limb_type == TAMX_LIMB_I8 && nslices >= 5) cands.push_back({128, 64, 16, 0});
  *best = cands[0];
  if (best_ms) *best_ms = 0.0f;
  if (sum.ntasks == 0) return TAMX_OK;
  TAMX_REQUIRE(C != nullptr || !sum.accumulates, "null C");
  const TuneKey key{ctx->device, 100 + 10 * limb_type + nslices, 0, bucket(sum.mean_m),
                    bucket(sum.mean_n),
                    bucket(sum.mean_k), bucket((double)sum.ntasks), 0};
  if (!(flags & TAMX_TUNE_FORCE)) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      *best = it->second.cfg;
      if (best_ms) *best_ms = it->second.ms;
      return TAMX_OK;
    }
  }
  const size_t cbytes = (size_t)std::max<int64_t>(sum.c_extent, 1) * sizeof(double);
  HipBuffer tmp;
  HipEvent e0, e1;
  TAMX_CHECK_HIP(hipMalloc(&tmp.p, cbytes));
  TAMX_CHECK_HIP(hipEventCreate(&e0.e));
  TAMX_CHECK_HIP(hipEventCreate(&e1.e));
  auto seed = [&]() -> int {
    if (sum.accumulates)
      TAMX_CHECK_HIP(hipMemcpyAsync(tmp.p, C, cbytes, hipMemcpyDeviceToDevice, s));
    return TAMX_OK;
  };
  auto launch = [&](const tamx_gemm_config& cfg) -> int {
    return tamx_grouped_gemm_split_config(ctx, wl, 1.0, a_plan, a_limbs, b_plan, b_limbs,
                                          static_cast<double*>(tmp.p), &cfg, s);
  };
  float best_t = -1.0f;
  tamx_gemm_config win = cands[0];
  for (const auto& cfg : cands) {
    if (!(flags & TAMX_TUNE_NO_WARMUP)) {
      if (int rc = seed()) return rc;
      if (int rc = launch(cfg)) return rc;
    }
    float total = 0.0f;
    for (int r = 0; r < reps; ++r) {
      if (int rc = seed()) return rc;
      TAMX_CHECK_HIP(hipEventRecord(e0.e, s));
      if (int rc = launch(cfg)) return rc;
      TAMX_CHECK_HIP(hipEventRecord(e1.e, s));
      TAMX_CHECK_HIP(hipEventSynchronize(e1.e));
      float ms = 0.0f;
      TAMX_CHECK_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
      total += ms;
    }
    const float mean = total / reps;
    if (best_t < 0.0f || mean < best_t) {
      best_t = mean;
      win = cfg;
    }
  }
  TAMX_CHECK_HIP(hipStreamSynchronize(s));
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_cache[key] = TuneEntry{win, best_t};
  }
  *best = win;
  if (best_ms) *best_ms = best_t;
  return TAMX_OK;
}

int tamx_autotune_clear(void) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune_cache.clear();
  return TAMX_OK;
}

int tamx_graph_begin(tamx_context* ctx, void* stream) {
  TAMX_REQUIRE(stream != nullptr, "graph capture needs a created stream, not the null stream");
  if (int rc = tamx::activate(ctx)) return rc;
  TAMX_CHECK_HIP(hipStreamBeginCapture(tamx::as_stream(stream), hipStreamCaptureModeThreadLocal));
  return TAMX_OK;
}

int tamx_graph_end(tamx_context* ctx, void* stream, tamx_graph** out) {
  TAMX_REQUIRE(out != nullptr, "null output pointer");
  *out = nullptr;
  TAMX_REQUIRE(stream != nullptr, "graph capture needs a created stream, not the null stream");
  if (int rc = tamx::activate(ctx)) return rc;
  auto* g = new (std::nothrow) tamx_graph();
  if (!g) return TAMX_ERR_OOM;
  g->device = ctx->device;
  hipError_t e = hipStreamEndCapture(tamx::as_stream(stream), &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    tamx::set_error(std::string("graph capture: ") + hipGetErrorString(e));
    tamx_graph_destroy(g);
    return TAMX_ERR_DEVICE;
  }
  *out = g;
  return TAMX_OK;
}

int tamx_graph_launch(tamx_graph* g, void* stream) {
  TAMX_REQUIRE(g != nullptr && g->exec != nullptr, "null graph");
  TAMX_CHECK_HIP(hipSetDevice(g->device));
  TAMX_CHECK_HIP(hipGraphLaunch(g->exec, tamx::as_stream(stream)));
  return TAMX_OK;
}

int tamx_graph_destroy(tamx_graph* g) {
  if (!g) return TAMX_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return TAMX_OK;
}

}  // extern "C"
