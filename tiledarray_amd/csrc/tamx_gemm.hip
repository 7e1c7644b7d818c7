// Plain entry points of the grouped GEMM and its item planner (the kernel
// is in tamx_gemm.hpp).
#include "tamx_gemm.hpp"
#include "tamx_scan.hpp"

namespace {

// Item planner: prefix[t] = number of BM x BN macro tiles of tasks 0..t-1,
// and the item -> task map when it fits in item_capacity entries.  Three
// launches (chunk sums over kScanBlocks workgroups, one-workgroup scan of the
// sums, chunk rescans), each walking the task list once.
__device__ inline int64_t planned_tasks(const int64_t* counters, int64_t capacity) {
  const int64_t n = counters[TAMX_WL_NTASKS];
  return (n < 0 || n > capacity || counters[TAMX_WL_STATUS] != 0) ? 0 : n;
}
__device__ inline int32_t task_items(const tamx_gemm_task* tasks, int64_t i, int bm,
                                     int bn) {
  const tamx_gemm_task t = tasks[i];
  return ((t.m + bm - 1) / bm) * ((t.n + bn - 1) / bn);
}

constexpr int kPlanThreads = 256;

__global__ __launch_bounds__(kPlanThreads) void plan_items_sum_kernel(
    const tamx_gemm_task* __restrict__ tasks, const int64_t* __restrict__ counters,
    int bm, int bn, int64_t capacity, int32_t* __restrict__ partial) {
  __shared__ int32_t red[kPlanThreads];
  const int64_t n = planned_tasks(counters, capacity);
  tamx::scan_phase1<kPlanThreads, int32_t>(
      n, red, [&](int64_t i) { return task_items(tasks, i, bm, bn); }, partial);
}

__global__ __launch_bounds__(1024) void plan_items_offsets_kernel(int32_t* partial) {
  __shared__ int32_t part[1024];
  tamx::scan_phase2<1024, int32_t>(tamx::kScanBlocks, part, partial);
}

__global__ __launch_bounds__(kPlanThreads) void plan_items_fill_kernel(
    const tamx_gemm_task* __restrict__ tasks, const int64_t* __restrict__ counters,
    int bm, int bn, int64_t capacity, const int32_t* __restrict__ partial,
    int32_t* __restrict__ prefix, int32_t* __restrict__ item_task,
    int64_t item_capacity) {
  __shared__ int32_t part[kPlanThreads];
  const int64_t n = planned_tasks(counters, capacity);
  const int32_t total = partial[tamx::kScanBlocks];
  const bool map = total <= item_capacity;
  auto count = [&](int64_t i) { return task_items(tasks, i, bm, bn); };
  tamx::scan_phase3<kPlanThreads, int32_t>(
      n, part, count,
      [&](int64_t i, int32_t q0) {
        prefix[i] = q0;
        if (map) {
          const int32_t q1 = q0 + count(i);
          for (int32_t q = q0; q < q1; ++q) item_task[q] = (int32_t)i;
        }
      },
      partial);
  if (blockIdx.x == 0 && threadIdx.x == 0) prefix[n] = total;
}

}  // namespace

int tamx::plan_items(tamx_worklist* wl, int bm, int bn, hipStream_t s) {
  auto* partial = static_cast<int32_t*>(wl->scan_partials);
  plan_items_sum_kernel<<<tamx::kScanBlocks, kPlanThreads, 0, s>>>(
      wl->tasks, wl->counters, bm, bn, wl->task_capacity, partial);
  TAMX_LAUNCH_CHECK();
  plan_items_offsets_kernel<<<1, 1024, 0, s>>>(partial);
  TAMX_LAUNCH_CHECK();
  plan_items_fill_kernel<<<tamx::kScanBlocks, kPlanThreads, 0, s>>>(
      wl->tasks, wl->counters, bm, bn, wl->task_capacity, partial, wl->item_prefix,
      wl->item_task, wl->item_capacity);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

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

}  // extern "C"

int tamx::plain_gemm_f64(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
                         double alpha, const double* A, const double* B, double* C,
                         const tamx_gemm_config* cfg, const GatherTile* c_scatter,
                         void* stream) {
  return grouped_gemm<double, 0>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, cfg, nullptr,
                                 nullptr, c_scatter);
}

int tamx::plain_gemm_f32(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
                         float alpha, const float* A, const float* B, float* C,
                         const tamx_gemm_config* cfg, const GatherTile* c_scatter,
                         void* stream) {
  return grouped_gemm<float, 0>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, cfg, nullptr,
                                nullptr, c_scatter);
}

extern "C" {

int tamx_grouped_gemm_f64_config(tamx_context* ctx, tamx_worklist* wl, int op_a,
                                 int op_b, double alpha, const double* A,
                                 const double* B, double* C,
                                 const tamx_gemm_config* cfg, void* stream) {
  return grouped_gemm<double, 0>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, cfg);
}

int tamx_grouped_gemm_f32_config(tamx_context* ctx, tamx_worklist* wl, int op_a,
                                 int op_b, float alpha, const float* A,
                                 const float* B, float* C,
                                 const tamx_gemm_config* cfg, void* stream) {
  return grouped_gemm<float, 0>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, cfg);
}

int tamx_grouped_gemm_f64(tamx_context* ctx, tamx_worklist* wl, int op_a,
                          int op_b, double alpha, const double* A,
                          const double* B, double* C, void* stream) {
  return grouped_gemm<double, 0>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, nullptr);
}

int tamx_grouped_gemm_f32(tamx_context* ctx, tamx_worklist* wl, int op_a,
                          int op_b, float alpha, const float* A,
                          const float* B, float* C, void* stream) {
  return grouped_gemm<float, 0>(ctx, wl, op_a, op_b, alpha, A, B, C, stream, nullptr);
}

}  // extern "C"
