// Context, status strings and work-list storage of libtamx.
#include <algorithm>
#include <cstring>
#include <new>

#include "tamx_internal.hpp"

namespace tamx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int activate(const Context* ctx) {
  TAMX_REQUIRE(ctx != nullptr, "null context");
  // HIP keeps the last failed call's error until hipGetLastError reads it;
  // a failure the caller (or this library, e.g. an allocation reported as
  // TAMX_ERR_OOM) has already seen must not be taken for this entry
  // point's launch failure
  (void)hipGetLastError();
  int cur = -1;
  TAMX_CHECK_HIP(hipGetDevice(&cur));
  if (cur != ctx->device) TAMX_CHECK_HIP(hipSetDevice(ctx->device));
  return TAMX_OK;
}

}  // namespace tamx

extern "C" {

const char* tamx_status_string(int status) {
  switch (status) {
    case TAMX_OK: return "ok";
    case TAMX_ERR_BAD_ARG: return "bad argument";
    case TAMX_ERR_OOM: return "out of device memory";
    case TAMX_ERR_UNSUPPORTED: return "unsupported";
    case TAMX_ERR_DEVICE: return "device error";
    case TAMX_ERR_CAPACITY: return "capacity exceeded";
    default: return "unknown status";
  }
}

const char* tamx_last_error(void) { return tamx::g_last_error.c_str(); }

const char* tamx_version(void) { return "tamx 0.1.0 gfx950"; }

int tamx_context_create(int device, tamx_context** out) {
  TAMX_REQUIRE(out != nullptr, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  TAMX_CHECK_HIP(hipGetDeviceCount(&ndev));
  TAMX_REQUIRE(device >= 0 && device < ndev, "device ordinal out of range");
  auto* ctx = new (std::nothrow) tamx_context();
  if (!ctx) return TAMX_ERR_OOM;
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipGetDeviceProperties(&ctx->prop, device);
  if (e != hipSuccess) {
    tamx::set_error(std::string("context: ") + hipGetErrorString(e));
    delete ctx;
    return TAMX_ERR_DEVICE;
  }
  ctx->num_cus = ctx->prop.multiProcessorCount;
  ctx->max_shared_per_block = static_cast<int>(ctx->prop.sharedMemPerBlock);
  if (std::strncmp(ctx->prop.gcnArchName, "gfx950", 6) != 0) {
    tamx::set_error(std::string("libtamx is built for gfx950, device is ") +
                    ctx->prop.gcnArchName);
    delete ctx;
    return TAMX_ERR_UNSUPPORTED;
  }
  if (int rc = tamx::norm_scratch_address(ctx)) {
    delete ctx;
    return rc;
  }
  *out = ctx;
  return TAMX_OK;
}

int tamx_context_destroy(tamx_context* ctx) {
  delete ctx;
  return TAMX_OK;
}

int tamx_context_num_cus(const tamx_context* ctx) {
  return ctx ? ctx->num_cus : 0;
}

int tamx_worklist_create(tamx_context* ctx, int64_t task_capacity,
                         int64_t pair_capacity, int64_t grid_capacity,
                         tamx_worklist** out) {
  TAMX_REQUIRE(out != nullptr, "null output pointer");
  *out = nullptr;
  TAMX_REQUIRE(task_capacity >= 0 && pair_capacity >= 0 && grid_capacity >= 0,
               "negative capacity");
  TAMX_REQUIRE(task_capacity < (int64_t(1) << 31) - 1 &&
                   pair_capacity < (int64_t(1) << 31) - 1 &&
                   grid_capacity < (int64_t(1) << 31) - 1,
               "capacity exceeds int32 indexing");
  if (int rc = tamx::activate(ctx)) return rc;
  auto* wl = new (std::nothrow) tamx_worklist();
  if (!wl) return TAMX_ERR_OOM;
  wl->ctx = ctx;
  wl->task_capacity = task_capacity;
  wl->pair_capacity = pair_capacity;
  wl->grid_capacity = grid_capacity;
  const int64_t tc = task_capacity > 0 ? task_capacity : 1;
  const int64_t pc = pair_capacity > 0 ? pair_capacity : 1;
  const int64_t gc = grid_capacity > 0 ? grid_capacity : 1;
  hipError_t e = hipSuccess;
  auto alloc = [&](auto** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  };
  alloc(&wl->tasks, sizeof(tamx_gemm_task) * tc);
  alloc(&wl->pairs, sizeof(tamx_gemm_pair) * pc);
  alloc(&wl->counters, sizeof(int64_t) * TAMX_WL_NCOUNTERS);
  alloc(&wl->tile_npairs, sizeof(int32_t) * gc);
  alloc(&wl->tile_volume, sizeof(int64_t) * gc);
  alloc(&wl->tile_flops, sizeof(int64_t) * gc);
  alloc(&wl->item_prefix, sizeof(int32_t) * (tc + 1));
  alloc(&wl->task_tile, sizeof(int32_t) * tc);
  // room for 16 macro tiles per task; larger lists fall back to a binary
  // search over item_prefix inside the GEMM kernel
  wl->item_capacity = std::min<int64_t>(tc * 16, (int64_t(1) << 31) - 1);
  alloc(&wl->item_task, sizeof(int32_t) * wl->item_capacity);
  alloc(&wl->queue, sizeof(int32_t) * 16);
  // one partial per work-list workgroup (32 result tiles each), at least
  // one per item-planner chunk
  const int64_t nparts = std::max<int64_t>(tamx::kScanBlocks, gc / 32 + 1) + 1;
  alloc(reinterpret_cast<char**>(&wl->scan_partials), 32 * nparts);
  if (e == hipSuccess) e = hipMemset(wl->counters, 0, sizeof(int64_t) * TAMX_WL_NCOUNTERS);
  if (e == hipSuccess) e = hipMemset(wl->queue, 0, sizeof(int32_t) * 16);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    tamx::set_error(std::string("worklist allocation: ") + hipGetErrorString(e));
    tamx_worklist_destroy(wl);
    return e == hipErrorOutOfMemory ? TAMX_ERR_OOM : TAMX_ERR_DEVICE;
  }
  *out = wl;
  return TAMX_OK;
}

int tamx_worklist_destroy(tamx_worklist* wl) {
  if (!wl) return TAMX_OK;
  void* ptrs[] = {wl->tasks,       wl->pairs,          wl->counters,
                  wl->tile_npairs, wl->tile_volume,    wl->tile_flops,
                  wl->item_prefix, wl->item_task, wl->queue, wl->scan_partials,
                  wl->task_tile};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete wl;
  return TAMX_OK;
}

int tamx_worklist_buffers(tamx_worklist* wl, tamx_gemm_task** tasks,
                          tamx_gemm_pair** pairs, int64_t** counters) {
  TAMX_REQUIRE(wl != nullptr, "null work list");
  if (tasks) *tasks = wl->tasks;
  if (pairs) *pairs = wl->pairs;
  if (counters) *counters = wl->counters;
  return TAMX_OK;
}

int tamx_worklist_read_counters(tamx_worklist* wl, int64_t* host, void* stream) {
  TAMX_REQUIRE(wl != nullptr && host != nullptr, "null argument");
  if (int rc = tamx::activate(wl->ctx)) return rc;
  TAMX_CHECK_HIP(hipMemcpyAsync(host, wl->counters,
                                sizeof(int64_t) * TAMX_WL_NCOUNTERS,
                                hipMemcpyDeviceToHost, tamx::as_stream(stream)));
  TAMX_CHECK_HIP(hipStreamSynchronize(tamx::as_stream(stream)));
  return TAMX_OK;
}

}  // extern "C"
