// Internal declarations shared by the HIP translation units of libtamx.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "tamx.h"

namespace tamx {

// Thread-local error message behind tamx_last_error().
void set_error(const std::string& msg);

struct Context {
  int device = 0;
  int num_cus = 0;
  int max_shared_per_block = 0;
  hipDeviceProp_t prop{};
  // device address of the static scratch of tamx_tile_norms_f64_noalloc
  void* norm_scratch = nullptr;
};

// Makes the context's device current for the calling thread.
int activate(const Context* ctx);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Descriptor tables a kernel only reads (work-list tasks and pairs, tile
// tables, all written by earlier launches), at wave-uniform indices: read
// through the constant address space, the loads go to the scalar unit and
// are not ordered behind the vector loads in flight, so fetching the next
// pair's descriptor never waits for a prefetched operand chunk.
template <typename T>
__device__ inline const __attribute__((address_space(4))) T* scalar_ptr(const T* p) {
  return (const __attribute__((address_space(4))) T*)(p);
}

// A pointer known to address global memory (so a select between two of
// them still compiles to global, not flat, loads).
template <typename T>
__device__ inline const __attribute__((address_space(1))) T* global_ptr(const T* p) {
  return (const __attribute__((address_space(1))) T*)(p);
}

}  // namespace tamx

struct tamx_context : tamx::Context {};

struct tamx_worklist {
  tamx_context* ctx = nullptr;
  tamx_gemm_task* tasks = nullptr;
  tamx_gemm_pair* pairs = nullptr;
  int64_t* counters = nullptr;  // TAMX_WL_NCOUNTERS int64
  int64_t task_capacity = 0;
  int64_t pair_capacity = 0;
  int64_t grid_capacity = 0;
  // scratch of the work-list builder (per result tile of the grid)
  int32_t* tile_npairs = nullptr;
  int64_t* tile_volume = nullptr;
  int64_t* tile_flops = nullptr;
  // scratch of the grouped GEMM: macro-tile prefix over tasks and the
  // persistent-kernel queue (next item, finished workgroups)
  int32_t* item_prefix = nullptr;  // task_capacity + 1
  // position of each task's result tile in the grid (written by
  // tamx_contraction_worklist; read by launches with a scatter plan)
  int32_t* task_tile = nullptr;
  int32_t* item_task = nullptr;    // item -> task map, item_capacity entries
  int64_t item_capacity = 0;
  int32_t* queue = nullptr;        // 16 int32 (per-XCD next item, done count), zero between launches
  // partial sums of the multi-workgroup scans: max(kScanBlocks, grid tiles
  // / 256 + 1) + 1 entries of 32 bytes
  void* scan_partials = nullptr;
};

namespace tamx {
constexpr int kScanBlocks = 256;  // workgroups of the multi-workgroup scans

// Fills wl->item_prefix (macro tiles of bm x bn per task, exclusive prefix,
// total at [ntasks]) and wl->item_task (item -> task) when the total fits
// wl->item_capacity.  Three launches on `s`; reads the task count from the
// work list's counters, so the sequence is graph-capturable.
int plan_items(tamx_worklist* wl, int bm, int bn, hipStream_t s);
// Resolves ctx->norm_scratch on ctx's device (which must be current).
int norm_scratch_address(tamx_context* ctx);
}

#define TAMX_CHECK_HIP(expr)                                                   \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      tamx::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));     \
      return TAMX_ERR_DEVICE;                                                  \
    }                                                                          \
  } while (0)

#define TAMX_REQUIRE(cond, msg)                                                \
  do {                                                                         \
    if (!(cond)) {                                                             \
      tamx::set_error(std::string("bad argument: ") + (msg));                 \
      return TAMX_ERR_BAD_ARG;                                                 \
    }                                                                          \
  } while (0)

#define TAMX_LAUNCH_CHECK()                                                    \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) {                                                    \
      tamx::set_error(std::string("kernel launch: ") + hipGetErrorString(e_)); \
      return TAMX_ERR_DEVICE;                                                  \
    }                                                                          \
  } while (0)
