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
};

// Makes the context's device current for the calling thread.
int activate(const Context* ctx);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

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
