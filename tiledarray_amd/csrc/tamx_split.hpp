// Declarations shared by the split-precision translation units: the limb
// tile table, the split plan, the GEMM kernel parameters and the LDS layout
// of limb chunks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tamx_internal.hpp"
#include "tamx_map.hpp"

namespace tamx {
namespace split {

struct SplitTile {
  int64_t src_off;   // element offset of the FP64 tile
  int64_t limb_off;  // element offset of limb plane 0 in the limb pool
  int32_t rows, cols;  // stored FP64 tile (row-major)
  int32_t out_rows;    // logical rows of the limb layout (m for A, n for B)
  int32_t out_k;       // logical K
  int32_t rows_pad;    // out_rows rounded up to 16
  int32_t k_pad;       // out_k rounded up to 32 (row stride of a plane)
  int32_t transpose;   // limb (row, k) reads stored (k, row)
  int32_t sigma;       // scale exponent, written by the max pass
  // general layout (gen_nr > 0): limb (row, k) reads source element
  // src_off + rowoff(row) + koff(k), each a mixed-radix sum over up to four
  // source modes (last fastest) with the modes' source strides; this is how
  // an operand that needs a general permutation is split straight from its
  // stored layout, without a permuted copy
  int32_t gen_nr, gen_nk;
  int32_t gr_ext[4], gk_ext[4];
  int64_t gr_str[4], gk_str[4];
};

typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef int i4_t __attribute__((ext_vector_type(4)));
typedef int i16_t __attribute__((ext_vector_type(16)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d4_t __attribute__((ext_vector_type(4)));

struct SplitParams {
  const bf8_t* __restrict__ LA;  // limb pools viewed as 8-element vectors
  const bf8_t* __restrict__ LB;
  const SplitTile* __restrict__ ta;
  const SplitTile* __restrict__ tb;
  double* __restrict__ C;
  const tamx_gemm_task* __restrict__ tasks;
  const tamx_gemm_pair* __restrict__ pairs;
  const int64_t* __restrict__ counters;
  const int32_t* __restrict__ prefix;
  const int32_t* __restrict__ item_task;
  int32_t* queue;
  int64_t task_capacity;
  int64_t item_capacity;
  double alpha;
  // result tiles stored through a scatter map (null: row-major with the
  // task's ldc); task_tile[t] is task t's position in the result grid
  const tamx::GatherTile* __restrict__ c_scatter;
  const int32_t* __restrict__ task_tile;
};

constexpr int kThreads = 256;
constexpr int kRegions = 8;     // XCDs of an MI355X; the queue holds kRegions + 1 counters
constexpr int SBK = 32;         // K per chunk (one BF16 MFMA 16x16x32)
constexpr int kFoldChunks = 4;  // fold FP32 levels into FP64 every 128 k
// INT8 limbs: K per chunk (one INT8 MFMA 16x16x64); the INT32 levels are
// exact and folded at pair ends or every 64 chunks (4096 k), far inside
// their range
constexpr int SBK8 = 64;
constexpr int kFoldChunks8 = 64;
// INT8 limb planes are stored in blocks of 32 k: within a block the 32 limbs
// of a row are contiguous and the block's rows_pad rows follow each other, so
// the chunk of a macro tile's rows that a GEMM loads is one contiguous run
// per plane and block (whole cache lines, rather than 32 or 64 bytes per row).
constexpr int kI8Block = 32;
// the 16-limb unit u (k 16u .. 16u + 15) of row r in a plane starting at
// unit `plane0`
__host__ __device__ inline int64_t i8_unit(int64_t plane0, int rows_pad, int r, int u) {
  return plane0 + ((int64_t)(u >> 1) * rows_pad + r) * 2 + (u & 1);
}
// rows of slack past the last tile of a limb pool: a macro tile may read up
// to this many rows beyond a tile's padded rows (those rows only reach
// results that are not stored)
constexpr int kLimbSlackRows = 128;

__device__ inline int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ inline int64_t uni64(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// LDS layout of one operand chunk: S planes x R rows x 32 k of BF16, as
// 16-byte units (8 k each); unit q of row r sits at r * 4 + (q ^ ((r >> 1) & 3)),
// so any 8 consecutive rows read at one unit cover all eight 16-byte slots
// of a 128-byte bank window (and the 4 units of 2 rows that 8 staging lanes
// write do too).
__device__ inline int lds_unit(int r, int q) { return r * 4 + (q ^ ((r >> 1) & 3)); }

// The grouped split GEMM with a bm x bn macro tile of 128 x 64 or 64 x 128
// for 2 to 4 limbs (tamx_split_wide.hip); plans its items and launches.
int launch_split_gemm_wide(int limb_type, int nslices, int bm, int bn, tamx_context* ctx,
                           tamx_worklist* wl, const SplitParams& prm, hipStream_t s);
// The grouped split GEMM for 4 or 5 INT8 limbs with two-wave workgroups,
// 64 x 64 macro tiles, A limbs through an LDS ring and B limbs straight into
// registers (tamx_split_i8.hip); plans its items and launches.
int launch_split_gemm_i8_2w(int nslices, tamx_context* ctx, tamx_worklist* wl,
                            const SplitParams& prm, hipStream_t s);
// The grouped split GEMM for INT8 limbs with four-wave workgroups, 64 x 64
// macro tiles on 32x32x32 blocks and a four-stage LDS ring of 32-k chunks
// (tamx_split_i8.hip); plans its items and launches.
int launch_split_gemm_i8_ring(int nslices, tamx_context* ctx, tamx_worklist* wl,
                              const SplitParams& prm, hipStream_t s);

}  // namespace split
}  // namespace tamx

struct tamx_split_plan {
  tamx_context* ctx = nullptr;
  int64_t ntiles = 0;
  int32_t nslices = 0;
  int32_t limb_type = 0;    // TAMX_LIMB_BF16 or TAMX_LIMB_I8
  int64_t limb_elems = 0;   // size of the limb pool in limbs (bf16 or int8)
  int64_t nblocks = 0;      // 16 x 128 blocks of the limb layouts over all tiles
  tamx::split::SplitTile* tiles = nullptr;  // device
  int64_t* block_prefix = nullptr;          // device, ntiles + 1
  unsigned long long* maxbits = nullptr;    // device, per-tile max |a| bits
};
