// Per-tile index maps: the element offsets of a tile whose rows and
// columns (or k) are each a mixed-radix combination of up to four modes with
// arbitrary strides.  They let the grouped GEMMs read operands that need a
// general permutation straight from their stored tiles (gather plans) and
// store result tiles straight into a permuted order (scatter plans over the
// result grid), so no permuted copy is made on either side.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tamx_internal.hpp"

namespace tamx {
// One mode of a gathered tile's index map: extent, stride in the stored
// tile, and the constants that divide by the extent with a multiply and
// shifts (Granlund and Montgomery, "Division by invariant integers using
// multiplication", 1994, Fig. 4.1): with l = ceil(log2 ext),
// magic = floor(2^32 (2^l - ext) / ext) + 1, t = mulhi(magic, x),
// x / ext = (t + ((x - t) >> min(l, 1))) >> max(l - 1, 0), exact for every
// 32-bit x.
struct GatherMode {
  uint32_t magic;
  uint32_t shifts;  // min(l, 1) | max(l - 1, 0) << 8
  int32_t ext;
  int32_t str;
};
__host__ __device__ inline uint32_t gather_div(uint32_t u, const GatherMode& m) {
  const uint32_t t = (uint32_t)(((uint64_t)m.magic * u) >> 32);
  return (t + ((u - t) >> (m.shifts & 0xff))) >> (m.shifts >> 8);
}
// The mode (ext, str) with its division constants (ext 0 divides as 1).
inline GatherMode make_gather_mode(int32_t ext, int32_t str) {
  const uint64_t d = ext > 0 ? (uint64_t)ext : 1u;
  int l = 0;
  while ((uint64_t(1) << l) < d) ++l;
  GatherMode m;
  m.magic = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
  m.shifts = (uint32_t)(l < 1 ? l : 1) | ((uint32_t)(l > 1 ? l - 1 : 0) << 8);
  m.ext = ext;
  m.str = str;
  return m;
}
// Index map of one gathered tile: element (row, k) lies at the tile's offset
// + rowoff(row) + koff(k), each a mixed-radix sum over four modes (the last
// fastest) with those modes' strides in the stored tile; a map of fewer
// source modes is padded at the slow end with (extent 1, stride 0) modes.
// Offsets inside a tile stay below 2^31.
struct GatherTile {
  GatherMode r[4], k[4];
};
}  // namespace tamx

struct tamx_gather_plan {
  tamx_context* ctx = nullptr;
  int64_t ntiles = 0;
  tamx::GatherTile* tiles = nullptr;
};

namespace tamx {

__device__ inline int uni_map(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Offset of element x of a gathered tile's row (or k) index: the
// mixed-radix sum over the four modes in m (wave-uniform), the last
// fastest; three multiply-shift divisions, no branches.
__device__ inline int32_t gather_offset(int x, const GatherMode (&m)[4]) {
  uint32_t off = 0, u = (uint32_t)x;
#pragma unroll
  for (int d = 3; d >= 1; --d) {
    const uint32_t q = gather_div(u, m[d]);
    off += (u - q * (uint32_t)m[d].ext) * (uint32_t)m[d].str;
    u = q;
  }
  return (int32_t)(off + u * (uint32_t)m[0].str);
}

__device__ inline void load_modes(const __attribute__((address_space(4))) GatherMode* src,
                                  GatherMode (&dst)[4]) {
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    dst[d].magic = src[d].magic;
    dst[d].shifts = src[d].shifts;
    dst[d].ext = src[d].ext;
    dst[d].str = src[d].str;
  }
}

// Runs store(index) with the index function of a task's result tile:
// index(row, col) is where element (row, col) (tile coordinates) is stored,
// relative to the task's c_off: through the tile's scatter map when the
// launch has one (task_tile gives the task's grid position, the map's
// ordinal), else row-major with the task's ldc, transposed when the task
// asks for STORE_TRANS.  Each case is its own instantiation of store, so
// the maps stay in registers.
template <class Store>
__device__ inline void with_result_index(const GatherTile* scatter, const int32_t* task_tile,
                                         int task, int flags, int ldc, Store&& store) {
  if (scatter != nullptr) {
    // the map is read with vector loads: the store loops run after the
    // main loop, when registers are free, and kernels near their scalar
    // register limit would otherwise spill
    const int g = uni_map(scalar_ptr(task_tile)[task]);
    const auto* sc = global_ptr(scatter + g);
    GatherMode rm[4], cm[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      rm[d].magic = sc->r[d].magic;
      rm[d].shifts = sc->r[d].shifts;
      rm[d].ext = sc->r[d].ext;
      rm[d].str = sc->r[d].str;
      cm[d].magic = sc->k[d].magic;
      cm[d].shifts = sc->k[d].shifts;
      cm[d].ext = sc->k[d].ext;
      cm[d].str = sc->k[d].str;
    }
    store([&](int row, int col) -> int64_t {
      return (int64_t)gather_offset(row, rm) + gather_offset(col, cm);
    });
  } else if (flags & TAMX_TASK_STORE_TRANS) {
    store([&](int row, int col) -> int64_t { return (int64_t)col * ldc + row; });
  } else {
    store([&](int row, int col) -> int64_t { return (int64_t)row * ldc + col; });
  }
}

}  // namespace tamx
