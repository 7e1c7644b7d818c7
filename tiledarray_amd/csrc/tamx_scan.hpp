// Exclusive scans used by the work-list builder and the grouped-GEMM item
// planner: a single-workgroup scan, and the three phases of a scan spread
// over many workgroups.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tamx {

// load(i) returns the i-th value; store(i, prefix) receives the exclusive
// prefix.  Returns the total to every thread.  Must be called by all NT
// threads of the block; `part` is NT elements of shared memory.
// Each thread owns a contiguous segment, so a scan of n values costs two
// passes over the data plus one log2(NT) step over per-thread sums.
template <int NT, typename V, typename Load, typename Store>
__device__ inline V block_exclusive_scan(int64_t n, V* part, Load load,
                                         Store store) {
  const int t = threadIdx.x;
  const int64_t seg = (n + NT - 1) / NT;
  const int64_t b = (int64_t)t * seg < n ? (int64_t)t * seg : n;
  const int64_t e = b + seg < n ? b + seg : n;
  V sum(0);
  for (int64_t i = b; i < e; ++i) sum += load(i);
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < NT; off <<= 1) {
    const V v = t >= off ? part[t - off] : V(0);
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  V run = t > 0 ? part[t - 1] : V(0);
  const V total = part[NT - 1];
  for (int64_t i = b; i < e; ++i) {
    const V v = load(i);
    store(i, run);
    run += v;
  }
  __syncthreads();
  return total;
}

// Multi-workgroup exclusive scan in three launches, for lists too long for
// one workgroup to walk quickly.  Element range of workgroup b of nb:
// [b * chunk, min(n, (b + 1) * chunk)) with chunk = ceil(n / nb).
__device__ inline void scan_chunk(int64_t n, int nb, int b, int64_t& lo, int64_t& hi) {
  const int64_t chunk = (n + nb - 1) / nb;
  lo = (int64_t)b * chunk < n ? (int64_t)b * chunk : n;
  hi = lo + chunk < n ? lo + chunk : n;
}

// Phase 1: partial[b] = sum of the chunk of workgroup b.
template <int NT, typename V, typename Load>
__device__ inline void scan_phase1(int64_t n, V* red, Load load, V* partial) {
  int64_t lo, hi;
  scan_chunk(n, gridDim.x, blockIdx.x, lo, hi);
  V s(0);
  for (int64_t i = lo + threadIdx.x; i < hi; i += NT) s += load(i);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Phase 2 (one workgroup): partial[0..nb) -> exclusive offsets, total at
// partial[nb].
template <int NT, typename V>
__device__ inline void scan_phase2(int nb, V* part_lds, V* partial) {
  const V total = block_exclusive_scan<NT, V>(
      nb, part_lds, [&](int64_t i) -> V { return partial[i]; },
      [&](int64_t i, const V& v) { partial[i] = v; });
  if (threadIdx.x == 0) partial[nb] = total;
}

// Phase 3: each workgroup scans its chunk from its offset; store(i, prefix).
template <int NT, typename V, typename Load, typename Store>
__device__ inline void scan_phase3(int64_t n, V* part_lds, Load load, Store store,
                                   const V* partial) {
  int64_t lo, hi;
  scan_chunk(n, gridDim.x, blockIdx.x, lo, hi);
  const V base = partial[blockIdx.x];
  block_exclusive_scan<NT, V>(
      hi - lo, part_lds, [&](int64_t i) -> V { return load(lo + i); },
      [&](int64_t i, const V& v) {
        V x = base;
        x += v;
        store(lo + i, x);
      });
}

}  // namespace tamx
