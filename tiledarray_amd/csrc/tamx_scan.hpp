// Single-workgroup exclusive scan used by the work-list builder and the
// grouped-GEMM item planner.  Each thread owns a contiguous segment, so a
// scan of n values costs two passes over the data plus one log2(NT) step
// over per-thread partial sums.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tamx {

// load(i) returns the i-th value; store(i, prefix) receives the exclusive
// prefix.  Returns the total to every thread.  Must be called by all NT
// threads of the block; `part` is NT elements of shared memory.
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

}  // namespace tamx
