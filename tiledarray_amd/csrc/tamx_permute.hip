// Tile permutation (TA "TO" convention) for many tiles in one launch.
//
// A plan fixes, per tile, the fused source extents and the source/result
// strides.  Consecutive source modes that stay consecutive in the result are
// fused first, as detail::permute does (tensor/permute.h:51-95).  Two cases:
//   - the fastest source mode stays fastest: contiguous runs are copied;
//   - otherwise the (fastest source mode s) x (mode t that becomes fastest
//     in the result) plane is moved through LDS so that both the loads and
//     the stores are contiguous along a row: whole when it has at most
//     96 x 96 elements, else in blocks of at most that many; every other
//     mode is a batch index.
// Batch indices run fastest-in-the-result first, so that a workgroup's
// consecutive items write neighbouring result rows; the partial cache line
// where two such rows meet is written by both items together.
// Both are HBM-bandwidth bound: 2 * volume * sizeof(T) bytes per tile.
#include <algorithm>
#include <new>
#include <vector>

#include "tamx_internal.hpp"

namespace {

constexpr int kMaxRank = 8;
constexpr int kCopyRun = 4096;  // elements per copy item (eight per thread)
// a transposed block holds at most kPlaneMax elements; its LDS rows are
// padded to an odd length, which needs at most kPlaneEdge more
constexpr int kPlaneMax = 96 * 96;
constexpr int kPlaneEdge = 96;
constexpr int kPermThreads = 512;

struct PermTile {
  int64_t src_off, dst_off;
  int64_t item_begin;
  int64_t sstride[kMaxRank];
  int64_t dstride[kMaxRank];
  int32_t ext[kMaxRank];
  int32_t q;      // fused rank
  int32_t s_dim;  // fastest source mode (q - 1)
  int32_t t_dim;  // source mode that is fastest in the result
  int32_t bls, blt;  // block extents along s and t
  int32_t nbs, nbt;  // blocks along s and t
  // batch modes (every mode but s and t; but s for a copy), fastest in the
  // result first: consecutive items of one workgroup write neighbouring
  // result rows, so the cache lines two rows share are filled in one L2
  int32_t nbatch;
  int32_t border[kMaxRank];
};

}  // namespace

struct tamx_permute_plan {
  tamx_context* ctx = nullptr;
  int64_t ntiles = 0;
  int64_t nitems = 0;
  int64_t elem_bytes_moved = 0;  // sum of volumes
  PermTile* tiles = nullptr;     // device
  int64_t* item_prefix = nullptr;  // device, ntiles + 1
};

namespace {

__device__ inline int64_t find_tile(const int64_t* prefix, int64_t n, int64_t item) {
  int64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Barrier for the LDS staging area only: this wave's LDS reads and writes
// have completed, its global stores may still be in flight (a
// __syncthreads() would wait for them too, draining the workgroup's stores
// at every item).
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// source and result offsets of item `local`'s batch indices
__device__ inline void batch_offsets(const PermTile& t, int local, int64_t& so, int64_t& doff) {
  so = t.src_off;
  doff = t.dst_off;
  for (int i = 0; i < t.nbatch; ++i) {
    const int d = t.border[i];
    const int x = local % t.ext[d];
    local /= t.ext[d];
    so += x * t.sstride[d];
    doff += x * t.dstride[d];
  }
}

// contiguous-run copy: item = (run block, other modes); every element a
// thread moves is loaded before any is stored
template <typename T>
__device__ inline void copy_item(const T* __restrict__ src, T* __restrict__ dst,
                                 const PermTile& t, int local) {
  const int sd = t.s_dim;
  const int nruns = (t.ext[sd] + kCopyRun - 1) / kCopyRun;
  const int rb = local % nruns;
  int64_t so, doff;
  batch_offsets(t, local / nruns, so, doff);
  const int b0 = rb * kCopyRun;
  const int n = min(kCopyRun, t.ext[sd] - b0);
  const T* __restrict__ sp = src + so + b0;
  T* __restrict__ dp = dst + doff + b0;
  constexpr int U = kCopyRun / kPermThreads;
  T v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int x = threadIdx.x + u * kPermThreads;
    v[u] = x < n ? sp[x] : T(0);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int x = threadIdx.x + u * kPermThreads;
    if (x < n) dp[x] = v[u];
  }
}

// One block of the (s, t) plane of a transposed tile: ns x nt elements
// (s fastest in the source, t fastest in the result), at most kPlaneMax.
struct Block {
  int64_t so, doff;  // element offsets of the block's first element
  int ns, nt;
};

__device__ inline int64_t uniform(int64_t x) {
  const uint64_t u = (uint64_t)x;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// item `local` = (s block, t block, batch indices)
__device__ inline Block block_of(const PermTile& t, int local) {
  const int bs = local % t.nbs;
  local /= t.nbs;
  const int bt = local % t.nbt;
  local /= t.nbt;
  Block b;
  batch_offsets(t, local, b.so, b.doff);
  const int s0 = bs * t.bls, t0 = bt * t.blt;
  b.ns = min(t.bls, t.ext[t.s_dim] - s0);
  b.nt = min(t.blt, t.ext[t.t_dim] - t0);
  b.so += s0 + (int64_t)t0 * t.sstride[t.t_dim];
  b.doff += t0 + (int64_t)s0 * t.dstride[t.s_dim];
  // the block is the same for every lane: keep it in scalar registers
  b.so = uniform(b.so);
  b.doff = uniform(b.doff);
  b.ns = __builtin_amdgcn_readfirstlane(b.ns);
  b.nt = __builtin_amdgcn_readfirstlane(b.nt);
  return b;
}

// Element e = tid + u * blockDim of an ns x nt block is row e / ns, column
// e % ns: one division, then each step advances with one carry.  Every
// element a thread moves is loaded into registers at once (the most bytes
// in flight).
constexpr int kPlaneRegs = (kPlaneMax + kPermThreads - 1) / kPermThreads;

template <typename T>
__device__ inline void block_load(const T* __restrict__ src, const PermTile& t, const Block& b,
                                  T (&v)[kPlaneRegs]) {
  const int ns = b.ns, n = ns * b.nt;
  const int qs = blockDim.x / ns, rs = blockDim.x - qs * ns;
  int y = threadIdx.x / ns, x = threadIdx.x - y * ns;
  // each pass recomputes its indices: shared across passes they would
  // hold two registers per element for the whole item
  asm volatile("" : "+v"(x), "+v"(y));
  // offsets inside one tile fit 32 bits (tile volumes are below 2^31)
  const T* __restrict__ sp = src + b.so;
  const uint32_t st = (uint32_t)t.sstride[t.t_dim];
#pragma unroll
  for (int u = 0; u < kPlaneRegs; ++u) {
    v[u] = threadIdx.x + u * (int)blockDim.x < n ? sp[(uint32_t)y * st + (uint32_t)x] : T(0);
    x += rs;
    y += qs;
    if (x >= ns) {
      x -= ns;
      ++y;
    }
  }
}

// the loaded elements into LDS, laid out [t][s] with an odd row stride
template <typename T>
__device__ inline void block_to_lds(const Block& b, const T (&v)[kPlaneRegs], T* pl) {
  const int ns = b.ns, n = ns * b.nt;
  const int ld = ns | 1;
  const int qs = blockDim.x / ns, rs = blockDim.x - qs * ns;
  int y = threadIdx.x / ns, x = threadIdx.x - y * ns;
  // each pass recomputes its indices: shared across passes they would
  // hold two registers per element for the whole item
  asm volatile("" : "+v"(x), "+v"(y));
#pragma unroll
  for (int u = 0; u < kPlaneRegs; ++u) {
    if (threadIdx.x + u * (int)blockDim.x < n) pl[y * ld + x] = v[u];
    x += rs;
    y += qs;
    if (x >= ns) {
      x -= ns;
      ++y;
    }
  }
}

// Cache-line ends of a result row of nt elements whose first element lies
// at byte address `lo` (low 32 bits suffice): the elements before its first
// line boundary (head) and after its last (tail).
template <typename T>
__device__ inline int head_len(uint32_t lo, int nt) {
  return min(nt, (int)(((128u - (lo & 127u)) & 127u) / sizeof(T)));
}
template <typename T>
__device__ inline int tail_len(uint32_t lo, int nt) {
  return (int)(((lo + (uint32_t)nt * sizeof(T)) & 127u) / sizeof(T));
}

// whether every result row of the block starts and ends on a cache line
template <typename T>
__device__ inline bool lines_whole(const T* dst, const PermTile& t, const Block& b) {
  const uint32_t lo = (uint32_t)(uintptr_t)(dst + b.doff);
  return ((lo | ((uint32_t)t.dstride[t.s_dim] * (uint32_t)sizeof(T)) |
           ((uint32_t)b.nt * (uint32_t)sizeof(T))) & 127u) == 0;
}

// The block from LDS to the result, rows along the result's fastest mode:
// columns [lo, hi) of every row, where lo skips the head when the previous
// item wrote it (`skip_head`) and hi the tail when the caller writes it
// later (`skip_tail`); `tail_only` writes just those tails.
template <typename T>
__device__ inline void block_store(T* __restrict__ dst, const PermTile& t, const Block& b,
                                   const T* pl, bool skip_head, bool skip_tail, bool tail_only) {
  const int ns = b.ns, nt = b.nt, n = ns * nt;
  const int ld = ns | 1;
  const int qt = blockDim.x / nt, rt = blockDim.x - qt * nt;
  int y = threadIdx.x / nt, x = threadIdx.x - y * nt;  // s index, t index
  const uint32_t st = (uint32_t)t.dstride[t.s_dim];
  T* __restrict__ dp = dst + b.doff;
  const uint32_t lo = (uint32_t)(uintptr_t)dp;
  // a loop rather than a full unroll: it indexes no register array, and
  // the next block's loads hold registers meanwhile
#pragma unroll 6
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t row = (uint32_t)y * st;  // offsets inside a tile fit 32 bits
    const int h = skip_head ? head_len<T>(lo + row * (uint32_t)sizeof(T), nt) : 0;
    const int e = skip_tail ? max(h, nt - tail_len<T>(lo + row * (uint32_t)sizeof(T), nt)) : nt;
    if (tail_only ? x >= e : (x >= h && x < e)) dp[row + (uint32_t)x] = pl[x * ld + y];
    x += rt;
    y += qt;
    if (x >= nt) {
      x -= nt;
      ++y;
    }
  }
}

// The heads of the next block's result rows, from the registers its loads
// landed in (the load order: element e is s index e % ns, t index e / ns).
template <typename T>
__device__ inline void block_store_heads(T* __restrict__ dst, const PermTile& t, const Block& b,
                                         const T (&v)[kPlaneRegs]) {
  const int ns = b.ns, nt = b.nt, n = ns * nt;
  const int qs = blockDim.x / ns, rs = blockDim.x - qs * ns;
  int y = threadIdx.x / ns, x = threadIdx.x - y * ns;  // t index, s index
  // each pass recomputes its indices: shared across passes they would
  // hold two registers per element for the whole item
  asm volatile("" : "+v"(x), "+v"(y));
  const uint32_t st = (uint32_t)t.dstride[t.s_dim];
  T* __restrict__ dp = dst + b.doff;
  const uint32_t lo = (uint32_t)(uintptr_t)dp;
#pragma unroll
  for (int u = 0; u < kPlaneRegs; ++u) {
    // a head has fewer elements than a cache line: t index below 128 / sizeof(T)
    if (u * (int)blockDim.x >= (128 / (int)sizeof(T)) * ns) break;
    if (threadIdx.x + u * (int)blockDim.x < n) {
      const uint32_t row = (uint32_t)x * st;
      if (y < head_len<T>(lo + row * (uint32_t)sizeof(T), nt)) dp[row + (uint32_t)y] = v[u];
    }
    x += rs;
    y += qs;
    if (x >= ns) {
      x -= ns;
      ++y;
    }
  }
}

// Each workgroup takes a contiguous range of items, so consecutive items
// mostly belong to one tile: the tile is found by binary search once and
// then followed forward.  Blocks of one tile are pipelined: the next block
// is loaded while this one is stored.
template <typename T>
__global__ __launch_bounds__(kPermThreads) void permute_kernel(
    const T* __restrict__ src, T* __restrict__ dst,
    const PermTile* __restrict__ tiles, const int64_t* __restrict__ prefix,
    int64_t ntiles) {
  static_assert(kPlaneRegs * kPermThreads >= kPlaneMax, "registers hold a block");
  static_assert(kCopyRun % kPermThreads == 0, "a copy item is whole passes of the workgroup");
  __shared__ T plane[kPlaneMax + kPlaneEdge];  // rows padded to an odd length
  const int64_t nitems = prefix[ntiles];
  const int64_t per = (nitems + gridDim.x - 1) / gridDim.x;
  const int64_t ib = min<int64_t>(nitems, (int64_t)blockIdx.x * per);
  const int64_t ie = min<int64_t>(nitems, ib + per);
  if (ib >= ie) return;
  int64_t ti = find_tile(prefix, ntiles, ib);
  int64_t tnext = prefix[ti + 1];
  T v[kPlaneRegs];
  bool have = false;  // v holds the block of `item`
  bool prev = false;  // the previous item wrote this block's row heads
  Block cur{}, nxt{};
  for (int64_t item = ib; item < ie; ++item) {
    while (item >= tnext) {
      ++ti;
      tnext = prefix[ti + 1];
    }
    const PermTile& t = tiles[ti];
    const int local = (int)(item - prefix[ti]);
    if (t.s_dim == t.t_dim) {
      copy_item<T>(src, dst, t, local);
      continue;
    }
    if (!have) {
      cur = block_of(t, local);
      block_load<T>(src, t, cur, v);
    } else {
      cur = nxt;
    }
    block_to_lds<T>(cur, v, plane);
    lds_barrier();
    have = item + 1 < ie && item + 1 < tnext;
    if (have) {
      nxt = block_of(t, local + 1);  // next block, same tile
      block_load<T>(src, t, nxt, v);
    }
    // When the next block's rows continue this one's in memory (the same
    // shape, its result starting where this block's first row ends, as for
    // consecutive indices of the batch mode just slower than t), each
    // row's last, partial cache line is shared with the next block's row:
    // this block's row tails and the next block's row heads are then
    // written back to back, so both parts of the line reach the L2
    // together instead of an item apart.
    const bool join = have && nxt.ns == cur.ns && nxt.nt == cur.nt &&
                      nxt.doff == cur.doff + cur.nt && !lines_whole<T>(dst, t, cur);
    block_store<T>(dst, t, cur, plane, prev, join, false);
    if (join) {
      block_store<T>(dst, t, cur, plane, prev, true, true);
      block_store_heads<T>(dst, t, nxt, v);
    }
    prev = join;
    lds_barrier();
  }
}

}  // namespace

// Blocks of an ns x nt plane: the whole plane when it fits the staging
// area, else the long side (or both) split evenly so that every block does.
static bool split_plane(int64_t ns, int64_t nt, PermTile& t) {
  auto fits = [](int64_t bs, int64_t bt) {
    return bs * bt <= kPlaneMax && (bs | 1) * bt <= kPlaneMax + kPlaneEdge;
  };
  int64_t nbs = 1, nbt = 1;
  if (!fits(ns, nt)) {
    if (ns <= 96) {
      nbt = (nt + kPlaneMax / (ns | 1) - 1) / (kPlaneMax / (ns | 1));
    } else if (nt <= 96) {
      nbs = (ns + kPlaneMax / nt - 1) / (kPlaneMax / nt);
    } else {
      nbs = (ns + 95) / 96;
      nbt = (nt + 95) / 96;
    }
  }
  t.bls = (int32_t)((ns + nbs - 1) / nbs);
  t.blt = (int32_t)((nt + nbt - 1) / nbt);
  t.nbs = (int32_t)((ns + t.bls - 1) / t.bls);
  t.nbt = (int32_t)((nt + t.blt - 1) / t.blt);
  return fits(t.bls, t.blt);
}

static unsigned permute_grid(const tamx_context* ctx, const tamx_permute_plan* plan) {
  // grid-stride over items: enough workgroups to fill every CU several times
  const int64_t cap = (int64_t)ctx->num_cus * 32;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(plan->nitems, cap));
}

extern "C" {

// Plan for permuting `ntiles` tiles of one rank with one permutation.
// host_extents: ntiles x rank source extents; host_src_off/host_dst_off:
// element offsets of each tile in the source and result pools.  Replaces
// per-tile Tensor(other, perm) / detail::permute (tensor/tensor.h:644-692,
// tensor/permute.h:121-207) and the LibreTT plan of the device build
// (external/librett.h:79-111).
int tamx_permute_plan_create(tamx_context* ctx, int rank, const int32_t* perm,
                             int64_t ntiles, const int64_t* host_extents,
                             const int64_t* host_src_off,
                             const int64_t* host_dst_off,
                             tamx_permute_plan** out) {
  TAMX_REQUIRE(out != nullptr, "null output pointer");
  *out = nullptr;
  TAMX_REQUIRE(rank >= 1 && rank <= kMaxRank, "rank must be in [1, 8]");
  TAMX_REQUIRE(perm != nullptr, "null permutation");
  TAMX_REQUIRE(ntiles >= 0, "negative tile count");
  TAMX_REQUIRE(ntiles == 0 || (host_extents && host_src_off && host_dst_off),
               "null tile table");
  {
    bool seen[kMaxRank] = {false};
    for (int d = 0; d < rank; ++d) {
      TAMX_REQUIRE(perm[d] >= 0 && perm[d] < rank && !seen[perm[d]],
                   "perm is not a permutation of 0..rank-1");
      seen[perm[d]] = true;
    }
  }
  if (int rc = tamx::activate(ctx)) return rc;
  std::vector<PermTile> tiles(ntiles);
  std::vector<int64_t> prefix(ntiles + 1, 0);
  int64_t items = 0, moved = 0;
  for (int64_t ti = 0; ti < ntiles; ++ti) {
    const int64_t* ext = host_extents + ti * rank;
    int64_t vol = 1;
    for (int d = 0; d < rank; ++d) {
      TAMX_REQUIRE(ext[d] >= 0 && ext[d] < (int64_t(1) << 31), "bad extent");
      vol *= ext[d];
    }
    // drop extent-1 modes (they move no data), renumbering the permutation
    int keep[kMaxRank], nk = 0;
    for (int d = 0; d < rank; ++d)
      if (ext[d] != 1) keep[nk++] = d;
    int64_t kext[kMaxRank];
    int kperm[kMaxRank];
    for (int a = 0; a < nk; ++a) {
      kext[a] = ext[keep[a]];
      int r = 0;  // rank of perm[keep[a]] among the kept result positions
      for (int b = 0; b < nk; ++b)
        if (perm[keep[b]] < perm[keep[a]]) ++r;
      kperm[a] = r;
    }
    if (nk == 0) {  // a single element (or empty) tile
      kext[0] = vol;
      kperm[0] = 0;
      nk = 1;
    }
    // result extents and strides of the reduced permutation
    int64_t rext[kMaxRank], rstride[kMaxRank], sstride[kMaxRank];
    for (int d = 0; d < nk; ++d) rext[kperm[d]] = kext[d];
    int64_t s = 1;
    for (int d = nk - 1; d >= 0; --d) { rstride[d] = s; s *= rext[d]; }
    s = 1;
    for (int d = nk - 1; d >= 0; --d) { sstride[d] = s; s *= kext[d]; }
    // fuse runs of source modes that stay adjacent and in order
    PermTile& t = tiles[ti];
    t = PermTile{};
    t.src_off = host_src_off[ti];
    t.dst_off = host_dst_off[ti];
    int q = 0;
    for (int d = 0; d < nk; ++d) {
      if (q > 0 && kperm[d] == kperm[d - 1] + 1) {
        // merge mode d into the previous fused mode (row-major within it)
        t.ext[q - 1] = (int32_t)(t.ext[q - 1] * kext[d]);
        t.sstride[q - 1] = sstride[d];
        t.dstride[q - 1] = rstride[kperm[d]];
      } else {
        t.ext[q] = (int32_t)kext[d];
        t.sstride[q] = sstride[d];
        t.dstride[q] = rstride[kperm[d]];
        ++q;
      }
    }
    TAMX_REQUIRE(vol < (int64_t(1) << 31), "tile volume exceeds int32 indexing");
    t.q = q;
    t.s_dim = q - 1;
    // the fused source mode with result stride 1 is fastest in the result
    t.t_dim = q - 1;
    for (int d = 0; d < q; ++d)
      if (t.dstride[d] == 1) t.t_dim = d;
    t.nbatch = 0;
    for (int d = 0; d < q; ++d)
      if (d != t.s_dim && d != t.t_dim) t.border[t.nbatch++] = d;
    std::sort(t.border, t.border + t.nbatch,
              [&](int x, int y) { return t.dstride[x] < t.dstride[y]; });
    int64_t nitems = 0;
    if (vol == 0) {
      nitems = 0;
    } else if (t.t_dim == t.s_dim) {
      nitems = (t.ext[t.s_dim] + kCopyRun - 1) / kCopyRun;
      for (int d = 0; d < q; ++d)
        if (d != t.s_dim) nitems *= t.ext[d];
    } else {
      TAMX_REQUIRE(split_plane(t.ext[t.s_dim], t.ext[t.t_dim], t),
                   "permute block exceeds the staging area");
      nitems = (int64_t)t.nbs * t.nbt;
      for (int d = 0; d < q; ++d)
        if (d != t.s_dim && d != t.t_dim) nitems *= t.ext[d];
    }
    t.item_begin = items;
    prefix[ti] = items;
    items += nitems;
    moved += vol;
  }
  prefix[ntiles] = items;
  TAMX_REQUIRE(items < (int64_t(1) << 31), "too many permute blocks for one launch");
  auto* plan = new (std::nothrow) tamx_permute_plan();
  if (!plan) return TAMX_ERR_OOM;
  plan->ctx = ctx;
  plan->ntiles = ntiles;
  plan->nitems = items;
  plan->elem_bytes_moved = moved;
  hipError_t e = hipMalloc(&plan->tiles, std::max<size_t>(1, tiles.size()) * sizeof(PermTile));
  if (e == hipSuccess) e = hipMalloc(&plan->item_prefix, prefix.size() * sizeof(int64_t));
  if (e == hipSuccess && !tiles.empty())
    e = hipMemcpy(plan->tiles, tiles.data(), tiles.size() * sizeof(PermTile), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(plan->item_prefix, prefix.data(), prefix.size() * sizeof(int64_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    tamx::set_error(std::string("permute plan: ") + hipGetErrorString(e));
    if (plan->tiles) (void)hipFree(plan->tiles);
    if (plan->item_prefix) (void)hipFree(plan->item_prefix);
    delete plan;
    return e == hipErrorOutOfMemory ? TAMX_ERR_OOM : TAMX_ERR_DEVICE;
  }
  *out = plan;
  return TAMX_OK;
}

int tamx_permute_plan_destroy(tamx_permute_plan* plan) {
  if (!plan) return TAMX_OK;
  if (plan->tiles) (void)hipFree(plan->tiles);
  if (plan->item_prefix) (void)hipFree(plan->item_prefix);
  delete plan;
  return TAMX_OK;
}

// Number of elements one execution of the plan moves (each read once and
// written once).
int64_t tamx_permute_plan_volume(const tamx_permute_plan* plan) {
  return plan ? plan->elem_bytes_moved : 0;
}

int tamx_permute_f64(tamx_context* ctx, const tamx_permute_plan* plan,
                     const double* src, double* dst, void* stream) {
  TAMX_REQUIRE(plan != nullptr, "null plan");
  if (plan->nitems == 0) return TAMX_OK;
  TAMX_REQUIRE(src != nullptr && dst != nullptr && src != dst,
               "src and dst must be distinct non-null pools");
  if (int rc = tamx::activate(ctx)) return rc;
  permute_kernel<double><<<permute_grid(ctx, plan), kPermThreads, 0, tamx::as_stream(stream)>>>(
      src, dst, plan->tiles, plan->item_prefix, plan->ntiles);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

int tamx_permute_f32(tamx_context* ctx, const tamx_permute_plan* plan,
                     const float* src, float* dst, void* stream) {
  TAMX_REQUIRE(plan != nullptr, "null plan");
  if (plan->nitems == 0) return TAMX_OK;
  TAMX_REQUIRE(src != nullptr && dst != nullptr && src != dst,
               "src and dst must be distinct non-null pools");
  if (int rc = tamx::activate(ctx)) return rc;
  permute_kernel<float><<<permute_grid(ctx, plan), kPermThreads, 0, tamx::as_stream(stream)>>>(
      src, dst, plan->tiles, plan->item_prefix, plan->ntiles);
  TAMX_LAUNCH_CHECK();
  return TAMX_OK;
}

}  // extern "C"
