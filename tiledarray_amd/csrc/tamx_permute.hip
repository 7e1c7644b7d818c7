// Tile permutation (TA "TO" convention) for many tiles in one launch.
//
// A plan fixes, per tile, the fused source extents and the source/result
// strides.  Consecutive source modes that stay consecutive in the result are
// fused first, as detail::permute does (tensor/permute.h:51-95).  Two cases:
//   - the fastest source mode stays fastest: contiguous runs are copied;
//   - otherwise the (fastest source mode) x (mode that becomes fastest in the
//     result) plane is moved through LDS so that both the loads and the
//     stores are contiguous along a row: whole when it has at most 96 x 96
//     elements, else in padded 32 x 32 tiles; every other mode is a batch
//     index.
// Both are HBM-bandwidth bound: 2 * volume * sizeof(T) bytes per tile.
#include <algorithm>
#include <new>
#include <vector>

#include "tamx_internal.hpp"

namespace {

constexpr int kMaxRank = 8;
constexpr int kTS = 32;       // transpose tile edge
constexpr int kCopyRun = 2048;  // elements per copy item
// a transposed plane of at most this many elements moves through LDS whole
// (one item per plane, no partial edge tiles)
constexpr int kPlaneMax = 96 * 96;
constexpr int kPermThreads = 512;
constexpr int kRowsPer = kPermThreads / kTS;  // tile rows a pass of the workgroup covers

struct PermTile {
  int64_t src_off, dst_off;
  int64_t item_begin;
  int64_t sstride[kMaxRank];
  int64_t dstride[kMaxRank];
  int32_t ext[kMaxRank];
  int32_t q;      // fused rank
  int32_t s_dim;  // fastest source mode (q - 1)
  int32_t t_dim;  // source mode that is fastest in the result
  int32_t nbs, nbt;  // blocks along s and t (0 for a whole-plane tile)
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

// `plane` is the workgroup's LDS staging area (a whole plane, or one padded
// 32 x 32 tile at its start).
template <typename T>
__device__ inline void permute_item(const T* __restrict__ src, T* __restrict__ dst,
                                    const PermTile& t, int local, T* plane) {
  const int q = t.q, sd = t.s_dim, td = t.t_dim;

  if (sd == td) {
    // contiguous-run copy: item = (other modes, run block)
    const int nruns = (t.ext[sd] + kCopyRun - 1) / kCopyRun;
    const int rb = local % nruns;
    local /= nruns;
    int64_t so = t.src_off, doff = t.dst_off;
    for (int d = q - 1; d >= 0; --d) {
      if (d == sd) continue;
      const int x = local % t.ext[d];
      local /= t.ext[d];
      so += x * t.sstride[d];
      doff += x * t.dstride[d];
    }
    const int b0 = rb * kCopyRun;
    const int n = min(kCopyRun, t.ext[sd] - b0);
    for (int x = threadIdx.x; x < n; x += blockDim.x) dst[doff + b0 + x] = src[so + b0 + x];
    return;
  }

  if (t.nbs == 0) {
    // whole plane: ns x nt elements, LDS laid out [t][s] with odd stride
    const int ns = t.ext[sd], nt = t.ext[td];
    int64_t so = t.src_off, doff = t.dst_off;
    for (int d = q - 1; d >= 0; --d) {
      if (d == sd || d == td) continue;
      const int x = local % t.ext[d];
      local /= t.ext[d];
      so += x * t.sstride[d];
      doff += x * t.dstride[d];
    }
    T* pl = plane;
    const int ld = ns | 1;
    const int n = ns * nt;
    // element e = tid + i * blockDim of the plane is (e / w, e % w) for a row
    // length w; one division per item, then each step of blockDim elements
    // advances by (blockDim / w, blockDim % w) with one carry, and four
    // elements per thread are loaded before any is stored
    constexpr int U = 4;
    {
      const int qs = blockDim.x / ns, rs = blockDim.x - qs * ns;
      int y = threadIdx.x / ns, x = threadIdx.x - y * ns;  // t index, s index
      const T* __restrict__ sp = src + so;
      const int64_t st = t.sstride[td];
      for (int e0 = threadIdx.x; e0 < n; e0 += U * blockDim.x) {
        T v[U];
        int yy[U], xx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          yy[u] = y;
          xx[u] = x;
          v[u] = e0 + u * (int)blockDim.x < n ? sp[(int64_t)y * st + x] : T(0);
          x += rs;
          y += qs;
          if (x >= ns) {
            x -= ns;
            ++y;
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (e0 + u * (int)blockDim.x < n) pl[yy[u] * ld + xx[u]] = v[u];
      }
    }
    __syncthreads();
    {
      const int qt = blockDim.x / nt, rt = blockDim.x - qt * nt;
      int y = threadIdx.x / nt, x = threadIdx.x - y * nt;  // s index, t index
      T* __restrict__ dp = dst + doff;
      const int64_t st = t.dstride[sd];
      for (int e0 = threadIdx.x; e0 < n; e0 += U * blockDim.x) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (e0 + u * (int)blockDim.x < n) dp[(int64_t)y * st + x] = pl[x * ld + y];
          x += rt;
          y += qt;
          if (x >= nt) {
            x -= nt;
            ++y;
          }
        }
      }
    }
    return;
  }

  const int bs = local % t.nbs;
  local /= t.nbs;
  const int bt = local % t.nbt;
  local /= t.nbt;
  int64_t so = t.src_off, doff = t.dst_off;
  for (int d = q - 1; d >= 0; --d) {
    if (d == sd || d == td) continue;
    const int x = local % t.ext[d];
    local /= t.ext[d];
    so += x * t.sstride[d];
    doff += x * t.dstride[d];
  }
  const int s0 = bs * kTS, t0 = bt * kTS;
  const int ns = min(kTS, t.ext[sd] - s0), nt = min(kTS, t.ext[td] - t0);
  const int tx = threadIdx.x % kTS, ty = threadIdx.x / kTS;  // 32 x kRowsPer
  T (*lds)[kTS + 1] = reinterpret_cast<T (*)[kTS + 1]>(plane);
  // load: rows along t, contiguous along s (source stride 1); all of a
  // thread's loads are issued before the LDS writes
  T v[kTS / kRowsPer];
#pragma unroll
  for (int k = 0; k < kTS / kRowsPer; ++k) {
    const int y = ty + kRowsPer * k;
    v[k] = (tx < ns && y < nt) ? src[so + (int64_t)(t0 + y) * t.sstride[td] + s0 + tx] : T(0);
  }
#pragma unroll
  for (int k = 0; k < kTS / kRowsPer; ++k) lds[ty + kRowsPer * k][tx] = v[k];
  __syncthreads();
  // store: rows along s, contiguous along t (result stride 1); the caller's
  // barrier after the item protects the LDS tile before its next reuse
#pragma unroll
  for (int k = 0; k < kTS / kRowsPer; ++k) {
    const int y = ty + kRowsPer * k;
    if (tx < nt && y < ns) dst[doff + (int64_t)(s0 + y) * t.dstride[sd] + t0 + tx] = lds[tx][y];
  }
}

// Each workgroup takes a contiguous range of items, so consecutive items
// mostly belong to one tile: the tile is found by binary search once and
// then followed forward.
template <typename T>
__global__ __launch_bounds__(kPermThreads) void permute_kernel(
    const T* __restrict__ src, T* __restrict__ dst,
    const PermTile* __restrict__ tiles, const int64_t* __restrict__ prefix,
    int64_t ntiles) {
  static_assert(kPlaneMax + 96 >= kTS * (kTS + 1), "staging area holds a tile");
  __shared__ T plane[kPlaneMax + 96];  // rows padded to an odd length
  const int64_t nitems = prefix[ntiles];
  const int64_t per = (nitems + gridDim.x - 1) / gridDim.x;
  const int64_t ib = min<int64_t>(nitems, (int64_t)blockIdx.x * per);
  const int64_t ie = min<int64_t>(nitems, ib + per);
  if (ib >= ie) return;
  int64_t ti = find_tile(prefix, ntiles, ib);
  int64_t tnext = prefix[ti + 1];
  for (int64_t item = ib; item < ie; ++item) {
    while (item >= tnext) {
      ++ti;
      tnext = prefix[ti + 1];
    }
    permute_item<T>(src, dst, tiles[ti], (int)(item - prefix[ti]), plane);
    __syncthreads();
  }
}

}  // namespace

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
    int64_t nitems = 0;
    if (vol == 0) {
      nitems = 0;
    } else if (t.t_dim == t.s_dim) {
      nitems = (t.ext[t.s_dim] + kCopyRun - 1) / kCopyRun;
      for (int d = 0; d < q; ++d)
        if (d != t.s_dim) nitems *= t.ext[d];
    } else {
      const bool whole = (int64_t)(t.ext[t.s_dim] | 1) * t.ext[t.t_dim] <= kPlaneMax + 96 &&
                         t.ext[t.s_dim] <= 96 && t.ext[t.t_dim] <= 96;
      t.nbs = whole ? 0 : (t.ext[t.s_dim] + kTS - 1) / kTS;
      t.nbt = whole ? 0 : (t.ext[t.t_dim] + kTS - 1) / kTS;
      nitems = whole ? 1 : (int64_t)t.nbs * t.nbt;
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
