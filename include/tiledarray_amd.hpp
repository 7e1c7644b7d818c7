// tiledarray_amd.hpp -- C++ host interface over the C ABI of include/tamx.h.
//
// Mirrors the parts of TiledArray's interface the contraction path sits
// behind, with the same names and meanings on one rank and one GPU:
//   World                  TA::World / TA::initialize (tiledarray.cpp:84-163)
//   TiledRange1/TiledRange tiled_range1.h / tiled_range.h (row-major tiles)
//   Permutation            permutation.h:115-150 ("TO" convention)
//   DistArray<T>           dist_array.h:63, tiles resident in HBM
//   make_sparse            make_array + SparseShape(norms) (make_array.h:152-205,
//                          sparse_shape.h:335-349)
//   A("i,j") * B("j,k")    expressions (dist_array.h:1311-1330, mult_expr.h:192)
//   C("i,k") = expr        Expr::eval_to (expressions/expr.h:377-427) running
//                          ContEngine + Summa + ContractReduce on the GPU
//   Exception              TiledArray::Exception (error.h): every non-OK status
//                          of the C ABI becomes one, as TA_ASSERT would throw.
// Header only; link against libtamx.so and the HIP runtime.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tamx.h"

namespace tamx {

class Exception : public std::runtime_error {
 public:
  Exception(int status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int rc, const char* what) {
  if (rc != TAMX_OK) {
    const char* m = tamx_last_error();
    throw Exception(rc, std::string(what) + ": " + (m && *m ? m : tamx_status_string(rc)));
  }
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Exception(TAMX_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Owning device allocation.
template <typename T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) : n_(n) {
    if (n) hip_check(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  ~DeviceBuffer() {
    if (p_) (void)hipFree(p_);
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  void upload(const T* h, size_t n, size_t off = 0) {
    if (n) hip_check(hipMemcpy(p_ + off, h, n * sizeof(T), hipMemcpyHostToDevice), "upload");
  }
  void download(T* h, size_t n, size_t off = 0) const {
    if (n) hip_check(hipMemcpy(h, p_ + off, n * sizeof(T), hipMemcpyDeviceToHost), "download");
  }
  std::vector<T> to_host() const {
    std::vector<T> v(n_);
    download(v.data(), n_);
    return v;
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

template <typename T>
DeviceBuffer<T> to_device(const std::vector<T>& v) {
  DeviceBuffer<T> d(std::max<size_t>(1, v.size()));
  d.upload(v.data(), v.size());
  return d;
}

// One rank, one GPU.
class World {
 public:
  explicit World(int device = 0) {
    check(tamx_context_create(device, &ctx_), "tamx_context_create");
    hip_check(hipStreamCreate(&stream_), "hipStreamCreate");
  }
  ~World() {
    if (stream_) (void)hipStreamDestroy(stream_);
    if (ctx_) tamx_context_destroy(ctx_);
  }
  World(const World&) = delete;
  World& operator=(const World&) = delete;
  tamx_context* ctx() const { return ctx_; }
  void* stream() const { return static_cast<void*>(stream_); }
  void fence() const { hip_check(hipStreamSynchronize(stream_), "fence"); }

 private:
  tamx_context* ctx_ = nullptr;
  hipStream_t stream_ = nullptr;
};

// A launch sequence captured on the world's stream (tamx_graph_begin/_end)
// and replayed by launch().  Only calls that neither synchronise nor
// allocate may run inside the capture (norms, shape estimate, work list,
// permutes, limb split, grouped GEMMs on buffers that outlive the graph).
class Graph {
 public:
  template <typename F>
  Graph(World& w, F&& fn) : w_(&w) {
    check(tamx_graph_begin(w.ctx(), w.stream()), "tamx_graph_begin");
    try {
      fn();
    } catch (...) {
      tamx_graph* g = nullptr;
      if (tamx_graph_end(w.ctx(), w.stream(), &g) == TAMX_OK) tamx_graph_destroy(g);
      throw;
    }
    check(tamx_graph_end(w.ctx(), w.stream(), &g_), "tamx_graph_end");
  }
  ~Graph() { tamx_graph_destroy(g_); }
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;
  void launch() const { check(tamx_graph_launch(g_, w_->stream()), "tamx_graph_launch"); }

 private:
  World* w_;
  tamx_graph* g_ = nullptr;
};

// Default SparseShape<float> threshold (sparse_shape.h:1941-1942).
inline float default_threshold() { return std::numeric_limits<float>::epsilon(); }

class TiledRange1 {
 public:
  TiledRange1() = default;
  TiledRange1(std::initializer_list<int64_t> b) : b_(b) { validate(); }
  explicit TiledRange1(std::vector<int64_t> b) : b_(std::move(b)) { validate(); }
  size_t ntiles() const { return b_.empty() ? 0 : b_.size() - 1; }
  int64_t extent(size_t t) const { return b_[t + 1] - b_[t]; }
  int64_t lobound(size_t t) const { return b_[t]; }
  int64_t elements() const { return b_.empty() ? 0 : b_.back() - b_.front(); }
  const std::vector<int64_t>& boundaries() const { return b_; }
  bool operator==(const TiledRange1& o) const { return b_ == o.b_; }
  bool operator!=(const TiledRange1& o) const { return !(*this == o); }

 private:
  void validate() const {
    for (size_t i = 1; i < b_.size(); ++i)
      if (b_[i] < b_[i - 1]) throw Exception(TAMX_ERR_BAD_ARG, "TiledRange1 boundaries decrease");
  }
  std::vector<int64_t> b_;
};

class TiledRange {
 public:
  TiledRange() = default;
  TiledRange(std::initializer_list<TiledRange1> r) : r_(r) {}
  explicit TiledRange(std::vector<TiledRange1> r) : r_(std::move(r)) {}
  size_t rank() const { return r_.size(); }
  const TiledRange1& dim(size_t d) const { return r_[d]; }
  size_t ntiles() const {
    size_t n = 1;
    for (auto& r : r_) n *= r.ntiles();
    return n;
  }
  std::vector<size_t> tile_index(size_t ord) const {
    std::vector<size_t> idx(rank());
    for (size_t d = rank(); d-- > 0;) {
      idx[d] = ord % r_[d].ntiles();
      ord /= r_[d].ntiles();
    }
    return idx;
  }
  std::vector<int64_t> tile_extents(size_t ord) const {
    auto idx = tile_index(ord);
    std::vector<int64_t> e(rank());
    for (size_t d = 0; d < rank(); ++d) e[d] = r_[d].extent(idx[d]);
    return e;
  }
  int64_t tile_volume(size_t ord) const {
    auto e = tile_extents(ord);
    return std::accumulate(e.begin(), e.end(), int64_t(1), std::multiplies<int64_t>());
  }
  // tiled range of the TO-permuted tensor: result[perm[d]] = this[d]
  TiledRange permute(const std::vector<int>& perm) const {
    std::vector<TiledRange1> out(rank());
    for (size_t d = 0; d < rank(); ++d) out[perm[d]] = r_[d];
    return TiledRange(out);
  }
  bool operator==(const TiledRange& o) const { return r_ == o.r_; }

 private:
  std::vector<TiledRange1> r_;
};

// Permutation in TA's one-line "TO" form (permutation.h:115-150).
class Permutation {
 public:
  Permutation() = default;
  explicit Permutation(std::vector<int> p) : p_(std::move(p)) {
    std::vector<bool> seen(p_.size(), false);
    for (int x : p_) {
      if (x < 0 || x >= (int)p_.size() || seen[x])
        throw Exception(TAMX_ERR_BAD_ARG, "not a permutation");
      seen[x] = true;
    }
  }
  size_t size() const { return p_.size(); }
  int operator[](size_t i) const { return p_[i]; }
  const std::vector<int>& data() const { return p_; }
  bool is_identity() const {
    for (size_t i = 0; i < p_.size(); ++i)
      if (p_[i] != (int)i) return false;
    return true;
  }
  Permutation inv() const {
    std::vector<int> q(p_.size());
    for (size_t i = 0; i < p_.size(); ++i) q[p_[i]] = (int)i;
    return Permutation(q);
  }

 private:
  std::vector<int> p_;
};

// BF16_LIMBS / INT8_LIMBS: FP64 operands split into nslices BF16 or INT8
// limbs (tamx_split_plan_create_typed), products on BF16 or INT8 MFMA, sums
// and C in FP64
enum class Precision { FP64, FP32, BF16_LIMBS, INT8_LIMBS };

struct ContractOptions {
  float threshold = default_threshold();  // result-shape threshold
  float pair_threshold = 0.0f;            // per-pair filter (off, as in TA)
  Precision precision = Precision::FP64;
  int nslices = 5;                        // limbs for BF16_LIMBS / INT8_LIMBS
  bool autotune = false;                  // pick the tile shape by timing (FP64/FP32)
  // limb precisions: a `general` operand is split straight from its stored tiles
  // through the permutation's index map (no permuted copy) unless set
  bool permute_copy = false;
};

template <typename T>
class DistArray;
template <typename T>
struct TsrExpr;
template <typename T>
struct ContExpr;

namespace detail {
template <typename T>
void contract(World& w, DistArray<T>& result, const std::string& target, const DistArray<T>& A,
              const std::string& la, const DistArray<T>& B, const std::string& lb, double factor,
              const ContractOptions& opt);
template <typename T>
DistArray<T> permute_array(World& w, const DistArray<T>& a, const Permutation& perm,
                           bool copy_data = true, std::vector<size_t>* src_of = nullptr);
}  // namespace detail

// A block-sparse tiled tensor resident in HBM: the non-zero tiles back to
// back in one pool (row-major tiles), offsets[t] < 0 for zero tiles, and the
// SparseShape (scaled float norms on the device) once made sparse.
template <typename T>
class DistArray {
 public:
  using value_type = T;
  DistArray() = default;
  DistArray(World& w, TiledRange tr, const std::vector<bool>& mask = {})
      : world_(&w), trange_(std::move(tr)) {
    offsets_.assign(trange_.ntiles(), -1);
    int64_t pos = 0;
    for (size_t t = 0; t < trange_.ntiles(); ++t)
      if (mask.empty() || mask[t]) {
        offsets_[t] = pos;
        pos += trange_.tile_volume(t);
      }
    pool_ = DeviceBuffer<T>(std::max<int64_t>(1, pos));
    pool_elems_ = pos;
    d_offsets_ = to_device(offsets_);
  }

  World& world() const { return *world_; }
  const TiledRange& trange() const { return trange_; }
  const std::vector<int64_t>& offsets() const { return offsets_; }
  T* pool() const { return pool_.get(); }
  int64_t pool_elems() const { return pool_elems_; }
  const int64_t* device_offsets() const { return d_offsets_.get(); }
  bool is_dense() const { return !shape_norms_.get(); }
  const float* shape_norms() const { return shape_norms_.get(); }
  float shape_threshold() const { return threshold_; }
  // a tile is zero when absent or screened by the shape (sparse_shape.h:495-499)
  bool is_zero(size_t t) const {
    if (offsets_[t] < 0) return true;
    if (is_dense()) return false;
    return host_norms()[t] < threshold_;
  }
  std::vector<float> host_norms() const {
    std::vector<float> h(trange_.ntiles(), 0.0f);
    if (!is_dense()) shape_norms_.download(h.data(), h.size());
    return h;
  }

  // tile t as a host vector (row-major); zeros for a zero tile
  std::vector<T> find(size_t t) const {
    std::vector<T> h(trange_.tile_volume(t), T(0));
    if (offsets_[t] >= 0) pool_.download(h.data(), h.size(), offsets_[t]);
    return h;
  }
  void set(size_t t, const std::vector<T>& data) {
    if (offsets_[t] < 0) throw Exception(TAMX_ERR_BAD_ARG, "set: tile is not allocated");
    if ((int64_t)data.size() != trange_.tile_volume(t))
      throw Exception(TAMX_ERR_BAD_ARG, "set: tile size mismatch");
    pool_.upload(data.data(), data.size(), offsets_[t]);
  }
  // init_tiles-like fill: op(t, extents) returns the tile's data
  template <typename Op>
  void init_tiles(Op&& op) {
    for (size_t t = 0; t < trange_.ntiles(); ++t)
      if (offsets_[t] >= 0) set(t, op(t, trange_.tile_extents(t)));
  }

  // SparseShape from tile norms, inverse-volume scaled and screened
  // (make_array semantics)
  void make_sparse(float threshold = default_threshold()) {
    std::vector<int64_t> vols(trange_.ntiles());
    for (size_t t = 0; t < vols.size(); ++t) vols[t] = trange_.tile_volume(t);
    tamx_norm_plan* plan = nullptr;
    check(tamx_norm_plan_create(world_->ctx(), offsets_.data(), vols.data(), (int64_t)vols.size(),
                                &plan),
          "tamx_norm_plan_create");
    std::unique_ptr<tamx_norm_plan, int (*)(tamx_norm_plan*)> guard(plan, tamx_norm_plan_destroy);
    shape_norms_ = DeviceBuffer<float>(std::max<size_t>(1, trange_.ntiles()));
    if (std::is_same<T, double>::value)
      check(tamx_tile_norms_f64(world_->ctx(), plan, reinterpret_cast<const double*>(pool()),
                                shape_norms_.get(), world_->stream()),
            "tamx_tile_norms_f64");
    else
      check(tamx_tile_norms_f32(world_->ctx(), plan, reinterpret_cast<const float*>(pool()),
                                shape_norms_.get(), world_->stream()),
            "tamx_tile_norms_f32");
    std::vector<int64_t> nt(trange_.rank());
    std::vector<float> sizes;
    for (size_t d = 0; d < trange_.rank(); ++d) {
      nt[d] = (int64_t)trange_.dim(d).ntiles();
      for (size_t i = 0; i < trange_.dim(d).ntiles(); ++i)
        sizes.push_back((float)trange_.dim(d).extent(i));
    }
    auto d_sizes = to_device(sizes);
    check(tamx_shape_scale(world_->ctx(), shape_norms_.get(), (int)trange_.rank(), nt.data(), 1,
                           d_sizes.get(), 0, threshold, nullptr, world_->stream()),
          "tamx_shape_scale");
    world_->fence();
    threshold_ = threshold;
  }

  TsrExpr<T> operator()(const std::string& annotation) { return TsrExpr<T>{*this, annotation}; }

 private:
  template <typename U>
  friend void detail::contract(World&, DistArray<U>&, const std::string&, const DistArray<U>&,
                               const std::string&, const DistArray<U>&, const std::string&,
                               double, const ContractOptions&);
  template <typename U>
  friend DistArray<U> detail::permute_array(World&, const DistArray<U>&, const Permutation&,
                                            bool, std::vector<size_t>*);

  World* world_ = nullptr;
  TiledRange trange_;
  DeviceBuffer<T> pool_;
  int64_t pool_elems_ = 0;
  std::vector<int64_t> offsets_;
  DeviceBuffer<int64_t> d_offsets_;
  DeviceBuffer<float> shape_norms_;
  float threshold_ = 0.0f;
};

template <typename T>
struct ContExpr {
  DistArray<T>* left;
  std::string left_annot;
  DistArray<T>* right;
  std::string right_annot;
  double factor = 1.0;
  ContractOptions options{};
  ContExpr& with(const ContractOptions& o) {
    options = o;
    return *this;
  }
};

template <typename T>
struct TsrExpr {
  DistArray<T>& array;
  std::string annotation;
  // C("i,k") = A("i,j") * B("j,k") (Expr::eval_to, expressions/expr.h:377-427)
  TsrExpr& operator=(const ContExpr<T>& e) {
    detail::contract(array.world(), array, annotation, *e.left, e.left_annot, *e.right,
                     e.right_annot, e.factor, e.options);
    return *this;
  }
};

template <typename T>
ContExpr<T> operator*(TsrExpr<T> l, TsrExpr<T> r) {
  return ContExpr<T>{&l.array, l.annotation, &r.array, r.annotation, 1.0, {}};
}
template <typename T>
ContExpr<T> operator*(double f, ContExpr<T> e) {
  e.factor *= f;
  return e;
}

namespace detail {

inline std::vector<int64_t> fused_extents(const std::vector<const TiledRange1*>& ranges) {
  std::vector<int64_t> e{1};
  for (auto* r : ranges) {
    std::vector<int64_t> n;
    n.reserve(e.size() * r->ntiles());
    for (int64_t x : e)
      for (size_t i = 0; i < r->ntiles(); ++i) n.push_back(x * r->extent(i));
    e.swap(n);
  }
  return e;
}

// recursive_outer_product of the contracted tile extents (float products,
// sparse_shape.h:106-131 with the identity op)
inline std::vector<float> k_sizes(const std::vector<const TiledRange1*>& ranges, size_t lo,
                                  size_t hi) {
  if (hi - lo == 1) {
    std::vector<float> v(ranges[lo]->ntiles());
    for (size_t i = 0; i < v.size(); ++i) v[i] = (float)ranges[lo]->extent(i);
    return v;
  }
  const size_t n = hi - lo, mid = (n >> 1) + (n & 1);
  auto l = k_sizes(ranges, lo, lo + mid), r = k_sizes(ranges, lo + mid, hi);
  std::vector<float> out;
  out.reserve(l.size() * r.size());
  for (float a : l)
    for (float b : r) out.push_back(a * b);
  return out;
}

// TO-permuted copy: tile data through tamx_permute, shape norms reordered.
// Without copy_data only the structure is permuted (tiled range, tile
// offsets into a's pool, shape) and the result has no pool of its own;
// src_of receives the source ordinal of every permuted tile.
template <typename T>
DistArray<T> permute_array(World& w, const DistArray<T>& a, const Permutation& perm,
                           bool copy_data, std::vector<size_t>* src_of) {
  const TiledRange& tr = a.trange();
  DistArray<T> out;
  out.world_ = &w;
  out.trange_ = tr.permute(perm.data());
  const size_t nt = tr.ntiles(), rank = tr.rank();
  out.offsets_.assign(nt, -1);
  if (src_of) src_of->assign(nt, 0);
  std::vector<int64_t> ext, src, dst;
  for (size_t t = 0; t < nt; ++t) {
    auto idx = tr.tile_index(t);
    std::vector<size_t> pidx(rank);
    for (size_t d = 0; d < rank; ++d) pidx[perm[d]] = idx[d];
    size_t o = 0;
    for (size_t d = 0; d < rank; ++d) o = o * out.trange_.dim(d).ntiles() + pidx[d];
    out.offsets_[o] = a.offsets_[t];  // a tile keeps its place in the pool
    if (src_of) (*src_of)[o] = t;
    if (a.offsets_[t] >= 0) {
      auto e = tr.tile_extents(t);
      ext.insert(ext.end(), e.begin(), e.end());
      src.push_back(a.offsets_[t]);
      dst.push_back(a.offsets_[t]);
    }
  }
  out.pool_elems_ = a.pool_elems_;
  out.d_offsets_ = to_device(out.offsets_);
  if (copy_data) out.pool_ = DeviceBuffer<T>(std::max<int64_t>(1, a.pool_elems_));
  tamx_permute_plan* plan = nullptr;
  check(tamx_permute_plan_create(w.ctx(), (int)rank, perm.data().data(), (int64_t)src.size(),
                                 ext.data(), src.data(), dst.data(), &plan),
        "tamx_permute_plan_create");
  std::unique_ptr<tamx_permute_plan, int (*)(tamx_permute_plan*)> guard(plan,
                                                                       tamx_permute_plan_destroy);
  if (copy_data && !src.empty()) {
    if (std::is_same<T, double>::value)
      check(tamx_permute_f64(w.ctx(), plan, reinterpret_cast<const double*>(a.pool()),
                             reinterpret_cast<double*>(out.pool()), w.stream()),
            "tamx_permute_f64");
    else
      check(tamx_permute_f32(w.ctx(), plan, reinterpret_cast<const float*>(a.pool()),
                             reinterpret_cast<float*>(out.pool()), w.stream()),
            "tamx_permute_f32");
  }
  if (!a.is_dense()) {
    auto hn = a.host_norms();
    std::vector<float> pn(nt);
    for (size_t t = 0; t < nt; ++t) {
      auto idx = tr.tile_index(t);
      std::vector<size_t> pidx(rank);
      for (size_t d = 0; d < rank; ++d) pidx[perm[d]] = idx[d];
      size_t o = 0;
      for (size_t d = 0; d < rank; ++d) o = o * out.trange_.dim(d).ntiles() + pidx[d];
      pn[o] = hn[t];
    }
    out.shape_norms_ = to_device(pn);
    out.threshold_ = a.threshold_;
  }
  w.fence();
  return out;
}

// One contraction on one rank: the planning of ContEngine
// (expressions/cont_engine.h:354-680), the schedule of Summa
// (dist_eval/contraction_eval.h:1094-1384) and the tile op ContractReduce
// (tile_op/contract_reduce.h:409-453), all tile work on the GPU.
template <typename T>
void contract(World& w, DistArray<T>& result, const std::string& target, const DistArray<T>& A0,
              const std::string& la, const DistArray<T>& B0, const std::string& lb, double factor,
              const ContractOptions& opt) {
  tamx_contraction_plan plan;
  check(tamx_plan_contraction(la.c_str(), lb.c_str(), target.c_str(), 1, &plan),
        "tamx_plan_contraction");
  if ((int)A0.trange().rank() != plan.left_rank || (int)B0.trange().rank() != plan.right_rank)
    throw Exception(TAMX_ERR_BAD_ARG, "annotation rank does not match the array rank");
  if (A0.is_dense() != B0.is_dense())
    throw Exception(TAMX_ERR_BAD_ARG, "operands must both be dense or both be sparse");
  auto perm_of = [](const int32_t* p, int n) { return Permutation(std::vector<int>(p, p + n)); };
  // operands whose layout is `general` are permuted (permute-on-read), or,
  // for the limb split, only their structure is, and the split reads the
  // stored tiles through the permutation's index map
  const bool limbs = opt.precision == Precision::BF16_LIMBS ||
                     opt.precision == Precision::INT8_LIMBS;
  const int limb_type = opt.precision == Precision::INT8_LIMBS ? TAMX_LIMB_I8 : TAMX_LIMB_BF16;
  const bool map = limbs && !opt.permute_copy;
  DistArray<T> Ap, Bp;
  const DistArray<T>* A = &A0;
  const DistArray<T>* B = &B0;
  std::vector<size_t> a_src, b_src;
  if (plan.left_permtype == 3) {
    Ap = permute_array(w, A0, perm_of(plan.left_perm, plan.left_rank), !map, &a_src);
    A = &Ap;
  }
  if (plan.right_permtype == 3) {
    Bp = permute_array(w, B0, perm_of(plan.right_perm, plan.right_rank), !map, &b_src);
    B = &Bp;
  }
  // element data of the operands in the layout the GEMM or split reads
  const T* a_data = (map && A != &A0) ? A0.pool() : A->pool();
  const T* b_data = (map && B != &B0) ? B0.pool() : B->pool();
  // fused structure (GemmHelper, math/gemm_helper.h:41-278)
  std::vector<const TiledRange1*> lo, li, ro, ri;
  for (int d = 0; d < plan.left_rank; ++d)
    (plan.left_gemm_is_inner[d] ? li : lo).push_back(&A->trange().dim(d));
  for (int d = 0; d < plan.right_rank; ++d)
    (plan.right_gemm_is_inner[d] ? ri : ro).push_back(&B->trange().dim(d));
  bool congruent = li.size() == ri.size();
  for (size_t d = 0; congruent && d < li.size(); ++d) congruent = *li[d] == *ri[d];
  if (!congruent)
    throw Exception(TAMX_ERR_BAD_ARG, "contracted modes have incongruent tilings");
  const auto mext = fused_extents(lo), next = fused_extents(ro), kext = fused_extents(li);
  const int64_t Mt = (int64_t)mext.size(), Nt = (int64_t)next.size(), Kt = (int64_t)kext.size();
  std::vector<int32_t> m32(mext.begin(), mext.end()), n32(next.begin(), next.end()),
      k32(kext.begin(), kext.end());
  auto d_m = to_device(m32), d_n = to_device(n32), d_k = to_device(k32);
  // result shape estimate (SparseShape::gemm) or dense
  const bool sparse = !A->is_dense();
  DeviceBuffer<float> c_norms(std::max<int64_t>(1, Mt * Nt));
  if (sparse) {
    auto ks = to_device(k_sizes(li, 0, li.size()));
    check(tamx_shape_gemm_f32(w.ctx(), A->shape_norms(), B->shape_norms(), plan.op_a, plan.op_b,
                              Mt, Nt, Kt, 1, ks.get(), (float)std::fabs(factor), opt.threshold,
                              c_norms.get(), nullptr, w.stream()),
          "tamx_shape_gemm_f32");
    w.fence();
  } else {
    c_norms.upload(std::vector<float>(Mt * Nt, 1.0f).data(), Mt * Nt);
  }
  // result tiles packed in row-major tile order
  int64_t cap = 0;
  for (int64_t i = 0; i < Mt; ++i) cap += mext[i];
  int64_t capn = 0;
  for (int64_t j = 0; j < Nt; ++j) capn += next[j];
  cap *= capn;
  DeviceBuffer<T> C(std::max<int64_t>(1, cap));
  DeviceBuffer<int64_t> c_off(std::max<int64_t>(1, Mt * Nt));
  tamx_worklist* wl = nullptr;
  check(tamx_worklist_create(w.ctx(), Mt * Nt, std::max<int64_t>(1, Mt * Nt * Kt), Mt * Nt, &wl),
        "tamx_worklist_create");
  std::unique_ptr<tamx_worklist, int (*)(tamx_worklist*)> wl_guard(wl, tamx_worklist_destroy);
  tamx_contraction_desc d{};
  d.op_a = plan.op_a;
  d.op_b = plan.op_b;
  d.Mt = Mt;
  d.Nt = Nt;
  d.Kt = Kt;
  d.nbatch = 1;
  d.mext = d_m.get();
  d.next = d_n.get();
  d.kext = d_k.get();
  d.a_off = A->device_offsets();
  d.b_off = B->device_offsets();
  d.a_norms = sparse ? A->shape_norms() : nullptr;
  d.b_norms = sparse ? B->shape_norms() : nullptr;
  d.c_norms = c_norms.get();
  d.a_threshold = sparse ? A->shape_threshold() : 0.0f;
  d.b_threshold = sparse ? B->shape_threshold() : 0.0f;
  d.c_threshold = sparse ? opt.threshold : 0.0f;
  d.pair_threshold = opt.pair_threshold;
  d.store_trans = plan.result_perm_matrix_transpose;
  d.c_off = c_off.get();
  d.c_capacity = cap;
  check(tamx_contraction_worklist(w.ctx(), &d, wl, w.stream()), "tamx_contraction_worklist");
  std::unique_ptr<tamx_split_plan, int (*)(tamx_split_plan*)> sa(nullptr, tamx_split_plan_destroy),
      sb(nullptr, tamx_split_plan_destroy);
  DeviceBuffer<uint8_t> la_limbs, lb_limbs;
  if (opt.precision == Precision::FP64 || opt.precision == Precision::FP32) {
    const bool f64 = std::is_same<T, double>::value;
    tamx_gemm_config cfg{};
    check(tamx_gemm_config_get(0, &cfg), "tamx_gemm_config_get");
    if (opt.autotune)
      check(tamx_autotune_grouped_gemm(w.ctx(), wl, f64 ? TAMX_PREC_F64 : TAMX_PREC_F32,
                                       plan.op_a, plan.op_b, a_data, b_data, C.get(),
                                       3, 0, &cfg, nullptr, w.stream()),
            "tamx_autotune_grouped_gemm");
    if (f64)
      check(tamx_grouped_gemm_f64_config(w.ctx(), wl, plan.op_a, plan.op_b, factor,
                                         reinterpret_cast<const double*>(a_data),
                                         reinterpret_cast<const double*>(b_data),
                                         reinterpret_cast<double*>(C.get()), &cfg, w.stream()),
            "tamx_grouped_gemm_f64_config");
    else
      check(tamx_grouped_gemm_f32_config(w.ctx(), wl, plan.op_a, plan.op_b, (float)factor,
                                         reinterpret_cast<const float*>(a_data),
                                         reinterpret_cast<const float*>(b_data),
                                         reinterpret_cast<float*>(C.get()), &cfg, w.stream()),
            "tamx_grouped_gemm_f32_config");
  } else {
    if (!std::is_same<T, double>::value)
      throw Exception(TAMX_ERR_UNSUPPORTED, "split precision takes FP64 operands");
    // tile tables of the fused operands in the ordinals the pairs carry
    auto table = [&](bool right, int op, std::vector<int32_t>& rows, std::vector<int32_t>& cols) {
      const auto& outer = right ? next : mext;
      const int64_t go = right ? Nt : Mt;
      const bool rows_first = right ? op == 1 : op == 0;
      for (int64_t x = 0; x < go * Kt; ++x) {
        const int64_t a = rows_first ? x / Kt : x / go, b = rows_first ? x % Kt : x % go;
        rows.push_back((int32_t)(rows_first ? outer[a] : kext[a]));
        cols.push_back((int32_t)(rows_first ? kext[b] : outer[b]));
      }
    };
    // the index map of a permuted operand: per tile of the permuted grid,
    // extents and stored strides of its outer modes (limb rows) and its
    // contracted modes (limb k), in GEMM order
    auto general_plan = [&](const DistArray<T>& src, const DistArray<T>& view,
                            const std::vector<size_t>& src_of, const int32_t* perm,
                            const int32_t* is_inner, int rank) {
      std::vector<int> inv(rank);
      for (int d = 0; d < rank; ++d) inv[perm[d]] = d;
      std::vector<int> rows_m, k_m;
      for (int d = 0; d < rank; ++d) (is_inner[d] ? k_m : rows_m).push_back(d);
      if (rows_m.empty() || k_m.empty() || rows_m.size() > 4 || k_m.size() > 4)
        throw Exception(TAMX_ERR_UNSUPPORTED, "index maps cover 1 to 4 outer and inner modes");
      const size_t nt = view.trange().ntiles();
      std::vector<int32_t> rext, kext;
      std::vector<int64_t> rstr, kstr;
      for (size_t t = 0; t < nt; ++t) {
        const auto e = src.trange().tile_extents(src_of[t]);
        std::vector<int64_t> st(rank, 1);
        for (int d = rank - 2; d >= 0; --d) st[d] = st[d + 1] * (int64_t)e[d + 1];
        for (int dp : rows_m) {
          rext.push_back((int32_t)e[inv[dp]]);
          rstr.push_back(st[inv[dp]]);
        }
        for (int dp : k_m) {
          kext.push_back((int32_t)e[inv[dp]]);
          kstr.push_back(st[inv[dp]]);
        }
      }
      tamx_split_plan* sp = nullptr;
      check(tamx_split_plan_create_general_typed(w.ctx(), limb_type, opt.nslices, (int64_t)nt,
                                                 view.offsets().data(), (int)rows_m.size(),
                                                 (int)k_m.size(), rext.data(), rstr.data(),
                                                 kext.data(), kstr.data(), &sp),
            "tamx_split_plan_create_general_typed");
      return sp;
    };
    std::vector<int32_t> ar, ac, br, bc;
    table(false, plan.op_a, ar, ac);
    table(true, plan.op_b, br, bc);
    tamx_split_plan *pa = nullptr, *pb = nullptr;
    if (map && A != &A0)
      pa = general_plan(A0, *A, a_src, plan.left_perm, plan.left_gemm_is_inner,
                        plan.left_rank);
    else
      check(tamx_split_plan_create_typed(w.ctx(), limb_type, opt.nslices, 0, plan.op_a,
                                         (int64_t)ar.size(), A->offsets().data(), ar.data(),
                                         ac.data(), &pa),
            "tamx_split_plan_create_typed");
    sa.reset(pa);
    if (map && B != &B0)
      pb = general_plan(B0, *B, b_src, plan.right_perm, plan.right_gemm_is_inner,
                        plan.right_rank);
    else
      check(tamx_split_plan_create_typed(w.ctx(), limb_type, opt.nslices, 1, plan.op_b,
                                         (int64_t)br.size(), B->offsets().data(), br.data(),
                                         bc.data(), &pb),
            "tamx_split_plan_create_typed");
    sb.reset(pb);
    const int64_t lbytes = limb_type == TAMX_LIMB_I8 ? 1 : 2;  // bytes per limb
    la_limbs = DeviceBuffer<uint8_t>(std::max<int64_t>(2, lbytes * tamx_split_plan_limb_elems(pa)));
    lb_limbs = DeviceBuffer<uint8_t>(std::max<int64_t>(2, lbytes * tamx_split_plan_limb_elems(pb)));
    check(tamx_split_f64(w.ctx(), pa, reinterpret_cast<const double*>(a_data), la_limbs.get(),
                         w.stream()),
          "tamx_split_f64");
    check(tamx_split_f64(w.ctx(), pb, reinterpret_cast<const double*>(b_data), lb_limbs.get(),
                         w.stream()),
          "tamx_split_f64");
    check(tamx_grouped_gemm_split(w.ctx(), wl, factor, pa, la_limbs.get(), pb, lb_limbs.get(),
                                  reinterpret_cast<double*>(C.get()), w.stream()),
          "tamx_grouped_gemm_split");
  }
  int64_t cnt[TAMX_WL_NCOUNTERS];
  check(tamx_worklist_read_counters(wl, cnt, w.stream()), "tamx_worklist_read_counters");
  if (cnt[TAMX_WL_STATUS] != 0)
    throw Exception(TAMX_ERR_CAPACITY, "contraction work list exceeded its capacity");
  // assemble the result in the target order
  std::vector<int64_t> off(Mt * Nt);
  c_off.download(off.data(), off.size());
  std::vector<TiledRange1> nat;
  for (auto* r : lo) nat.push_back(*r);
  for (auto* r : ro) nat.push_back(*r);
  TiledRange ntr(nat);
  Permutation rp = perm_of(plan.result_perm, plan.result_rank);
  DistArray<T> R;
  R.world_ = &w;
  R.trange_ = plan.result_perm_identity ? ntr : ntr.permute(rp.data());
  R.offsets_.assign(Mt * Nt, -1);
  std::vector<float> hn;
  if (sparse) {
    hn.resize(Mt * Nt);
    c_norms.download(hn.data(), hn.size());
  }
  std::vector<float> rn(sparse ? Mt * Nt : 0);
  std::vector<int64_t> ext, src;
  for (int64_t t = 0; t < Mt * Nt; ++t) {
    auto idx = ntr.tile_index(t);
    std::vector<size_t> pidx(idx.size());
    for (size_t q = 0; q < idx.size(); ++q) pidx[rp[q]] = idx[q];
    size_t o = 0;
    for (size_t q = 0; q < idx.size(); ++q) o = o * R.trange_.dim(q).ntiles() + pidx[q];
    R.offsets_[o] = off[t];
    if (sparse) rn[o] = hn[t];
    if (off[t] >= 0 && !plan.result_perm_identity && !plan.result_perm_matrix_transpose) {
      auto e = ntr.tile_extents(t);
      ext.insert(ext.end(), e.begin(), e.end());
      src.push_back(off[t]);
    }
  }
  if (!src.empty()) {  // general result permutation (contract_reduce.h:370-378)
    DeviceBuffer<T> P(C.size());
    tamx_permute_plan* pp = nullptr;
    check(tamx_permute_plan_create(w.ctx(), (int)rp.size(), rp.data().data(), (int64_t)src.size(),
                                   ext.data(), src.data(), src.data(), &pp),
          "tamx_permute_plan_create");
    std::unique_ptr<tamx_permute_plan, int (*)(tamx_permute_plan*)> g(pp,
                                                                     tamx_permute_plan_destroy);
    if (std::is_same<T, double>::value)
      check(tamx_permute_f64(w.ctx(), pp, reinterpret_cast<const double*>(C.get()),
                             reinterpret_cast<double*>(P.get()), w.stream()),
            "tamx_permute_f64");
    else
      check(tamx_permute_f32(w.ctx(), pp, reinterpret_cast<const float*>(C.get()),
                             reinterpret_cast<float*>(P.get()), w.stream()),
            "tamx_permute_f32");
    w.fence();
    C = std::move(P);
  }
  R.pool_ = std::move(C);
  R.pool_elems_ = cap;
  R.d_offsets_ = to_device(R.offsets_);
  if (sparse) {
    R.shape_norms_ = to_device(rn);
    R.threshold_ = opt.threshold;
  }
  w.fence();
  result = std::move(R);
}

}  // namespace detail

}  // namespace tamx
