// C++ parity tests of the host interface (include/tiledarray_amd.hpp), in the
// style of TiledArray's own tests: integer-valued tiles (rand() % 27) so every
// product is exact and results are compared for equality against a host
// reference (tests/tile_op_contract_reduce.cpp:109-222,
// tests/dist_eval_contraction_eval.cpp:293-472, tests/expressions_impl.h
// cont* cases); random FP64 data against a relative Frobenius bound.
// Prints one line per case and "ALL PASSED" at the end; exit status 0 on
// success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "tiledarray_amd.hpp"

using namespace tamx;

static int g_failures = 0;
#define CHECK(cond, ...)                                      \
  do {                                                        \
    if (!(cond)) {                                            \
      std::printf("  FAILED %s:%d: ", __FILE__, __LINE__);    \
      std::printf(__VA_ARGS__);                               \
      std::printf("\n");                                      \
      ++g_failures;                                           \
    }                                                         \
  } while (0)

// dense host copy of a tiled array (row-major over the element ranges)
template <typename T>
std::vector<double> dense(const DistArray<T>& a) {
  const TiledRange& tr = a.trange();
  const size_t r = tr.rank();
  std::vector<int64_t> ext(r);
  for (size_t d = 0; d < r; ++d) ext[d] = tr.dim(d).elements();
  int64_t vol = 1;
  for (auto e : ext) vol *= e;
  std::vector<double> out(vol, 0.0);
  for (size_t t = 0; t < tr.ntiles(); ++t) {
    if (a.offsets()[t] < 0) continue;
    auto tile = a.find(t);
    auto idx = tr.tile_index(t);
    auto te = tr.tile_extents(t);
    std::vector<int64_t> lo(r), e(r, 0);
    for (size_t d = 0; d < r; ++d) lo[d] = tr.dim(d).lobound(idx[d]);
    for (size_t x = 0; x < tile.size(); ++x) {
      int64_t o = 0;
      for (size_t d = 0; d < r; ++d) o = o * ext[d] + lo[d] + e[d];
      out[o] = (double)tile[x];
      for (size_t d = r; d-- > 0;) {
        if (++e[d] < te[d]) break;
        e[d] = 0;
      }
    }
  }
  return out;
}

template <typename T>
DistArray<T> make_array(World& w, const TiledRange& tr, std::mt19937& rng, bool integer,
                        double fill = 1.0) {
  std::vector<bool> mask(tr.ntiles(), true);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (size_t t = 0; t < mask.size(); ++t) mask[t] = u(rng) < fill;
  DistArray<T> a(w, tr, mask);
  a.init_tiles([&](size_t, const std::vector<int64_t>& e) {
    int64_t n = 1;
    for (auto x : e) n *= x;
    std::vector<T> v(n);
    for (auto& x : v) x = integer ? (T)(rng() % 27) : (T)(2.0 * u(rng) - 1.0);
    return v;
  });
  return a;
}

// generic einsum on dense host tensors: result[target] = sum A[la] B[lb]
std::vector<double> einsum(const std::vector<double>& A, const std::string& la,
                           const std::vector<int64_t>& ea, const std::vector<double>& B,
                           const std::string& lb, const std::vector<int64_t>& eb,
                           const std::string& target, std::vector<int64_t>& er) {
  std::map<char, int64_t> ext;
  std::string idx;
  auto add = [&](const std::string& l, const std::vector<int64_t>& e) {
    size_t d = 0;
    for (char c : l)
      if (c != ',') {
        ext[c] = e[d++];
        if (idx.find(c) == std::string::npos) idx.push_back(c);
      }
  };
  add(la, ea);
  add(lb, eb);
  std::string lA, lB, lT;
  for (char c : la) if (c != ',') lA.push_back(c);
  for (char c : lb) if (c != ',') lB.push_back(c);
  for (char c : target) if (c != ',') lT.push_back(c);
  er.clear();
  int64_t rv = 1;
  for (char c : lT) er.push_back(ext[c]), rv *= ext[c];
  std::vector<double> R(rv, 0.0);
  std::map<char, int64_t> val;
  for (char c : idx) val[c] = 0;
  auto ord = [&](const std::string& l) {
    int64_t o = 0;
    for (char c : l) o = o * ext[c] + val[c];
    return o;
  };
  for (;;) {
    R[ord(lT)] += A[ord(lA)] * B[ord(lB)];
    size_t d = idx.size();
    for (; d-- > 0;) {
      if (++val[idx[d]] < ext[idx[d]]) break;
      val[idx[d]] = 0;
    }
    if (d == (size_t)-1) break;
  }
  return R;
}

template <typename T>
std::vector<int64_t> extents(const DistArray<T>& a) {
  std::vector<int64_t> e;
  for (size_t d = 0; d < a.trange().rank(); ++d) e.push_back(a.trange().dim(d).elements());
  return e;
}

double rel_fro(const std::vector<double>& x, const std::vector<double>& r) {
  double n = 0, d = 0;
  for (size_t i = 0; i < x.size(); ++i) n += (x[i] - r[i]) * (x[i] - r[i]), d += r[i] * r[i];
  return d > 0 ? std::sqrt(n / d) : std::sqrt(n);
}

static void run(const char* name, const std::function<void()>& f) {
  const int before = g_failures;
  try {
    f();
  } catch (const std::exception& e) {
    std::printf("  exception: %s\n", e.what());
    ++g_failures;
  }
  std::printf("%s %s\n", g_failures == before ? "ok  " : "FAIL", name);
  std::fflush(stdout);
}

int main() {
  World world(0);
  std::mt19937 rng(42);
  TiledRange1 ti{0, 3, 8, 10, 17}, tj{0, 5, 6, 13}, tk{0, 2, 9, 12, 20};

  // matrix_multiply in all four Trans combinations, factor 3, accumulate
  // over k tiles (tile_op_contract_reduce.cpp:109-222)
  struct Case { const char *la, *lb; };
  for (Case c : {Case{"i,k", "k,j"}, Case{"k,i", "k,j"}, Case{"i,k", "j,k"}, Case{"k,i", "j,k"}}) {
    run((std::string("matrix_multiply ") + c.la + " x " + c.lb).c_str(), [&] {
      auto trA = std::string(c.la) == "i,k" ? TiledRange{ti, tk} : TiledRange{tk, ti};
      auto trB = std::string(c.lb) == "k,j" ? TiledRange{tk, tj} : TiledRange{tj, tk};
      auto A = make_array<double>(world, trA, rng, true, 0.8);
      auto B = make_array<double>(world, trB, rng, true, 0.8);
      DistArray<double> R(world, TiledRange{ti, tj});
      R("i,j") = 3.0 * (A(c.la) * B(c.lb));
      std::vector<int64_t> er;
      auto ref = einsum(dense(A), c.la, extents(A), dense(B), c.lb, extents(B), "i,j", er);
      for (auto& x : ref) x *= 3.0;
      auto got = dense(R);
      CHECK(got == ref, "result differs from the host product");
    });
  }

  run("sparse_eval block-sparse with screened shapes", [&] {
    auto A = make_array<double>(world, TiledRange{ti, tk}, rng, true, 0.6);
    auto B = make_array<double>(world, TiledRange{tk, tj}, rng, true, 0.6);
    A.make_sparse(1e-3f);
    B.make_sparse(1e-3f);
    DistArray<double> C(world, TiledRange{ti, tj});
    C("i,j") = A("i,k") * B("k,j");
    std::vector<int64_t> er;
    auto ref = einsum(dense(A), "i,k", extents(A), dense(B), "k,j", extents(B), "i,j", er);
    auto got = dense(C);
    CHECK(got == ref, "sparse result differs");
    // a result tile is present exactly when some k pair survives
    for (size_t t = 0; t < C.trange().ntiles(); ++t) {
      auto idx = C.trange().tile_index(t);
      bool any = false;
      for (size_t k = 0; k < tk.ntiles(); ++k)
        any = any || (!A.is_zero(idx[0] * tk.ntiles() + k) && !B.is_zero(k * tj.ntiles() + idx[1]));
      CHECK(any == !C.is_zero(t), "pattern of tile %zu", t);
    }
  });

  TiledRange1 tv{0, 4, 9, 11}, to{0, 3, 6};
  run("cont_permute ladder R(a,b,i,j) = T(c,d,i,j) * V(a,c,b,d)", [&] {
    auto T = make_array<double>(world, TiledRange{tv, tv, to, to}, rng, false);
    auto V = make_array<double>(world, TiledRange{tv, tv, tv, tv}, rng, false);
    DistArray<double> R(world, TiledRange{tv, tv, to, to});
    R("a,b,i,j") = T("c,d,i,j") * V("a,c,b,d");
    std::vector<int64_t> er;
    auto ref = einsum(dense(T), "c,d,i,j", extents(T), dense(V), "a,c,b,d", extents(V),
                      "a,b,i,j", er);
    const double err = rel_fro(dense(R), ref);
    CHECK(err <= 1e-13, "relative error %g", err);
  });

  run("cont_permute_permute general result order", [&] {
    auto T = make_array<double>(world, TiledRange{tv, tv, to, to}, rng, true);
    auto V = make_array<double>(world, TiledRange{tv, tv, tv, tv}, rng, true);
    DistArray<double> R(world, TiledRange{to, tv, tv, to});
    R("i,b,a,j") = T("c,d,i,j") * V("a,b,c,d");
    std::vector<int64_t> er;
    auto ref = einsum(dense(T), "c,d,i,j", extents(T), dense(V), "a,b,c,d", extents(V),
                      "i,b,a,j", er);
    CHECK(dense(R) == ref, "permuted result differs");
  });

  run("split precision, 5 BF16 limbs, ladder", [&] {
    auto T = make_array<double>(world, TiledRange{tv, tv, to, to}, rng, false);
    auto V = make_array<double>(world, TiledRange{tv, tv, tv, tv}, rng, false);
    DistArray<double> R(world, TiledRange{tv, tv, to, to});
    ContractOptions opt;
    opt.precision = Precision::BF16_LIMBS;
    opt.nslices = 5;
    R("a,b,i,j") = (T("c,d,i,j") * V("a,c,b,d")).with(opt);
    std::vector<int64_t> er;
    auto ref = einsum(dense(T), "c,d,i,j", extents(T), dense(V), "a,c,b,d", extents(V),
                      "a,b,i,j", er);
    const double err = rel_fro(dense(R), ref);
    CHECK(err <= 1e-10, "relative error %g", err);
  });

  run("split precision, general V read through the index map = permuted copy", [&] {
    auto T = make_array<double>(world, TiledRange{tv, tv, to, to}, rng, false);
    auto V = make_array<double>(world, TiledRange{tv, tv, tv, tv}, rng, false);
    DistArray<double> Rm(world, TiledRange{tv, tv, to, to}), Rc(world, TiledRange{tv, tv, to, to});
    ContractOptions opt;
    opt.precision = Precision::BF16_LIMBS;
    opt.nslices = 3;
    Rm("a,b,i,j") = (T("c,d,i,j") * V("a,c,b,d")).with(opt);
    opt.permute_copy = true;
    Rc("a,b,i,j") = (T("c,d,i,j") * V("a,c,b,d")).with(opt);
    CHECK(dense(Rm) == dense(Rc), "index-map split differs from the permuted-copy split");
  });

  run("split precision, 5 INT8 limbs, ladder through the index map", [&] {
    auto T = make_array<double>(world, TiledRange{tv, tv, to, to}, rng, false);
    auto V = make_array<double>(world, TiledRange{tv, tv, tv, tv}, rng, false);
    DistArray<double> R(world, TiledRange{tv, tv, to, to});
    ContractOptions opt;
    opt.precision = Precision::INT8_LIMBS;
    opt.nslices = 5;
    R("a,b,i,j") = (T("c,d,i,j") * V("a,c,b,d")).with(opt);
    std::vector<int64_t> er;
    auto ref = einsum(dense(T), "c,d,i,j", extents(T), dense(V), "a,c,b,d", extents(V),
                      "a,b,i,j", er);
    const double err = rel_fro(dense(R), ref);
    CHECK(err <= 3e-8, "relative error %g", err);
  });

  run("autotuned tile shape, block-sparse", [&] {
    TiledRange1 tb{0, 37, 90, 150, 151, 260, 333};
    auto A = make_array<double>(world, TiledRange{tb, tb}, rng, true, 0.7);
    auto B = make_array<double>(world, TiledRange{tb, tb}, rng, true, 0.7);
    DistArray<double> C(world, TiledRange{tb, tb});
    ContractOptions opt;
    opt.autotune = true;
    C("i,j") = (A("i,k") * B("k,j")).with(opt);
    std::vector<int64_t> er;
    auto ref = einsum(dense(A), "i,k", extents(A), dense(B), "k,j", extents(B), "i,j", er);
    CHECK(dense(C) == ref, "autotuned result differs");
  });

  run("graph capture and replay of a tile permutation", [&] {
    auto V = make_array<double>(world, TiledRange{tv, to, tv}, rng, true);
    const int64_t n = V.pool_elems();
    DeviceBuffer<double> dst(n);
    const int32_t perm[3] = {2, 0, 1};
    std::vector<int64_t> ext, src_off, dst_off;
    for (size_t t = 0; t < V.trange().ntiles(); ++t) {
      auto e = V.trange().tile_extents(t);
      ext.insert(ext.end(), e.begin(), e.end());
      src_off.push_back(V.offsets()[t]);
      dst_off.push_back(V.offsets()[t]);
    }
    tamx_permute_plan* plan = nullptr;
    check(tamx_permute_plan_create(world.ctx(), 3, perm, (int64_t)V.trange().ntiles(), ext.data(),
                                   src_off.data(), dst_off.data(), &plan),
          "tamx_permute_plan_create");
    check(tamx_permute_f64(world.ctx(), plan, V.pool(), dst.get(), world.stream()),
          "tamx_permute_f64");
    std::vector<double> eager(n), replay(n);
    hip_check(hipMemcpyAsync(eager.data(), dst.get(), n * sizeof(double), hipMemcpyDeviceToHost,
                             static_cast<hipStream_t>(world.stream())), "copy");
    world.fence();
    hip_check(hipMemset(dst.get(), 0, n * sizeof(double)), "clear");
    {
      Graph g(world, [&] {
        check(tamx_permute_f64(world.ctx(), plan, V.pool(), dst.get(), world.stream()),
              "tamx_permute_f64");
      });
      for (int rep = 0; rep < 3; ++rep) {
        hip_check(hipMemset(dst.get(), 0, n * sizeof(double)), "clear");
        g.launch();
        hip_check(hipMemcpyAsync(replay.data(), dst.get(), n * sizeof(double),
                                 hipMemcpyDeviceToHost, static_cast<hipStream_t>(world.stream())),
                  "copy");
        world.fence();
        CHECK(replay == eager, "replay %d differs from the eager permute", rep);
      }
    }
    tamx_permute_plan_destroy(plan);
  });

  run("float tiles", [&] {
    auto A = make_array<float>(world, TiledRange{ti, tk}, rng, true);
    auto B = make_array<float>(world, TiledRange{tk, tj}, rng, true);
    DistArray<float> C(world, TiledRange{ti, tj});
    C("i,j") = A("i,k") * B("k,j");
    std::vector<int64_t> er;
    auto ref = einsum(dense(A), "i,k", extents(A), dense(B), "k,j", extents(B), "i,j", er);
    CHECK(dense(C) == ref, "float result differs");
  });

  run("incongruent contracted tilings throw", [&] {
    auto A = make_array<double>(world, TiledRange{ti, tk}, rng, true);
    auto B = make_array<double>(world, TiledRange{tj, tj}, rng, true);
    DistArray<double> C(world, TiledRange{ti, tj});
    bool threw = false;
    try {
      C("i,j") = A("i,k") * B("k,j");
    } catch (const Exception& e) {
      threw = e.status() == TAMX_ERR_BAD_ARG;
    }
    CHECK(threw, "no exception");
  });

  run("bad annotation throws", [&] {
    auto A = make_array<double>(world, TiledRange{ti, tk}, rng, true);
    auto B = make_array<double>(world, TiledRange{tk, tj}, rng, true);
    DistArray<double> C(world, TiledRange{ti, tj});
    bool threw = false;
    try {
      C("i,q") = A("i,k") * B("k,j");
    } catch (const Exception&) {
      threw = true;
    }
    CHECK(threw, "no exception");
  });

  if (g_failures) {
    std::printf("%d FAILURES\n", g_failures);
    return 1;
  }
  std::printf("ALL PASSED\n");
  return 0;
}
