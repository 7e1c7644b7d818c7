// C-ABI tests of two ways TiledArray's own callers reach the contraction
// tile work, each against the oracle (oracle/ta_oracle.h) on integer-valued
// tiles, so every result is exact and compared for equality:
//
//  1. The per-pair binding: ContractReduce's gemm customisation point
//     (tile_op/contract_reduce.h:409-453 -> tile_op/tile_interface.h:805-830)
//     called once per tile pair of a result tile, the first call with
//     beta = 0 into a fresh result, later calls with beta = 1; each call is
//     one task of one pair in a work list that the caller fills itself
//     (tamx_worklist_buffers), the tiles' own device pointers as the pools.
//     Reference: ora_dgemm chained the same way.
//
//  2. Concurrent submission: four host threads share one context, each with
//     its own stream, work list and block-sparse problem, and each runs the
//     device schedule (tamx_contraction_worklist) and the grouped GEMM
//     several times; TA's pool threads call the tile ops concurrently
//     (reduce_task.h:451-539, serialised only per result tile) and the
//     device build gives each range its own stream (device/blas.h:57-79).
//     Reference: ora_contract_f64 on each thread's problem.
//
// Prints one line per case and "ALL PASSED" at the end; exit status 0 on
// success.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "tamx.h"
#include "../../oracle/ta_oracle.h"

static std::atomic<int> g_failures{0};

#define CHECK(cond, ...)                                      \
  do {                                                        \
    if (!(cond)) {                                            \
      std::printf("  FAILED %s:%d: ", __FILE__, __LINE__);    \
      std::printf(__VA_ARGS__);                               \
      std::printf("\n");                                      \
      ++g_failures;                                           \
    }                                                         \
  } while (0)

#define HIP_OK(expr)                                                             \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      std::printf("  HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                     \
      std::exit(2);                                                              \
    }                                                                            \
  } while (0)

#define TAMX_OK_(expr)                                                         \
  do {                                                                         \
    int rc_ = (expr);                                                          \
    if (rc_ != TAMX_OK) {                                                      \
      std::printf("  tamx error %d (%s) at %s:%d\n", rc_, tamx_last_error(),  \
                  __FILE__, __LINE__);                                         \
      std::exit(2);                                                            \
    }                                                                          \
  } while (0)

template <typename T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  explicit Dev(size_t n_) : n(n_) { HIP_OK(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T))); }
  Dev(const std::vector<T>& h) : Dev(h.size()) {
    if (n) HIP_OK(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
  }
  ~Dev() { (void)hipFree(p); }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
};

// ------------------------------------------------ 1. per-pair binding --

// One result tile C (m x n) = factor * sum_p op(A_p) op(B_p), one call per
// pair p, through a one-task one-pair work list; returns the failures.
static void per_pair_case(tamx_context* ctx, int op_a, int op_b, std::mt19937& rng) {
  const int m = 37, n = 50;
  const std::vector<int> ks = {16, 1, 64, 23, 130};
  const double factor = 3.0;
  std::uniform_int_distribution<int> val(-8, 8);
  const int np = (int)ks.size();
  std::vector<std::vector<double>> Ah(np), Bh(np);
  std::vector<Dev<double>*> Ad, Bd;
  for (int p = 0; p < np; ++p) {
    Ah[p].resize((size_t)m * ks[p]);
    Bh[p].resize((size_t)ks[p] * n);
    for (auto& x : Ah[p]) x = val(rng);
    for (auto& x : Bh[p]) x = val(rng);
    Ad.push_back(new Dev<double>(Ah[p]));
    Bd.push_back(new Dev<double>(Bh[p]));
  }
  // the result tile holds stale data before the first (beta = 0) call
  std::vector<double> Ch((size_t)m * n, 1234.0);
  Dev<double> Cd(Ch);

  hipStream_t s;
  HIP_OK(hipStreamCreate(&s));
  tamx_worklist* wl = nullptr;
  TAMX_OK_(tamx_worklist_create(ctx, 1, 1, 1, &wl));
  tamx_gemm_task* dt;
  tamx_gemm_pair* dp;
  int64_t* dc;
  TAMX_OK_(tamx_worklist_buffers(wl, &dt, &dp, &dc));
  // the descriptors of every call in pinned host memory, one slot each, so
  // the asynchronous copies never read a slot that a later call rewrites
  tamx_gemm_task* ht;
  tamx_gemm_pair* hp;
  int64_t* hc;
  HIP_OK(hipHostMalloc((void**)&ht, np * sizeof(tamx_gemm_task), hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&hp, np * sizeof(tamx_gemm_pair), hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&hc, np * TAMX_WL_NCOUNTERS * sizeof(int64_t),
                       hipHostMallocDefault));
  for (int p = 0; p < np; ++p) {
    const int k = ks[p];
    ht[p] = tamx_gemm_task{0, m, n, n, 0, 1, p == 0 ? 0 : TAMX_TASK_BETA1};
    // TA's leading dimensions (tensor/kernels.h:146-158)
    hp[p] = tamx_gemm_pair{0, 0, k, op_a ? m : k, op_b ? k : n, 0, 0, 0};
    int64_t* c = hc + p * TAMX_WL_NCOUNTERS;
    std::memset(c, 0, TAMX_WL_NCOUNTERS * sizeof(int64_t));
    c[TAMX_WL_NTASKS] = 1;
    c[TAMX_WL_NPAIRS] = 1;
    c[TAMX_WL_CELEMS] = (int64_t)m * n;
    HIP_OK(hipMemcpyAsync(dt, ht + p, sizeof(tamx_gemm_task), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(dp, hp + p, sizeof(tamx_gemm_pair), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(dc, c, TAMX_WL_NCOUNTERS * sizeof(int64_t), hipMemcpyHostToDevice, s));
    // the operand tiles are stored as op() asks: A k x m when op_a, B n x k
    // when op_b (the same data read transposed)
    TAMX_OK_(tamx_grouped_gemm_f64(ctx, wl, op_a, op_b, factor, Ad[p]->p, Bd[p]->p, Cd.p, s));
  }
  HIP_OK(hipMemcpyAsync(Ch.data(), Cd.p, Ch.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));

  std::vector<double> ref((size_t)m * n, 1234.0);
  for (int p = 0; p < np; ++p) {
    const int k = ks[p];
    ora_dgemm(op_a, op_b, m, n, k, factor, Ah[p].data(), op_a ? m : k, Bh[p].data(),
              op_b ? k : n, p == 0 ? 0.0 : 1.0, ref.data(), n);
  }
  CHECK(Ch == ref, "op_a %d op_b %d: result differs from the oracle's chained gemm", op_a, op_b);

  (void)hipHostFree(ht);
  (void)hipHostFree(hp);
  (void)hipHostFree(hc);
  tamx_worklist_destroy(wl);
  (void)hipStreamDestroy(s);
  for (auto* d : Ad) delete d;
  for (auto* d : Bd) delete d;
}

// ------------------------------------------ 2. concurrent submission --

struct Problem {
  int op_a, op_b;
  int64_t Mt, Nt, Kt;
  std::vector<int32_t> mext, next, kext;
  std::vector<int64_t> a_off, b_off;
  std::vector<double> A, B;
  int64_t c_capacity = 0;
};

static std::vector<int32_t> widths(std::mt19937& rng, int64_t nt, int lo, int hi) {
  std::uniform_int_distribution<int> w(lo, hi);
  std::vector<int32_t> v(nt);
  for (auto& x : v) x = w(rng);
  return v;
}

static Problem make_problem(unsigned seed) {
  std::mt19937 rng(seed);
  Problem P;
  P.op_a = (int)(seed & 1);
  P.op_b = (int)((seed >> 1) & 1);
  P.Mt = 6;
  P.Nt = 5;
  P.Kt = 7;
  P.mext = widths(rng, P.Mt, 20, 90);
  P.next = widths(rng, P.Nt, 20, 90);
  P.kext = widths(rng, P.Kt, 1, 80);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::uniform_int_distribution<int> val(-8, 8);
  auto fill = [&](std::vector<int64_t>& off, std::vector<double>& pool, int64_t R, int64_t Cc,
                  const std::vector<int32_t>& re, const std::vector<int32_t>& ce) {
    off.assign(R * Cc, -1);
    int64_t pos = 0;
    for (int64_t r = 0; r < R; ++r)
      for (int64_t c = 0; c < Cc; ++c)
        if (u(rng) < 0.7) {
          off[r * Cc + c] = pos;
          pos += (int64_t)re[r] * ce[c];
        }
    pool.resize(std::max<int64_t>(1, pos));
    for (auto& x : pool) x = val(rng);
  };
  // storage grids: A Mt x Kt (Kt x Mt when op_a), B Kt x Nt (Nt x Kt when op_b)
  if (P.op_a) fill(P.a_off, P.A, P.Kt, P.Mt, P.kext, P.mext);
  else fill(P.a_off, P.A, P.Mt, P.Kt, P.mext, P.kext);
  if (P.op_b) fill(P.b_off, P.B, P.Nt, P.Kt, P.next, P.kext);
  else fill(P.b_off, P.B, P.Kt, P.Nt, P.kext, P.next);
  int64_t sm = 0, sn = 0;
  for (auto x : P.mext) sm += x;
  for (auto x : P.next) sn += x;
  P.c_capacity = sm * sn;
  return P;
}

// One thread: its own stream and work list on the shared context; runs the
// device schedule and the grouped GEMM `iters` times and checks every result
// against the oracle's.
static void thread_body(tamx_context* ctx, const Problem& P, const std::vector<double>& ref,
                        const std::vector<int64_t>& ref_coff, int tid, int iters) {
  // the result shape: 1 where a tile pair survives, else 0 (screened out
  // with threshold 0.5), so the schedule allocates exactly those tiles
  std::vector<float> est(ref_coff.size());
  for (size_t t = 0; t < est.size(); ++t) est[t] = ref_coff[t] >= 0 ? 1.0f : 0.0f;
  HIP_OK(hipSetDevice(0));
  hipStream_t s;
  HIP_OK(hipStreamCreate(&s));
  tamx_worklist* wl = nullptr;
  const int64_t grid = P.Mt * P.Nt;
  TAMX_OK_(tamx_worklist_create(ctx, grid, grid * P.Kt, grid, &wl));
  Dev<int32_t> mext(P.mext), next(P.next), kext(P.kext);
  Dev<int64_t> a_off(P.a_off), b_off(P.b_off), c_off(grid);
  Dev<float> c_norms(est);
  Dev<double> A(P.A), B(P.B), C(P.c_capacity);
  tamx_contraction_desc d{};
  d.op_a = P.op_a;
  d.op_b = P.op_b;
  d.Mt = P.Mt;
  d.Nt = P.Nt;
  d.Kt = P.Kt;
  d.nbatch = 1;
  d.mext = mext.p;
  d.next = next.p;
  d.kext = kext.p;
  d.a_off = a_off.p;
  d.b_off = b_off.p;
  d.c_norms = c_norms.p;
  d.c_threshold = 0.5f;
  d.c_off = c_off.p;
  d.c_capacity = P.c_capacity;
  std::vector<double> got(P.c_capacity);
  std::vector<int64_t> coff(grid);
  for (int it = 0; it < iters; ++it) {
    HIP_OK(hipMemsetAsync(C.p, 0, C.n * sizeof(double), s));
    TAMX_OK_(tamx_contraction_worklist(ctx, &d, wl, s));
    TAMX_OK_(tamx_grouped_gemm_f64(ctx, wl, P.op_a, P.op_b, -2.0, A.p, B.p, C.p, s));
    HIP_OK(hipMemcpyAsync(got.data(), C.p, got.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(coff.data(), c_off.p, coff.size() * sizeof(int64_t),
                          hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    CHECK(coff == ref_coff, "thread %d iteration %d: result tile offsets differ", tid, it);
    bool same = true;
    for (int64_t t = 0; t < grid; ++t) {
      if (ref_coff[t] < 0) continue;
      const int64_t vol = (int64_t)P.mext[t / P.Nt] * P.next[t % P.Nt];
      for (int64_t e = 0; e < vol; ++e) same = same && got[ref_coff[t] + e] == ref[ref_coff[t] + e];
    }
    CHECK(same, "thread %d iteration %d: result differs from the oracle", tid, it);
  }
  tamx_worklist_destroy(wl);
  (void)hipStreamDestroy(s);
}

static void run(const char* name, void (*f)(tamx_context*), tamx_context* ctx) {
  const int before = g_failures;
  f(ctx);
  std::printf("%s %s\n", g_failures == before ? "ok  " : "FAIL", name);
  std::fflush(stdout);
}

static void per_pair_all_ops(tamx_context* ctx) {
  std::mt19937 rng(7);
  for (int op_a = 0; op_a < 2; ++op_a)
    for (int op_b = 0; op_b < 2; ++op_b) per_pair_case(ctx, op_a, op_b, rng);
}

static void four_threads(tamx_context* ctx) {
  constexpr int kThreads = 4, kIters = 6;
  std::vector<Problem> probs;
  std::vector<std::vector<double>> refs;
  std::vector<std::vector<int64_t>> coffs;
  for (int t = 0; t < kThreads; ++t) {
    probs.push_back(make_problem(100 + t));
    const Problem& P = probs.back();
    // every result tile with at least one surviving pair, packed in
    // row-major tile order (the work list's allocation order)
    std::vector<int64_t> coff(P.Mt * P.Nt, -1);
    int64_t pos = 0;
    for (int64_t i = 0; i < P.Mt; ++i)
      for (int64_t j = 0; j < P.Nt; ++j) {
        bool any = false;
        for (int64_t k = 0; k < P.Kt; ++k) {
          const int64_t ai = P.op_a ? k * P.Mt + i : i * P.Kt + k;
          const int64_t bi = P.op_b ? j * P.Kt + k : k * P.Nt + j;
          any = any || (P.a_off[ai] >= 0 && P.b_off[bi] >= 0);
        }
        if (any) {
          coff[i * P.Nt + j] = pos;
          pos += (int64_t)P.mext[i] * P.next[j];
        }
      }
    std::vector<double> ref(P.c_capacity, 0.0);
    ora_contract_f64(P.op_a, P.op_b, P.Mt, P.Nt, P.Kt, P.mext.data(), P.next.data(),
                     P.kext.data(), P.A.data(), P.a_off.data(), P.B.data(), P.b_off.data(),
                     ref.data(), coff.data(), -2.0, nullptr, nullptr, 0.0f);
    refs.push_back(std::move(ref));
    coffs.push_back(std::move(coff));
  }
  std::vector<std::thread> ths;
  for (int t = 0; t < kThreads; ++t)
    ths.emplace_back(thread_body, ctx, std::cref(probs[t]), std::cref(refs[t]),
                     std::cref(coffs[t]), t, kIters);
  for (auto& th : ths) th.join();
}

int main() {
  HIP_OK(hipSetDevice(0));
  tamx_context* ctx = nullptr;
  TAMX_OK_(tamx_context_create(0, &ctx));
  run("per-pair binding: one task per call, beta 0 then 1, all four ops", per_pair_all_ops, ctx);
  run("four threads, own streams and work lists, one context", four_threads, ctx);
  tamx_context_destroy(ctx);
  if (g_failures) {
    std::printf("%d FAILURES\n", g_failures.load());
    return 1;
  }
  std::printf("ALL PASSED\n");
  return 0;
}
