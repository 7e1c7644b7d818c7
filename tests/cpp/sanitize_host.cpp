// Host code of the path under AddressSanitizer and UndefinedBehaviorSanitizer
// (built by tests/test_sanitize_host.py with g++ -fsanitize=address,undefined):
// the contraction planner (tiledarray_amd/csrc/tamx_plan.cpp) on hand cases,
// malformed annotations and random index patterns, and every function of the
// CPU oracle (oracle/ta_oracle.c) on small random block-sparse inputs,
// checked against brute-force loops.  A sanitizer report or a failed check
// ends the run with a non-zero status; "SANITIZE OK" on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "tamx.h"
#include "../../oracle/ta_oracle.h"

// The planner reports argument errors through tamx::set_error, which the
// library defines beside its device code; this host-only build records the
// message here instead.
#include <string>
namespace tamx {
static std::string g_msg;
void set_error(const std::string& msg) { g_msg = msg; }
}  // namespace tamx

static int g_fail = 0;
#define CHECK(c, ...)                          \
  do {                                         \
    if (!(c)) {                                \
      std::printf("FAILED %d: ", __LINE__);    \
      std::printf(__VA_ARGS__);                \
      std::printf("\n");                       \
      ++g_fail;                                \
    }                                          \
  } while (0)

static void planner_cases() {
  // ok: whether the planner accepts the pattern (fused Hadamard indices such
  // as n below are planned by the fused-mode product, not here)
  struct Case { const char *l, *r, *t; int ok; };
  const Case cases[] = {
      {"m,k", "k,n", "m,n", 1},          {"k,m", "k,n", "m,n", 1},
      {"c,d,i,j", "a,b,c,d", "a,b,i,j", 1}, {"c,d,i,j", "a,c,b,d", "a,b,i,j", 1},
      {"n,i,k", "n,k,j", "n,i,j", 0},    {"i,k", "k,j", "j,i", 1},
      {"a,b,c,d,e,f,g,h", "h,g,x", "a,b,c,d,e,f,x", 1},
      {"i,k", "k,j", "i,q", 0},          {"", "k,j", "j", 0},
      {"i,,k", "k,j", "i,j", 0},         {"a,b,c,d,e,f,g,h,i", "i,x", "x", 0},
      {"i,i", "i,j", "j", 0},
  };
  for (const Case& c : cases) {
    for (int pref = 0; pref < 2; ++pref) {
      tamx_contraction_plan p;
      const int rc = tamx_plan_contraction(c.l, c.r, c.t, pref, &p);
      CHECK((rc == TAMX_OK) == (c.ok == 1), "%s * %s -> %s: rc %d (%s)", c.l, c.r, c.t, rc,
            tamx::g_msg.c_str());
      if (rc == TAMX_OK) {
        CHECK(p.result_rank <= TAMX_PLAN_MAX_RANK && p.left_rank <= TAMX_PLAN_MAX_RANK,
              "ranks out of range");
        // the result permutation is a permutation
        std::vector<int> seen(p.result_rank, 0);
        for (int d = 0; d < p.result_rank; ++d)
          if (p.result_perm[d] >= 0 && p.result_perm[d] < p.result_rank) ++seen[p.result_perm[d]];
        for (int d = 0; d < p.result_rank; ++d) CHECK(seen[d] == 1, "result_perm not a permutation");
      }
    }
  }
  // random index patterns: every plan's operand permutations are permutations
  std::mt19937 rng(7);
  const std::string names = "abcdefgh";
  for (int trial = 0; trial < 2000; ++trial) {
    const int nl = 1 + rng() % 3, nr = 1 + rng() % 3, nk = 1 + rng() % 3;
    std::vector<std::string> L, R, K, T;
    int at = 0;
    for (int i = 0; i < nk; ++i) K.push_back(std::string(1, names[at++]));
    for (int i = 0; i < nl; ++i) L.push_back(std::string(1, names[at++]));
    for (int i = 0; i < nr && at < 8; ++i) R.push_back(std::string(1, names[at++]));
    std::vector<std::string> left = L, right = R;
    left.insert(left.end(), K.begin(), K.end());
    right.insert(right.end(), K.begin(), K.end());
    T = L;
    T.insert(T.end(), R.begin(), R.end());
    std::shuffle(left.begin(), left.end(), rng);
    std::shuffle(right.begin(), right.end(), rng);
    std::shuffle(T.begin(), T.end(), rng);
    auto join = [](const std::vector<std::string>& v) {
      std::string s;
      for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + v[i];
      return s;
    };
    tamx_contraction_plan p;
    const int rc = tamx_plan_contraction(join(left).c_str(), join(right).c_str(),
                                         join(T).c_str(), (int)(rng() % 2), &p);
    CHECK(rc == TAMX_OK, "random pattern %d: %s", trial, tamx::g_msg.c_str());
    if (rc != TAMX_OK) continue;
    for (int side = 0; side < 2; ++side) {
      const int rank = side ? p.right_rank : p.left_rank;
      const int32_t* perm = side ? p.right_perm : p.left_perm;
      std::vector<int> seen(rank, 0);
      for (int d = 0; d < rank; ++d)
        if (perm[d] >= 0 && perm[d] < rank) ++seen[perm[d]];
      for (int d = 0; d < rank; ++d) CHECK(seen[d] == 1, "operand permutation, trial %d", trial);
    }
  }
}

static void oracle_cases() {
  std::mt19937 rng(11);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  // gemm against the definition, all ops
  for (int op_a = 0; op_a < 2; ++op_a)
    for (int op_b = 0; op_b < 2; ++op_b) {
      const int m = 7, n = 5, k = 9;
      std::vector<double> A(m * k), B(k * n), C(m * n, 0.5), R(m * n, 0.5);
      for (auto& x : A) x = std::round(8 * u(rng));
      for (auto& x : B) x = std::round(8 * u(rng));
      ora_dgemm(op_a, op_b, m, n, k, 2.0, A.data(), op_a ? m : k, B.data(), op_b ? k : n, 1.0,
                C.data(), n);
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
          double s = 0;
          for (int l = 0; l < k; ++l)
            s += (op_a ? A[l * m + i] : A[i * k + l]) * (op_b ? B[j * k + l] : B[l * n + j]);
          R[i * n + j] += 2.0 * s;
        }
      CHECK(C == R, "ora_dgemm op %d %d", op_a, op_b);
    }
  // norms, scaling, shape GEMM and the one-rank contraction on a 3 x 4 x 2
  // block-sparse problem
  const int Mt = 3, Nt = 2, Kt = 4;
  const std::vector<int32_t> mext = {3, 1, 4}, next = {2, 5}, kext = {1, 3, 2, 6};
  std::vector<int64_t> a_off(Mt * Kt, -1), b_off(Kt * Nt, -1), c_off(Mt * Nt, -1);
  std::vector<double> A, B;
  for (int i = 0; i < Mt; ++i)
    for (int k = 0; k < Kt; ++k)
      if ((i + k) % 3) {
        a_off[i * Kt + k] = (int64_t)A.size();
        for (int e = 0; e < mext[i] * kext[k]; ++e) A.push_back(std::round(8 * u(rng)));
      }
  for (int k = 0; k < Kt; ++k)
    for (int j = 0; j < Nt; ++j)
      if ((k + 2 * j) % 4) {
        b_off[k * Nt + j] = (int64_t)B.size();
        for (int e = 0; e < kext[k] * next[j]; ++e) B.push_back(std::round(8 * u(rng)));
      }
  std::vector<int64_t> a_vol(Mt * Kt), b_vol(Kt * Nt);
  for (int i = 0; i < Mt; ++i)
    for (int k = 0; k < Kt; ++k) a_vol[i * Kt + k] = mext[i] * kext[k];
  for (int k = 0; k < Kt; ++k)
    for (int j = 0; j < Nt; ++j) b_vol[k * Nt + j] = kext[k] * next[j];
  std::vector<float> an(Mt * Kt), bn(Kt * Nt);
  ora_tile_norms_f64(A.data(), a_off.data(), a_vol.data(), Mt * Kt, an.data());
  ora_tile_norms_f64(B.data(), b_off.data(), b_vol.data(), Kt * Nt, bn.data());
  const int64_t ant[2] = {Mt, Kt}, bnt[2] = {Kt, Nt};
  std::vector<float> asz, bsz;
  for (int i = 0; i < Mt; ++i) asz.push_back((float)mext[i]);
  for (int k = 0; k < Kt; ++k) asz.push_back((float)kext[k]);
  for (int k = 0; k < Kt; ++k) bsz.push_back((float)kext[k]);
  for (int j = 0; j < Nt; ++j) bsz.push_back((float)next[j]);
  ora_scale_tile_norms(an.data(), 2, ant, asz.data(), 1e-3f);
  ora_scale_tile_norms(bn.data(), 2, bnt, bsz.data(), 1e-3f);
  std::vector<float> ksz(kext.begin(), kext.end()), est(Mt * Nt);
  ora_shape_gemm(an.data(), bn.data(), 0, 0, Mt, Nt, Kt, ksz.data(), 1.0f, 1e-3f, est.data());
  int64_t cvol = 0;
  for (int i = 0; i < Mt; ++i)
    for (int j = 0; j < Nt; ++j)
      if (est[i * Nt + j] > 0) {
        c_off[i * Nt + j] = cvol;
        cvol += mext[i] * next[j];
      }
  std::vector<double> C(cvol + 1, 0.0);
  ora_set_num_threads(2);
  const int64_t np = ora_contract_f64(0, 0, Mt, Nt, Kt, mext.data(), next.data(), kext.data(),
                                      A.data(), a_off.data(), B.data(), b_off.data(), C.data(),
                                      c_off.data(), 1.0, nullptr, nullptr, 0.0f);
  CHECK(np > 0, "no pairs");
  // brute force over the same pattern
  for (int i = 0; i < Mt; ++i)
    for (int j = 0; j < Nt; ++j) {
      if (c_off[i * Nt + j] < 0) continue;
      for (int r = 0; r < mext[i]; ++r)
        for (int c = 0; c < next[j]; ++c) {
          double s = 0;
          for (int k = 0; k < Kt; ++k) {
            if (a_off[i * Kt + k] < 0 || b_off[k * Nt + j] < 0) continue;
            for (int l = 0; l < kext[k]; ++l)
              s += A[a_off[i * Kt + k] + r * kext[k] + l] * B[b_off[k * Nt + j] + l * next[j] + c];
          }
          CHECK(C[c_off[i * Nt + j] + r * next[j] + c] == s, "contract (%d,%d)", i, j);
        }
    }
  // permute round trip, ranks 2-6
  for (int rank = 2; rank <= 6; ++rank) {
    std::vector<int64_t> ext(rank);
    int64_t vol = 1;
    for (int d = 0; d < rank; ++d) vol *= (ext[d] = 1 + rng() % 4);
    std::vector<int32_t> perm(rank), inv(rank);
    for (int d = 0; d < rank; ++d) perm[d] = d;
    std::shuffle(perm.begin(), perm.end(), rng);
    for (int d = 0; d < rank; ++d) inv[perm[d]] = d;
    std::vector<double> x(vol), y(vol), z(vol);
    for (auto& v : x) v = u(rng);
    ora_permute_f64(x.data(), y.data(), rank, ext.data(), perm.data());
    std::vector<int64_t> pext(rank);
    for (int d = 0; d < rank; ++d) pext[perm[d]] = ext[d];
    ora_permute_f64(y.data(), z.data(), rank, pext.data(), inv.data());
    CHECK(z == x, "permute round trip, rank %d", rank);
  }
}

int main() {
  planner_cases();
  oracle_cases();
  if (g_fail) {
    std::printf("%d FAILURES\n", g_fail);
    return 1;
  }
  std::printf("SANITIZE OK\n");
  return 0;
}
