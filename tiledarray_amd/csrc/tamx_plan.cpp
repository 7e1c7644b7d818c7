// Host-side contraction planner: the decisions ContEngine makes before any
// tile is touched, restated for one rank.
//   - which operand tiles must be permuted and which permutations fold into
//     the GEMM as NoTrans/Trans (GEMMPermutationOptimizer,
//     expressions/permopt.h:254-381, to_cblas_op :47-57);
//   - the natural result index order and the permutation to the target
//     (ContEngine::init_struct, expressions/cont_engine.h:354-470);
//   - the fused GEMM structure (GemmHelper, math/gemm_helper.h:41-278).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "tamx.h"
#include "tamx_internal.hpp"

namespace {

std::vector<std::string> parse_indices(const char* s) {
  std::vector<std::string> out;
  std::string cur;
  for (const char* p = s; *p; ++p) {
    if (*p == ',') {
      out.push_back(cur);
      cur.clear();
    } else if (*p != ' ' && *p != '\t') {
      cur.push_back(*p);
    }
  }
  if (!cur.empty() || !out.empty()) out.push_back(cur);
  return out;
}

int find(const std::vector<std::string>& v, const std::string& x) {
  auto it = std::find(v.begin(), v.end(), x);
  return it == v.end() ? -1 : (int)(it - v.begin());
}

// permopt.h:254-381 (argument-only form; the target-indices constructor of
// GEMMPermutationOptimizer ignores the target, permopt.h:203-213).
void index_list_contraction(const std::vector<std::string>& left,
                            const std::vector<std::string>& right,
                            bool prefer_to_permute_left,
                            std::vector<std::string>& res_left,
                            std::vector<std::string>& res_right,
                            std::vector<std::string>& res, int& left_type,
                            int& right_type) {
  const int lrank = (int)left.size(), rrank = (int)right.size();
  res_left.clear();
  res_right.clear();
  res.clear();
  for (const auto& v : left) {
    if (find(right, v) < 0) {
      res_left.push_back(v);
      res.push_back(v);
    } else {
      res_right.push_back(v);
    }
  }
  const int inner_rank = (int)res_right.size();
  const int left_outer_rank = (int)res_left.size();
  const int right_outer_rank = rrank - inner_rank;
  if (inner_rank == 0) {
    for (const auto& v : right) {
      res_right.push_back(v);
      res.push_back(v);
    }
    left_type = right_type = 3;
    return;
  }
  bool ordered = true, l_no = true, l_tr = true, r_no = true, r_tr = true;
  const bool perm_left =
      (lrank < rrank) || (lrank == rrank && prefer_to_permute_left);
  for (int i = 0; i < rrank; ++i) {
    const std::string& idx = right[i];
    const int j = find(left, idx);
    if (j < 0) {
      res_right.push_back(idx);
      res.push_back(idx);
      continue;
    }
    const int x = (int)res_left.size() - left_outer_rank;
    ordered = ordered && (res_right[x] == idx);
    l_no = l_no && (j >= left_outer_rank);
    l_tr = l_tr && (j < inner_rank);
    r_no = r_no && (i < inner_rank);
    r_tr = r_tr && (i >= right_outer_rank);
    if (ordered) {
      res_left.push_back(idx);
    } else if (perm_left) {
      res_left.push_back(idx);
      res_right[x] = idx;
      l_no = l_tr = false;
    } else {
      res_left.push_back(res_right[x]);
      r_no = r_tr = false;
    }
  }
  auto op = [](bool no, bool tr) { return no ? 1 : (tr ? 2 : 3); };
  left_type = op(l_no, l_tr);
  right_type = op(r_no, r_tr);
}

}  // namespace

extern "C" {

// Plans C(target) = A(left) * B(right) for plain (non-nested) tiles.
// Annotations are comma-separated index names as in TA ("a,b,i,j").
int tamx_plan_contraction(const char* left, const char* right,
                          const char* target, int prefer_to_permute_left,
                          tamx_contraction_plan* out) {
  TAMX_REQUIRE(left && right && target && out, "null argument");
  std::memset(out, 0, sizeof(*out));
  const auto L = parse_indices(left), R = parse_indices(right),
             T = parse_indices(target);
  const int kMax = TAMX_PLAN_MAX_RANK;
  TAMX_REQUIRE(!L.empty() && !R.empty(), "empty operand annotation");
  TAMX_REQUIRE((int)L.size() <= kMax && (int)R.size() <= kMax &&
                   (int)T.size() <= kMax,
               "rank exceeds TAMX_PLAN_MAX_RANK");
  for (const auto* v : {&L, &R, &T})
    for (size_t a = 0; a < v->size(); ++a) {
      TAMX_REQUIRE(!(*v)[a].empty(), "empty index name");
      for (size_t b = a + 1; b < v->size(); ++b)
        TAMX_REQUIRE((*v)[a] != (*v)[b], "repeated index in an annotation");
    }
  std::vector<std::string> tl, tr, res;
  int lt = 0, rt = 0;
  index_list_contraction(L, R, prefer_to_permute_left != 0, tl, tr, res, lt, rt);
  // every result index must appear in the target and vice versa
  TAMX_REQUIRE(res.size() == T.size(),
               "target annotation does not match the free indices");
  for (const auto& v : res)
    TAMX_REQUIRE(find(T, v) >= 0, "target annotation does not match the free indices");
  const int inner = (int)(L.size() + R.size() - res.size()) / 2;
  TAMX_REQUIRE(inner > 0, "outer products are not on the contraction path");
  out->left_rank = (int)L.size();
  out->right_rank = (int)R.size();
  out->result_rank = (int)res.size();
  out->inner_rank = inner;
  out->left_permtype = lt;
  out->right_permtype = rt;
  // to_cblas_op (permopt.h:47-57): only matrix_transpose maps to Trans; a
  // general permutation is applied to the tiles, leaving NoTrans.
  out->op_a = lt == 2 ? TAMX_TRANS : TAMX_NO_TRANS;
  out->op_b = rt == 2 ? TAMX_TRANS : TAMX_NO_TRANS;
  // TO permutations: source mode d goes to position perm[d]
  for (int d = 0; d < (int)L.size(); ++d)
    out->left_perm[d] = lt == 3 ? find(tl, L[d]) : d;
  for (int d = 0; d < (int)R.size(); ++d)
    out->right_perm[d] = rt == 3 ? find(tr, R[d]) : d;
  bool ident = true;
  for (int d = 0; d < (int)res.size(); ++d) {
    out->result_perm[d] = find(T, res[d]);
    ident = ident && out->result_perm[d] == d;
  }
  out->result_perm_identity = ident ? 1 : 0;
  // A result permutation that swaps the (left-outer) and (right-outer)
  // groups is the matrix transpose of the fused result tile.
  const int lo = out->left_rank - inner, ro = out->right_rank - inner;
  bool mt = !ident && lo > 0 && ro > 0;
  for (int d = 0; d < (int)res.size() && mt; ++d)
    mt = out->result_perm[d] == (d < lo ? d + ro : d - lo);
  out->result_perm_matrix_transpose = mt ? 1 : 0;
  // fused modes of the GEMM operands in their GEMM layout: left is
  // (outer, inner) for NoTrans and (inner, outer) for Trans
  // (gemm_helper.h:72-96)
  const auto& lg = lt == 3 ? tl : L;
  const auto& rg = rt == 3 ? tr : R;
  for (int d = 0; d < (int)lg.size(); ++d)
    out->left_gemm_is_inner[d] = find(R, lg[d]) >= 0 ? 1 : 0;
  for (int d = 0; d < (int)rg.size(); ++d)
    out->right_gemm_is_inner[d] = find(L, rg[d]) >= 0 ? 1 : 0;
  auto cpy = [](char* dst, const std::vector<std::string>& v) {
    std::string s;
    for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + v[i];
    std::strncpy(dst, s.c_str(), TAMX_PLAN_NAME_LEN - 1);
  };
  cpy(out->left_target, lg);
  cpy(out->right_target, rg);
  cpy(out->result_natural, res);
  return TAMX_OK;
}

}  // extern "C"
