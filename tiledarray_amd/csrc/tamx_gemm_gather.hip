// Gathered entry points of the grouped GEMM: an operand that needs a general
// permutation before the GEMM is read straight from its stored tiles through
// a per-tile index map (tamx_gather_plan), so no permuted copy is made.  The
// kernels are the plain one's instantiations with a gather mode GM > 0
// (tamx_gemm.hpp); this unit holds the plan and the FP64 entry point.
#include <new>
#include <vector>

#include "tamx_gemm.hpp"

extern "C" {

int tamx_gather_plan_create(tamx_context* ctx, int64_t ntiles, int nrow_modes, int nk_modes,
                            const int32_t* host_row_ext, const int64_t* host_row_str,
                            const int32_t* host_k_ext, const int64_t* host_k_str,
                            tamx_gather_plan** out) {
  TAMX_REQUIRE(out != nullptr, "null output pointer");
  *out = nullptr;
  TAMX_REQUIRE(ntiles >= 0 && ntiles < (int64_t(1) << 31), "tile count out of range");
  TAMX_REQUIRE(nrow_modes >= 1 && nrow_modes <= 4 && nk_modes >= 1 && nk_modes <= 4,
               "row and k modes must number 1 to 4 each");
  TAMX_REQUIRE(ntiles == 0 || (host_row_ext && host_row_str && host_k_ext && host_k_str),
               "null tile table");
  if (int rc = tamx::activate(ctx)) return rc;
  std::vector<tamx::GatherTile> tiles(ntiles);
  for (int64_t t = 0; t < ntiles; ++t) {
    // modes right-aligned: padding (extent 1, stride 0) at the slow end
    tamx::GatherTile& g = tiles[t];
    for (int d = 0; d < 4; ++d) g.r[d] = g.k[d] = tamx::make_gather_mode(1, 0);
    // the largest offset the map reaches must fit 31 bits
    int64_t reach = 0;
    for (int d = 0; d < nrow_modes + nk_modes; ++d) {
      const bool row = d < nrow_modes;
      const int64_t i = row ? t * nrow_modes + d : t * nk_modes + (d - nrow_modes);
      const int64_t e = row ? host_row_ext[i] : host_k_ext[i];
      const int64_t st = row ? host_row_str[i] : host_k_str[i];
      TAMX_REQUIRE(e >= 0 && st >= 0, "negative extent or stride");
      TAMX_REQUIRE(e < (int64_t(1) << 31) && st < (int64_t(1) << 31), "extent or stride too large");
      if (e > 0) reach += (e - 1) * st;
      TAMX_REQUIRE(reach < (int64_t(1) << 31), "tile offsets exceed 31 bits");
      if (row)
        g.r[4 - nrow_modes + d] = tamx::make_gather_mode((int32_t)e, (int32_t)st);
      else
        g.k[4 - nk_modes + (d - nrow_modes)] = tamx::make_gather_mode((int32_t)e, (int32_t)st);
    }
  }
  auto* plan = new (std::nothrow) tamx_gather_plan();
  if (!plan) return TAMX_ERR_OOM;
  plan->ctx = ctx;
  plan->ntiles = ntiles;
  hipError_t e = hipMalloc(&plan->tiles, std::max<size_t>(1, tiles.size()) * sizeof(tamx::GatherTile));
  if (e == hipSuccess && !tiles.empty())
    e = hipMemcpy(plan->tiles, tiles.data(), tiles.size() * sizeof(tamx::GatherTile),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    tamx::set_error(std::string("gather plan: ") + hipGetErrorString(e));
    if (plan->tiles) (void)hipFree(plan->tiles);
    delete plan;
    return e == hipErrorOutOfMemory ? TAMX_ERR_OOM : TAMX_ERR_DEVICE;
  }
  *out = plan;
  return TAMX_OK;
}

int tamx_gather_plan_destroy(tamx_gather_plan* plan) {
  if (!plan) return TAMX_OK;
  if (plan->tiles) (void)hipFree(plan->tiles);
  delete plan;
  return TAMX_OK;
}

int tamx_grouped_gemm_gather_f64(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
                                 double alpha, const double* A,
                                 const tamx_gather_plan* a_gather, const double* B,
                                 const tamx_gather_plan* b_gather, double* C,
                                 const tamx_gemm_config* cfg, void* stream) {
  return gemm_gathered<double>(ctx, wl, op_a, op_b, alpha, A, a_gather, B, b_gather, C,
                               nullptr, cfg, stream);
}

int tamx_grouped_gemm_scatter_f64(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
                                  double alpha, const double* A,
                                  const tamx_gather_plan* a_gather, const double* B,
                                  const tamx_gather_plan* b_gather, double* C,
                                  const tamx_gather_plan* c_scatter,
                                  const tamx_gemm_config* cfg, void* stream) {
  TAMX_REQUIRE(c_scatter != nullptr, "null scatter plan");
  return gemm_gathered<double>(ctx, wl, op_a, op_b, alpha, A, a_gather, B, b_gather, C,
                               c_scatter, cfg, stream);
}

}  // extern "C"
