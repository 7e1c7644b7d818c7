// FP32 gathered entry point of the grouped GEMM (its own translation unit
// so the FP64 and FP32 instantiations compile in parallel).
#include "tamx_gemm.hpp"

extern "C" {

int tamx_grouped_gemm_gather_f32(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
                                 float alpha, const float* A,
                                 const tamx_gather_plan* a_gather, const float* B,
                                 const tamx_gather_plan* b_gather, float* C,
                                 const tamx_gemm_config* cfg, void* stream) {
  return gemm_gathered<float>(ctx, wl, op_a, op_b, alpha, A, a_gather, B, b_gather, C, nullptr,
                              cfg, stream);
}

int tamx_grouped_gemm_scatter_f32(tamx_context* ctx, tamx_worklist* wl, int op_a, int op_b,
                                  float alpha, const float* A,
                                  const tamx_gather_plan* a_gather, const float* B,
                                  const tamx_gather_plan* b_gather, float* C,
                                  const tamx_gather_plan* c_scatter,
                                  const tamx_gemm_config* cfg, void* stream) {
  TAMX_REQUIRE(c_scatter != nullptr, "null scatter plan");
  return gemm_gathered<float>(ctx, wl, op_a, op_b, alpha, A, a_gather, B, b_gather, C, c_scatter,
                              cfg, stream);
}

}  // extern "C"
