/*
 * ta_oracle.h -- CPU restatement of TiledArray's one-rank contraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in tiledarray_amd/: only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.  Nothing in the product path
 * links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the TiledArray source tree).
 */
#ifndef TA_ORACLE_H
#define TA_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Name of the BLAS used by ora_dgemm/ora_sgemm ("mkl", or "loops"). */
const char* ora_blas_name(void);
/* Number of OpenMP threads ora_contract_* uses (one task per result tile). */
int ora_num_threads(void);
void ora_set_num_threads(int n);

/* Tensor<double>::norm(): sqrt of a strictly sequential double sum of squares,
 * cast to the float shape type.  tensor/tensor.h:3768-3787,
 * math/vector_op.h:562-577, conversions/make_array.h:170-172. */
float ora_norm_f64(const double* x, int64_t n);
/* Tensor<float>::norm(): the scalar type is float, so the sum is in float. */
float ora_norm_f32(const float* x, int64_t n);
/* Norms of many tiles stored in one pool; offsets[t] < 0 means "no tile"
 * (norm 0). */
void ora_tile_norms_f64(const double* pool, const int64_t* offsets,
                        const int64_t* volumes, int64_t ntiles, float* out);
void ora_tile_norms_f32(const float* pool, const int64_t* offsets,
                        const int64_t* volumes, int64_t ntiles, float* out);

/* Norms of the outer tiles of a tensor of tensors
 * (Tensor<Tensor<T>>::squared_norm, tensor/tensor.h:3768-3776): per tile,
 * the cells [tile_cells[t], tile_cells[t+1]) in order, each cell's
 * sequential sum of squares added to the tile's accumulator; empty cells
 * (cell_off < 0) skipped. */
void ora_tot_tile_norms_f64(const double* pool, const int64_t* tile_cells,
                            const int64_t* cell_off, const int64_t* cell_vol,
                            int64_t ntiles, float* out);
void ora_tot_tile_norms_f32(const float* pool, const int64_t* tile_cells,
                            const int64_t* cell_off, const int64_t* cell_vol,
                            int64_t ntiles, float* out);

/* SparseShape::scale_tile_norms<InverseVolume, Screen=true>
 * (sparse_shape.h:150-217, recursive_outer_product :106-131).
 * norms: row-major tile-norm tensor of the given rank, scaled in place.
 * ntiles[d]: number of tiles of mode d; sizes: concatenation over modes of the
 * tile extents (as float, sparse_shape.h:219-239).  Returns the zero count. */
int64_t ora_scale_tile_norms(float* norms, int rank, const int64_t* ntiles,
                             const float* sizes, float threshold);

/* SparseShape::gemm (sparse_shape.h:1589-1681) on fused M x K and K x N norm
 * matrices given in storage order (op_a/op_b: 0 = NoTrans, 1 = Trans).
 * k_sizes[K]: recursive outer product of the contracted modes' tile extents.
 * The k_sizes scaling walks contiguous runs of K as the reference does
 * (:1632-1647).  The float GEMM uses the canonical order of this project:
 * est = abs_factor * sum_{k ascending} (l*r) with every product and sum
 * rounded to float.  Entries < threshold are set to 0 and counted. */
int64_t ora_shape_gemm(const float* left, const float* right, int op_a,
                       int op_b, int64_t M, int64_t N, int64_t K,
                       const float* k_sizes, float abs_factor, float threshold,
                       float* result);

/* Row-major C = alpha*op(A)*op(B) + beta*C with TiledArray's leading
 * dimensions: lda = NoTrans ? k : m, ldb = NoTrans ? n : k, ldc = n.
 * tensor/kernels.h:146-158,226; math/blas.h:171-177. */
void ora_dgemm(int op_a, int op_b, int64_t m, int64_t n, int64_t k,
               double alpha, const double* A, int64_t lda, const double* B,
               int64_t ldb, double beta, double* C, int64_t ldc);
void ora_sgemm(int op_a, int op_b, int64_t m, int64_t n, int64_t k,
               float alpha, const float* A, int64_t lda, const float* B,
               int64_t ldb, float beta, float* C, int64_t ldc);

/* Tensor(other, perm) / detail::permute (tensor/tensor.h:644-692,
 * tensor/permute.h:121-207) in the "TO" convention of permutation.h:115-150:
 * result extent[perm[d]] = extent[d]; result[perm(idx)] = src[idx]. */
void ora_permute_f64(const double* src, double* dst, int rank,
                     const int64_t* extents, const int32_t* perm);
void ora_permute_f32(const float* src, float* dst, int rank,
                     const int64_t* extents, const int32_t* perm);

/* One-rank Summa + ReducePairTask + ContractReduce over fused block-sparse
 * matrices (dist_eval/contraction_eval.h:1094-1161, :1348-1384, :974-1001;
 * reduce_task.h:545-595; tile_op/contract_reduce.h:409-453).
 * Tile grids: A is Mt x Kt (storage Kt x Mt if op_a), B is Kt x Nt (storage
 * Nt x Kt if op_b).  a_off/b_off: element offset of each tile in its pool in
 * storage order, < 0 for a zero tile.  mext/next/kext: tile extents.
 * c_off (Mt x Nt): offsets of non-zero result tiles in C, < 0 for a zero
 * result tile.  For each non-zero C tile every k (ascending) with both
 * operands present contributes; the first contribution uses beta = 0, the
 * rest beta = 1.  pair_filter_threshold > 0 enables the per-pair rule that
 * the reference compiles out (contraction_eval.h:1386-1441): skip a pair when
 * a_norm*b_norm < pair_filter_threshold / Kt (needs a_norms, b_norms).
 * One OpenMP task per result tile, serial BLAS inside (tiledarray.cpp:112).
 * Returns the number of tile pairs multiplied. */
int64_t ora_contract_f64(int op_a, int op_b, int64_t Mt, int64_t Nt,
                         int64_t Kt, const int32_t* mext, const int32_t* next,
                         const int32_t* kext, const double* A,
                         const int64_t* a_off, const double* B,
                         const int64_t* b_off, double* C, const int64_t* c_off,
                         double alpha, const float* a_norms,
                         const float* b_norms, float pair_filter_threshold);
int64_t ora_contract_f32(int op_a, int op_b, int64_t Mt, int64_t Nt,
                         int64_t Kt, const int32_t* mext, const int32_t* next,
                         const int32_t* kext, const float* A,
                         const int64_t* a_off, const float* B,
                         const int64_t* b_off, float* C, const int64_t* c_off,
                         float alpha, const float* a_norms,
                         const float* b_norms, float pair_filter_threshold);

#ifdef __cplusplus
}
#endif
#endif
