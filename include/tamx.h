/*
 * tamx.h -- C ABI of the MI355X-native TiledArray contraction hot path.
 *
 * Plain C types only: device pointers, sizes and an opaque context.  A stream
 * argument is a hipStream_t passed as void* (NULL = the default stream).
 * Every entry point returns a tamx_status; tamx_last_error() returns a
 * thread-local message describing the last failure.  Entry points that
 * touch the device first clear HIP's pending last-error state, so a failure
 * already reported (for example an allocation returned as TAMX_ERR_OOM) is
 * not taken for a later launch's.
 *
 * Conventions shared with TiledArray:
 *   - tiles are dense row-major blocks (TA::Tensor storage, tensor/tensor.h:454-458);
 *   - GEMM is row-major C[m x n] = alpha*op(A)*op(B) + beta*C with
 *     lda = NoTrans ? k : m, ldb = NoTrans ? n : k, ldc = n
 *     (tensor/kernels.h:146-158,226; math/blas.h:171-177);
 *   - permutations use the "TO" convention of permutation.h:115-150:
 *     result extent[perm[d]] = source extent[d];
 *   - a tile is zero when its per-element (volume-scaled) Frobenius norm is
 *     below the shape threshold (sparse_shape.h:150-217, :495-499).
 *   The shape threshold is always an explicit argument, never a global.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef TAMX_H
#define TAMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status --- */

typedef enum tamx_status {
  TAMX_OK = 0,
  TAMX_ERR_BAD_ARG = 1,     /* replaces TA_ASSERT failures on arguments */
  TAMX_ERR_OOM = 2,         /* device allocation failed */
  TAMX_ERR_UNSUPPORTED = 3, /* shape/rank/precision this build does not handle */
  TAMX_ERR_DEVICE = 4,      /* a HIP runtime call failed */
  TAMX_ERR_CAPACITY = 5     /* a work list or output pool was too small */
} tamx_status;

typedef enum tamx_op { TAMX_NO_TRANS = 0, TAMX_TRANS = 1 } tamx_op;

/* Arithmetic of a grouped GEMM launch. */
typedef enum tamx_precision {
  TAMX_PREC_F64 = 0,    /* FP64 MFMA (v_mfma_f64_16x16x4) */
  TAMX_PREC_F32 = 1,    /* FP32 MFMA on FP32 data (tensor-of-tensor FP32) */
  TAMX_PREC_BF16X3 = 2, /* FP64 data split into 3 BF16 limbs */
  TAMX_PREC_BF16XS = 3, /* FP64 data split into nslices BF16 limbs */
  TAMX_PREC_F32X2 = 4   /* FP64 data split into 2 FP32 limbs */
} tamx_precision;

const char* tamx_status_string(int status);
const char* tamx_last_error(void);
/* Library version string ("tamx <semver> gfx950"). */
const char* tamx_version(void);

/* ----------------------------------------------------------- context --- */

typedef struct tamx_context tamx_context;

/* Replaces the device part of TA::initialize (tiledarray.cpp:32-57,106):
 * binds the calling thread's work to HIP device `device`. */
int tamx_context_create(int device, tamx_context** ctx);
/* Replaces TA::finalize's device teardown (tiledarray.cpp:167-186). */
int tamx_context_destroy(tamx_context* ctx);
/* Number of compute units of the context's device. */
int tamx_context_num_cus(const tamx_context* ctx);

/* --------------------------------------------------- tile-pair work --- */

/* One result tile of a grouped GEMM: C tile = alpha * sum_p op(A_p)op(B_p)
 * (+ C tile when beta = 1).  Mirrors one ReducePairTask
 * (reduce_task.h:961-1026): the first contribution of ContractReduce has
 * beta = 0 on an empty result, later ones beta = 1
 * (tile_op/contract_reduce.h:409-453). */
typedef struct tamx_gemm_task {
  int64_t c_off;      /* element offset of the result tile in C */
  int32_t m, n;       /* result tile extents (fused rows, fused columns) */
  int32_t ldc;        /* leading dimension of the stored tile */
  int32_t pair_begin; /* first contributing pair in the pair array */
  int32_t pair_count; /* number of contributing pairs */
  int32_t flags;      /* TAMX_TASK_* bits */
} tamx_gemm_task;

enum {
  TAMX_TASK_BETA1 = 1,       /* accumulate into the existing C tile */
  TAMX_TASK_STORE_TRANS = 2  /* store C^T (n x m, ldc >= m): fused result
                                permutation that is a matrix transpose of
                                the (left-outer, right-outer) result,
                                contract_reduce.h:370-378 */
};

/* One tile pair (A_ik, B_kj) contributing to a task. */
typedef struct tamx_gemm_pair {
  int64_t a_off;  /* element offset of A_ik (storage k x m when op_a = TRANS) */
  int64_t b_off;  /* element offset of B_kj (storage n x k when op_b = TRANS) */
  int32_t k;      /* contracted extent */
  int32_t lda, ldb;
  int32_t a_tile; /* ordinal of A_ik in its operand's tile table (used by the
                     split-precision path to find the tile's limbs) */
  int32_t b_tile; /* ordinal of B_kj */
  int32_t reserved;
} tamx_gemm_pair;

/* Device-resident work list of tasks and pairs with fixed capacity, so that
 * building and consuming it needs no host synchronisation and can be
 * captured in a hipGraph. */
typedef struct tamx_worklist tamx_worklist;

/* Layout of the int64 counters the work list keeps on the device. */
enum {
  TAMX_WL_NTASKS = 0,  /* number of tasks */
  TAMX_WL_NPAIRS = 1,  /* number of pairs */
  TAMX_WL_CELEMS = 2,  /* result elements allocated in C */
  TAMX_WL_STATUS = 3,  /* 0, or a TAMX_ERR_CAPACITY overflow mark */
  TAMX_WL_FLOPS = 4,   /* sum over pairs of 2*m*n*k */
  TAMX_WL_NCOUNTERS = 8
};

int tamx_worklist_create(tamx_context* ctx, int64_t task_capacity,
                         int64_t pair_capacity, int64_t grid_capacity,
                         tamx_worklist** wl);
int tamx_worklist_destroy(tamx_worklist* wl);
/* Device pointers of the work list, for callers that fill tasks and pairs
 * themselves (e.g. a ContractReduce-compatible op recording pairs).  Such a
 * caller writes counters[TAMX_WL_NTASKS] and counters[TAMX_WL_NPAIRS]. */
int tamx_worklist_buffers(tamx_worklist* wl, tamx_gemm_task** tasks,
                          tamx_gemm_pair** pairs, int64_t** counters);
/* Copies the counters to host memory (synchronises `stream`). */
int tamx_worklist_read_counters(tamx_worklist* wl, int64_t* host_counters,
                                void* stream);

/* ---------------------------------------------------------- planner --- */

#define TAMX_PLAN_MAX_RANK 8
#define TAMX_PLAN_NAME_LEN 64

/* Host-side plan of C(target) = A(left) * B(right): the decisions of
 * ContEngine::init_indices/init_struct (expressions/cont_engine.h:354-470)
 * and GEMMPermutationOptimizer (expressions/permopt.h:254-381). */
typedef struct tamx_contraction_plan {
  int32_t left_rank, right_rank, result_rank, inner_rank;
  int32_t left_permtype, right_permtype; /* 1 identity, 2 matrix_transpose,
                                            3 general (PermutationType) */
  int32_t op_a, op_b;                    /* to_cblas_op of the permtypes */
  int32_t left_perm[TAMX_PLAN_MAX_RANK];  /* TO permutation applied to each
                                             left tile (identity unless
                                             general) */
  int32_t right_perm[TAMX_PLAN_MAX_RANK];
  int32_t result_perm[TAMX_PLAN_MAX_RANK]; /* natural result order ->
                                              target, TO convention */
  int32_t result_perm_identity;
  int32_t result_perm_matrix_transpose;  /* result permute == transpose of
                                            the fused result tile */
  int32_t left_gemm_is_inner[TAMX_PLAN_MAX_RANK];  /* per mode of the left
                                                      GEMM layout */
  int32_t right_gemm_is_inner[TAMX_PLAN_MAX_RANK];
  char left_target[TAMX_PLAN_NAME_LEN];   /* left index order fed to GEMM */
  char right_target[TAMX_PLAN_NAME_LEN];
  char result_natural[TAMX_PLAN_NAME_LEN]; /* index order GEMM produces */
} tamx_contraction_plan;

/* Annotations are comma-separated index names ("a,b,i,j").
 * prefer_to_permute_left mirrors `left_type::leaves <= right_type::leaves`
 * (expressions/binary_engine.h:120-123), 1 for two leaf arrays. */
int tamx_plan_contraction(const char* left, const char* right,
                          const char* target, int prefer_to_permute_left,
                          tamx_contraction_plan* plan);

/* ------------------------------------------------------- screening --- */

/* Per-tile Frobenius norms, cast to float.  Replaces Tensor::norm()
 * (tensor/tensor.h:3768-3787) as called by make_array for sparse arrays
 * (conversions/make_array.h:152-205).  A plan fixes the tile table so a
 * launch needs no host data; tiles with offset < 0 get norm 0. */
typedef struct tamx_norm_plan tamx_norm_plan;
int tamx_norm_plan_create(tamx_context* ctx, const int64_t* host_offsets,
                          const int64_t* host_volumes, int64_t ntiles,
                          tamx_norm_plan** plan);
/* A plan over the outer tiles of a tensor of tensors: tile t is the cells
 * [host_tile_cells[t], host_tile_cells[t+1]) of the cell table, cell c at
 * element offset host_cell_off[c] (< 0: empty cell) with host_cell_vol[c]
 * elements, wherever they lie in the pool.  The norm of tile t is
 * sqrt(sum over its cells, in order, of each cell's sum of squares), as
 * Tensor<Tensor<T>>::squared_norm reduces its cells
 * (tensor/tensor.h:3768-3776), cast to float.  Used with
 * tamx_tile_norms_f64 / _f32 like any norm plan. */
int tamx_norm_plan_create_cells(tamx_context* ctx, int64_t ntiles,
                                const int64_t* host_tile_cells,
                                const int64_t* host_cell_off,
                                const int64_t* host_cell_vol,
                                tamx_norm_plan** plan);
int tamx_norm_plan_destroy(tamx_norm_plan* plan);
int tamx_tile_norms_f64(tamx_context* ctx, const tamx_norm_plan* plan,
                        const double* pool, float* norms, void* stream);
/* The same norms (bit-identical), and in the same pass each tile's max |a|
 * (0 for absent tiles) into maxabs[ntiles] (device double): the scale input
 * of tamx_split_f64_premax, so the limb split needs no separate max pass.
 * No reference counterpart (TA's split-free path needs no max). */
int tamx_tile_norms_max_f64(tamx_context* ctx, const tamx_norm_plan* plan,
                            const double* pool, float* norms, double* maxabs,
                            void* stream);
int tamx_tile_norms_f32(tamx_context* ctx, const tamx_norm_plan* plan,
                        const float* pool, float* norms, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the norms of tamx_tile_norms_f64, bit for bit, without
 * a plan.  offsets and volumes are the tile table in device memory (int64
 * [ntiles]; offset < 0: no tile, norm 0); pool_elems is the length of pool,
 * inside which the tiles lie.  The chunk table and partial sums live in a
 * static buffer of the library (about 10 MiB per GPU, held until the process
 * ends), so nothing is allocated.  Limits: ntiles <= TAMX_NOALLOC_MAX_TILES
 * and pool_elems / 65536 + ntiles <= TAMX_NOALLOC_MAX_CHUNKS, else
 * TAMX_ERR_CAPACITY before any launch.  A table the device finds invalid (a
 * present tile with negative volume, or tiles outside the pool) gives NaN
 * norms.  Every call on one GPU uses the same buffer: calls that overlap on
 * one GPU (other streams, other threads) race. */
#define TAMX_NOALLOC_MAX_TILES 262144
#define TAMX_NOALLOC_MAX_CHUNKS 1048576
int tamx_tile_norms_f64_noalloc(tamx_context* ctx, const int64_t* offsets,
                                const int64_t* volumes, int64_t ntiles,
                                const double* pool, int64_t pool_elems,
                                float* norms, void* stream);

/* Scales raw tile norms in place by inverse tile volume and screens them.
 * Replaces SparseShape's norm constructor (sparse_shape.h:335-349, scaling
 * :150-217).  norms: row-major tile grid of `rank` modes (nbatch grids back
 * to back); host_ntiles[d]: tiles of mode d; sizes: device floats, the tile
 * extents of mode 0, then mode 1, ... (per batch when sizes_per_batch != 0).
 * zero_count (device int64, may be NULL) receives the number of zeroed
 * tiles (added atomically). */
int tamx_shape_scale(tamx_context* ctx, float* norms, int rank,
                     const int64_t* host_ntiles, int64_t nbatch,
                     const float* sizes, int sizes_per_batch, float threshold,
                     int64_t* zero_count, void* stream);

/* Result shape estimate of a contraction.  Replaces SparseShape::gemm
 * (sparse_shape.h:1589-1681): left is the M x K tile-norm matrix in storage
 * order (K x M if op_a), right K x N (N x K if op_b), k_sizes the K volumes
 * of the contracted tiles.  est = abs_factor * sum_k l*r computed in float,
 * k ascending, products and sums rounded separately; est < threshold -> 0.
 * K = 0 is the outer product est = abs_factor * l_i * r_j.  nbatch
 * independent problems may be stacked (batch strides M*K, K*N, K, M*N; M
 * and N for the operands of an outer product).  zero_count as above. */
int tamx_shape_gemm_f32(tamx_context* ctx, const float* left,
                        const float* right, int op_a, int op_b, int64_t M,
                        int64_t N, int64_t K, int64_t nbatch,
                        const float* k_sizes, float abs_factor,
                        float threshold, float* result, int64_t* zero_count,
                        void* stream);

/* A block-sparse contraction on one rank, fused to matrix form.  Device
 * arrays; nbatch independent problems of the same tile-grid size may be
 * stacked (the batched product of tile_op/batched_contract_reduce.h). */
typedef struct tamx_contraction_desc {
  int32_t op_a, op_b;      /* tamx_op of the fused operands */
  int64_t Mt, Nt, Kt;      /* tile-grid extents of one problem */
  int64_t nbatch;          /* >= 1 */
  const int32_t* mext;     /* [nbatch*Mt] fused row extent of row tile i */
  const int32_t* next;     /* [nbatch*Nt] */
  const int32_t* kext;     /* [nbatch*Kt] */
  const int64_t* a_off;    /* [nbatch*Mt*Kt] tile offsets in A, < 0 absent
                              (grid in storage order: Kt x Mt if op_a) */
  const int64_t* b_off;    /* [nbatch*Kt*Nt] (Nt x Kt if op_b) */
  const float* a_norms;    /* optional scaled norms (same grid as a_off) */
  const float* b_norms;    /* optional */
  const float* c_norms;    /* [nbatch*Mt*Nt] result shape estimate */
  float a_threshold;       /* tile present iff offset >= 0 and norm >= thr */
  float b_threshold;
  float c_threshold;       /* result tile non-zero iff c_norm >= thr */
  float pair_threshold;    /* > 0: skip pairs with a_norm*b_norm <
                              pair_threshold/Kt, the per-pair rule the
                              reference compiles out
                              (contraction_eval.h:1386-1441); 0: off */
  int32_t store_trans;     /* 1: result tiles stored transposed */
  int64_t* c_off;          /* out [nbatch*Mt*Nt]: offsets in C, -1 zero */
  int64_t c_capacity;      /* elements available in C */
} tamx_contraction_desc;

/* Builds the work list of a contraction on the device: one task per non-zero
 * result tile, pairs for every k (ascending) with both operands non-zero.
 * Replaces the single-rank schedule of Summa (contraction_eval.h:1094-1161
 * initialize, :974-1001 iterate_sparse, :1348-1384 contract) and the
 * allocation of result tiles by ReducePairTask.  Result tiles are packed in
 * C in row-major tile order. */
int tamx_contraction_worklist(tamx_context* ctx,
                              const tamx_contraction_desc* desc,
                              tamx_worklist* wl, void* stream);

/* -------------------------------------------------------- permute --- */

/* Permutes many tiles of one rank with one permutation in one launch.
 * Replaces per-tile Tensor(other, perm) / detail::permute
 * (tensor/tensor.h:644-692, tensor/permute.h:121-207) reached from the
 * permute-on-read path (dist_eval/array_eval.h:330-365 ->
 * tile_interface/permute.h:74-134) and the result permute of
 * ContractReduce (contract_reduce.h:370-378); also replaces the per-call
 * LibreTT plan of the device build (external/librett.h:79-111).
 * perm: "TO" convention, result extent[perm[d]] = extent[d].
 * host_extents: ntiles x rank source extents (row-major tiles);
 * host_src_off / host_dst_off: element offsets of each tile in the source
 * and result pools.  The plan keeps the tile table on the device. */
typedef struct tamx_permute_plan tamx_permute_plan;
int tamx_permute_plan_create(tamx_context* ctx, int rank, const int32_t* perm,
                             int64_t ntiles, const int64_t* host_extents,
                             const int64_t* host_src_off,
                             const int64_t* host_dst_off,
                             tamx_permute_plan** plan);
int tamx_permute_plan_destroy(tamx_permute_plan* plan);
/* Elements moved by one execution (each read once, written once). */
int64_t tamx_permute_plan_volume(const tamx_permute_plan* plan);
int tamx_permute_f64(tamx_context* ctx, const tamx_permute_plan* plan,
                     const double* src, double* dst, void* stream);
int tamx_permute_f32(tamx_context* ctx, const tamx_permute_plan* plan,
                     const float* src, float* dst, void* stream);

/* ---------------------------------------------------- grouped GEMM --- */

/* All tasks of a work list in one persistent launch.  Replaces the per-pair
 * gemm customisation points (tile_op/tile_interface.h:805-830) that
 * ContractReduce calls (contract_reduce.h:446-451), reaching
 * Tensor::gemm -> detail::gemm -> math::blas::gemm (tensor.h:3131-3219,
 * kernels.h:90-231, blas.h:171-177). */
int tamx_grouped_gemm_f64(tamx_context* ctx, tamx_worklist* wl, int op_a,
                          int op_b, double alpha, const double* A,
                          const double* B, double* C, void* stream);
/* Launch configuration of a grouped GEMM: a BM x BN macro tile per
 * workgroup, K staged through LDS in chunks of BK, and at most wgs_per_cu
 * resident workgroups per compute unit (0 = as many as fit).  The compiled
 * shapes are enumerated by tamx_gemm_config_count/tamx_gemm_config_get;
 * the autotuner picks among them. */
typedef struct tamx_gemm_config {
  int32_t bm, bn, bk, wgs_per_cu;
} tamx_gemm_config;
int tamx_gemm_config_count(void);
int tamx_gemm_config_get(int index, tamx_gemm_config* cfg);
int tamx_grouped_gemm_f64_config(tamx_context* ctx, tamx_worklist* wl,
                                 int op_a, int op_b, double alpha,
                                 const double* A, const double* B, double* C,
                                 const tamx_gemm_config* cfg, void* stream);
int tamx_grouped_gemm_f32_config(tamx_context* ctx, tamx_worklist* wl,
                                 int op_a, int op_b, float alpha,
                                 const float* A, const float* B, float* C,
                                 const tamx_gemm_config* cfg, void* stream);
/* FP32 data, FP32 MFMA: tensor-of-tensor cells in single precision. */
int tamx_grouped_gemm_f32(tamx_context* ctx, tamx_worklist* wl, int op_a,
                          int op_b, float alpha, const float* A,
                          const float* B, float* C, void* stream);

/* Operands read through an index map.  An operand that needs a general
 * permutation into GEMM layout (permopt.h:254-381) is permuted per tile on
 * read in the reference (array_eval.h:330-365 -> Tensor(other, perm),
 * tensor.h:644-692) before gemm; here the GEMM reads its stored tiles
 * directly and no permuted copy exists.  A gather plan lists, for every
 * tile ordinal the work list's pairs carry (a_tile / b_tile), the extents
 * and stored strides of the tile's outer modes (rows of op(X) in GEMM
 * order, host_row_ext/str[t * nrow_modes + d]) and of its contracted modes
 * (host_k_ext/str[t * nk_modes + d]), 1 to 4 of each, the last fastest:
 * element (row, k) of tile t lies at the pair's tile offset + rowoff(row)
 * + koff(k).  The products of the extents must equal the tile's m (or n)
 * and k in the work list. */
typedef struct tamx_gather_plan tamx_gather_plan;
int tamx_gather_plan_create(tamx_context* ctx, int64_t ntiles, int nrow_modes,
                            int nk_modes, const int32_t* host_row_ext,
                            const int64_t* host_row_str, const int32_t* host_k_ext,
                            const int64_t* host_k_str, tamx_gather_plan** plan);
int tamx_gather_plan_destroy(tamx_gather_plan* plan);
/* The grouped GEMM with either operand (or both) gathered; a null plan
 * reads that operand as tamx_grouped_gemm_*_config does.  For a gathered
 * operand the pair's leading dimension is unused and its op only picks the
 * index the loads run along (A: NO_TRANS k, TRANS m; B: NO_TRANS n, TRANS
 * k): pass the one whose last mode is the stored stride-1 mode.  Results
 * equal those on a permuted copy bit for bit. */
int tamx_grouped_gemm_gather_f64(tamx_context* ctx, tamx_worklist* wl, int op_a,
                                 int op_b, double alpha, const double* A,
                                 const tamx_gather_plan* a_gather, const double* B,
                                 const tamx_gather_plan* b_gather, double* C,
                                 const tamx_gemm_config* cfg, void* stream);
int tamx_grouped_gemm_gather_f32(tamx_context* ctx, tamx_worklist* wl, int op_a,
                                 int op_b, float alpha, const float* A,
                                 const tamx_gather_plan* a_gather, const float* B,
                                 const tamx_gather_plan* b_gather, float* C,
                                 const tamx_gemm_config* cfg, void* stream);

/* Result tiles stored through an index map.  Replaces the permute of each
 * finished result tile into the target order (ContractReduce's post-op,
 * tile_op/contract_reduce.h:370-378, with the shape's permute,
 * sparse_shape.h:1687-1691): the GEMM writes every element of the natural
 * (left outer, right outer) result tile straight to its place in the
 * permuted tile.  c_scatter is a gather plan over the result tile grid, by
 * the grid position tamx_contraction_worklist gives each task (batch-major,
 * then row-major over Mt x Nt): its row modes are the left-outer modes,
 * its k modes the right-outer modes, each with its stride in the stored
 * target-order tile, so element (row, col) of result tile g lies at the
 * task's c_off + rowoff(row) + koff(col).  The work list must come from
 * tamx_contraction_worklist, which records each task's grid position; the
 * task's ldc and STORE_TRANS flag are unused.  Operand gathers as in
 * tamx_grouped_gemm_gather_*. */
int tamx_grouped_gemm_scatter_f64(tamx_context* ctx, tamx_worklist* wl, int op_a,
                                  int op_b, double alpha, const double* A,
                                  const tamx_gather_plan* a_gather, const double* B,
                                  const tamx_gather_plan* b_gather, double* C,
                                  const tamx_gather_plan* c_scatter,
                                  const tamx_gemm_config* cfg, void* stream);
int tamx_grouped_gemm_scatter_f32(tamx_context* ctx, tamx_worklist* wl, int op_a,
                                  int op_b, float alpha, const float* A,
                                  const tamx_gather_plan* a_gather, const float* B,
                                  const tamx_gather_plan* b_gather, float* C,
                                  const tamx_gather_plan* c_scatter,
                                  const tamx_gemm_config* cfg, void* stream);

/* ------------------------------------------------ split precision --- */

/* Splitting of FP64 operand tiles into nslices BF16 limbs with one
 * power-of-two scale per tile (sigma = exponent of the tile's max |a| plus
 * one).  The first limb takes 8 bits of the tile-scaled value, every later
 * limb 9:  a = 2^(sigma-8) * sum_s L_s + r,  |r| <= 2^(sigma - 9*nslices),
 * with L_s = l_s 2^-9s, l_s an integer, |l_s| <= 256, so every L_s is held
 * exactly in BF16.  The limbs are laid out for the
 * MFMA: the logical op(A) tile as m x K and op(B) as n x K, K contiguous,
 * rows padded to 16 and K to 32 with zeros.  No reference counterpart: the
 * reference multiplies in FP64 only (math/blas.h:171-177); this is the
 * split-precision path that stands beside it.
 * The plan lists the operand's tiles by the ordinals the work list's pairs
 * carry (a_tile / b_tile): host_offsets[t] (< 0: absent), and the stored
 * (row-major) extents host_rows[t] x host_cols[t].  is_right selects the
 * right operand layout, op its tamx_op. */
typedef struct tamx_split_plan tamx_split_plan;
int tamx_split_plan_create(tamx_context* ctx, int nslices, int is_right, int op,
                           int64_t ntiles, const int64_t* host_offsets,
                           const int32_t* host_rows, const int32_t* host_cols,
                           tamx_split_plan** plan);
/* A plan whose limb layout is read from the operand's stored tiles through
 * a general index map, so an operand that needs a general permutation
 * (permopt.h:254-381; permuted per tile on read, array_eval.h:330-365) is
 * split without a permuted copy: limb (row, k) of tile t reads source
 * element host_offsets[t] + rowoff(row) + koff(k), where rowoff is the
 * mixed-radix sum over nrow_modes modes (extents host_row_ext[t*nrow_modes
 * + d], source strides host_row_str[...], last mode fastest) and koff the
 * same over nk_modes modes (1 to 4 each).  The limb layout and GEMM use are
 * those of tamx_split_plan_create. */
int tamx_split_plan_create_general(tamx_context* ctx, int nslices, int64_t ntiles,
                                   const int64_t* host_offsets, int nrow_modes,
                                   int nk_modes, const int32_t* host_row_ext,
                                   const int64_t* host_row_str,
                                   const int32_t* host_k_ext,
                                   const int64_t* host_k_str, tamx_split_plan** plan);
/* Limb types.  TAMX_LIMB_BF16: the BF16 limbs above.  TAMX_LIMB_I8: INT8
 * limbs, a = 2^(sigma-7) * sum_s l_s 2^(-7 s) + r with |l_0| <= 127,
 * |l_s| <= 64 after it and |r| <= 2^(sigma - 7*nslices - 1) (sigma as above,
 * one more where the tile max would round to 128), multiplied on the INT8
 * MFMA with one exact INT32 accumulator per level s + t (an integer-slice
 * scheme in the manner of Ozaki et al.); one byte per limb, the logical
 * tile as m x K (n x K) with rows padded to 16 and K to 64, each plane
 * stored in blocks of 32 K: block b holds K 32b .. 32b + 31 of every row,
 * rows one after the other (limb (r, k) at ((k / 32) * m_pad + r) * 32 +
 * k % 32), so the chunk a GEMM loads for a macro tile is one contiguous run
 * per plane and block. */
#define TAMX_LIMB_BF16 0
#define TAMX_LIMB_I8 1
int tamx_split_plan_create_typed(tamx_context* ctx, int limb_type, int nslices,
                                 int is_right, int op, int64_t ntiles,
                                 const int64_t* host_offsets, const int32_t* host_rows,
                                 const int32_t* host_cols, tamx_split_plan** plan);
int tamx_split_plan_create_general_typed(tamx_context* ctx, int limb_type, int nslices,
                                         int64_t ntiles, const int64_t* host_offsets,
                                         int nrow_modes, int nk_modes,
                                         const int32_t* host_row_ext,
                                         const int64_t* host_row_str,
                                         const int32_t* host_k_ext,
                                         const int64_t* host_k_str, tamx_split_plan** plan);
int tamx_split_plan_limb_type(const tamx_split_plan* plan);
int tamx_split_plan_destroy(tamx_split_plan* plan);
/* Limb count of a plan (0 for a null plan). */
int tamx_split_plan_nslices(const tamx_split_plan* plan);
/* Limbs (BF16 or INT8 elements) the limb pool of a plan needs. */
int64_t tamx_split_plan_limb_elems(const tamx_split_plan* plan);
/* Copies the per-tile scale exponents to the host (synchronous). */
int tamx_split_plan_sigmas(const tamx_split_plan* plan, int32_t* host_sigma);
/* Writes the limbs of every tile of `pool` (device buffer of
 * tamx_split_plan_limb_elems BF16 or INT8 elements, by the plan's type). */
int tamx_split_f64(tamx_context* ctx, const tamx_split_plan* plan,
                   const double* pool, void* limbs, void* stream);
/* As tamx_split_f64 with each tile's max |a| given (device double, by the
 * plan's tile ordinals, e.g. from tamx_tile_norms_max_f64): one pass over
 * the operand instead of two. */
int tamx_split_f64_premax(tamx_context* ctx, const tamx_split_plan* plan,
                          const double* pool, const double* tile_maxabs,
                          void* limbs, void* stream);
/* Grouped GEMM on split operands: per pair the limb products with
 * s + t < nslices on BF16 MFMA with FP32 accumulation (the s + t = 0 and
 * s + t = 1 sums exact, the lower levels together), folded into FP64 every
 * 128 k (cross-limb sums, the 2^(sigma_A + sigma_B) tile scaling and the C
 * accumulation in FP64).  INT8 limbs: the products on INT8 MFMA, one exact
 * INT32 accumulator per level, folded at pair ends (at least every 4096 k).
 * Both plans must have the same limb type and count.  Same work list and
 * beta/transpose semantics as tamx_grouped_gemm_f64. */
int tamx_grouped_gemm_split(tamx_context* ctx, tamx_worklist* wl, double alpha,
                            const tamx_split_plan* a_plan, const void* a_limbs,
                            const tamx_split_plan* b_plan, const void* b_limbs,
                            double* C, void* stream);
/* The same with a macro tile: bm x bn of 64 x 64 (the default, any limb
 * count), 128 x 64 or 64 x 128 (2 to 4 BF16 or 3 to 4 INT8 limbs; one wave
 * per SIMD, a quarter fewer limb bytes loaded per result), 128 x 64 for 5 or
 * 6 INT8 limbs (eight waves of 32 x 32 in one workgroup per CU).  For INT8 limbs
 * at 64 x 64, bk picks the kernel: 32, each wave's 32 x 32 quadrant on the
 * INT8 MFMA 32x32x32 with 32-k chunks through a four-stage LDS ring (three
 * chunks in flight; chunks of k padding skipped); 64, the same blocks with
 * two 64-k buffers; 128 (4 or 5 limbs), two-wave workgroups with A limbs
 * through an LDS ring and B limbs straight into registers; otherwise the
 * 16x16x64 MFMA.  All give the same result bit for bit.  wgs_per_cu is
 * ignored. */
int tamx_grouped_gemm_split_config(tamx_context* ctx, tamx_worklist* wl, double alpha,
                                   const tamx_split_plan* a_plan, const void* a_limbs,
                                   const tamx_split_plan* b_plan, const void* b_limbs,
                                   double* C, const tamx_gemm_config* cfg, void* stream);
/* The same storing result tiles through a scatter plan (as
 * tamx_grouped_gemm_scatter_f64; null: as _config). */
int tamx_grouped_gemm_split_scatter(tamx_context* ctx, tamx_worklist* wl, double alpha,
                                    const tamx_split_plan* a_plan, const void* a_limbs,
                                    const tamx_split_plan* b_plan, const void* b_limbs,
                                    double* C, const tamx_gather_plan* c_scatter,
                                    const tamx_gemm_config* cfg, void* stream);

/* Two FP32 limbs: FP64 tiles split while staged (hi = fp32(a),
 * lo = fp32(a - hi)); hi*hi and hi*lo + lo*hi on FP32 MFMA with FP32
 * accumulation over 16-k chunks, folded into FP64.  FP64 operands in the
 * FP64 work list's layout; no plan needed. */
int tamx_grouped_gemm_f32x2(tamx_context* ctx, tamx_worklist* wl, int op_a,
                            int op_b, double alpha, const double* A,
                            const double* B, double* C, void* stream);
/* The same storing result tiles through a scatter plan (as
 * tamx_grouped_gemm_scatter_f64; null: as tamx_grouped_gemm_f32x2). */
int tamx_grouped_gemm_f32x2_scatter(tamx_context* ctx, tamx_worklist* wl, int op_a,
                                    int op_b, double alpha, const double* A,
                                    const double* B, double* C,
                                    const tamx_gather_plan* c_scatter, void* stream);

/* ------------------------------------------------------ autotuning --- */

/* Times every compiled tile configuration (tamx_gemm_config_get) of the
 * FP64 (precision TAMX_PREC_F64) or FP32 (TAMX_PREC_F32) grouped GEMM on a
 * built work list and returns the fastest in *best (and its mean launch
 * time, item planning included, in *best_ms).  The caller's C is not
 * modified: the trial launches write into a temporary buffer, seeded from C
 * when a task accumulates into it (TAMX_TASK_BETA1).  Each configuration
 * runs once untimed (unless TAMX_TUNE_NO_WARMUP) and then `reps` times.
 * Synchronises `stream` and allocates; not for use under graph capture.
 * Results are cached per device and shape class: precision, ops, and the
 * half-octave buckets of the mean task m and n, the mean pair k and the
 * task count.  A cached class returns without timing (*best_ms = the time
 * measured then) unless flags has TAMX_TUNE_FORCE.
 * No reference counterpart: the reference calls a vendor BLAS with its own
 * fixed blocking per tile pair (math/blas.h:171-177). */
#define TAMX_TUNE_FORCE 1
#define TAMX_TUNE_NO_WARMUP 2
int tamx_autotune_grouped_gemm(tamx_context* ctx, tamx_worklist* wl,
                               int precision, int op_a, int op_b,
                               const void* A, const void* B, const void* C,
                               int reps, int flags, tamx_gemm_config* best,
                               float* best_ms, void* stream);
/* The same for tamx_grouped_gemm_gather_f64 / _f32 (null plans allowed). */
int tamx_autotune_grouped_gemm_gather(tamx_context* ctx, tamx_worklist* wl,
                                      int precision, int op_a, int op_b,
                                      const void* A, const tamx_gather_plan* a_gather,
                                      const void* B, const tamx_gather_plan* b_gather,
                                      const void* C, int reps, int flags,
                                      tamx_gemm_config* best, float* best_ms,
                                      void* stream);
/* The same for the split GEMM (tamx_grouped_gemm_split_config): times its
 * macro tiles (64x64; up to four limbs also 128x64 and 64x128) and returns
 * the fastest; cached per device, limb count and shape class. */
int tamx_autotune_grouped_gemm_split(tamx_context* ctx, tamx_worklist* wl,
                                     const tamx_split_plan* a_plan, const void* a_limbs,
                                     const tamx_split_plan* b_plan, const void* b_limbs,
                                     const double* C, int reps, int flags,
                                     tamx_gemm_config* best, float* best_ms, void* stream);
/* Drops every cached tuning result of the process. */
int tamx_autotune_clear(void);

/* --------------------------------------------------------- graphs --- */

/* Capture of a launch sequence on a stream into a replayable graph.  Every
 * entry point above that neither synchronises nor allocates may be called
 * between tamx_graph_begin and tamx_graph_end (norms, shape scaling and
 * estimate, work-list build, permutes, limb split, grouped GEMMs); the
 * graph replays them with the arguments of capture time, reading the
 * current contents of the buffers they name.  `stream` must be a created
 * stream (not the null stream).  No reference counterpart: the reference
 * launches per tile pair (contract_reduce.h:446-451). */
typedef struct tamx_graph tamx_graph;
int tamx_graph_begin(tamx_context* ctx, void* stream);
int tamx_graph_end(tamx_context* ctx, void* stream, tamx_graph** graph);
int tamx_graph_launch(tamx_graph* graph, void* stream);
int tamx_graph_destroy(tamx_graph* graph);

#ifdef __cplusplus
}
#endif
#endif /* TAMX_H */
