/*
 * ta_oracle.c -- CPU restatement of TiledArray's one-rank contraction path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load
 * this library; the product path never does.
 *
 * Compiled with -ffp-contract=off so every float/double operation is rounded
 * on its own, in the order written here (the order each function cites).
 * Tile GEMMs go through MKL's cblas_dgemm/cblas_sgemm (dlopen'd, sequential
 * threading layer) when libmkl_rt is present, else through plain loops; the
 * reference likewise calls a vendor BLAS through BLAS++ (math/blas.h:163-197).
 */
#define _GNU_SOURCE
#include "ta_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- BLAS -- */

enum { CBLAS_ROW_MAJOR = 101, CBLAS_NO_TRANS = 111, CBLAS_TRANS = 112 };
typedef void (*cblas_dgemm_fn)(int, int, int, int, int, int, double,
                               const double*, int, const double*, int, double,
                               double*, int);
typedef void (*cblas_sgemm_fn)(int, int, int, int, int, int, float,
                               const float*, int, const float*, int, float,
                               float*, int);
typedef int (*mkl_set_layer_fn)(int);

static cblas_dgemm_fn g_dgemm = NULL;
static cblas_sgemm_fn g_sgemm = NULL;
static int g_blas_probed = 0;
static int g_threads = 0;

static void probe_blas(void) {
  if (g_blas_probed) return;
#pragma omp critical(ora_probe)
  {
    if (!g_blas_probed) {
      const char* env = getenv("TA_ORACLE_BLAS");
      if (!(env && strcmp(env, "loops") == 0)) {
        const char* cands[] = {env && env[0] && strcmp(env, "mkl") ? env : NULL,
                               "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so",
                               "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so.1"};
        for (unsigned c = 0; c < sizeof(cands) / sizeof(cands[0]); ++c) {
          if (!cands[c]) continue;
          void* h = dlopen(cands[c], RTLD_NOW | RTLD_LOCAL);
          if (!h) continue;
          /* MKL_THREADING_SEQUENTIAL = 1: serial BLAS inside each tile task,
           * as TA::initialize sets (tiledarray.cpp:111-112). */
          mkl_set_layer_fn layer =
              (mkl_set_layer_fn)dlsym(h, "MKL_Set_Threading_Layer");
          if (layer) layer(1);
          g_dgemm = (cblas_dgemm_fn)dlsym(h, "cblas_dgemm");
          g_sgemm = (cblas_sgemm_fn)dlsym(h, "cblas_sgemm");
          if (g_dgemm && g_sgemm) break;
          g_dgemm = NULL;
          g_sgemm = NULL;
        }
      }
      g_blas_probed = 1;
    }
  }
}

const char* ora_blas_name(void) {
  probe_blas();
  return g_dgemm ? "mkl" : "loops";
}

int ora_num_threads(void) {
  return g_threads > 0 ? g_threads : omp_get_max_threads();
}
void ora_set_num_threads(int n) { g_threads = n; }

#define DEFINE_GEMM(NAME, T, CBLAS)                                            \
  void NAME(int op_a, int op_b, int64_t m, int64_t n, int64_t k, T alpha,      \
            const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* C,    \
            int64_t ldc) {                                                     \
    probe_blas();                                                              \
    if (CBLAS && m > 0 && n > 0) {                                             \
      CBLAS(CBLAS_ROW_MAJOR, op_a ? CBLAS_TRANS : CBLAS_NO_TRANS,              \
            op_b ? CBLAS_TRANS : CBLAS_NO_TRANS, (int)m, (int)n, (int)k,       \
            alpha, A, (int)lda, B, (int)ldb, beta, C, (int)ldc);               \
      return;                                                                  \
    }                                                                          \
    for (int64_t i = 0; i < m; ++i)                                            \
      for (int64_t j = 0; j < n; ++j) {                                        \
        T s = 0;                                                               \
        for (int64_t p = 0; p < k; ++p) {                                      \
          const T a = op_a ? A[p * lda + i] : A[i * lda + p];                  \
          const T b = op_b ? B[j * ldb + p] : B[p * ldb + j];                  \
          s += a * b;                                                          \
        }                                                                      \
        C[i * ldc + j] =                                                       \
            beta == (T)0 ? alpha * s : alpha * s + beta * C[i * ldc + j];      \
      }                                                                        \
  }

DEFINE_GEMM(ora_dgemm, double, g_dgemm)
DEFINE_GEMM(ora_sgemm, float, g_sgemm)

/* --------------------------------------------------------------- norms -- */

/* squared_norm: reduce_op_serial walks the elements in order, adding
 * x*x to a double (vector_op.h:562-577); norm = sqrt(double) (tensor.h:3782),
 * stored into a float tensor by make_array (make_array.h:170-172). */
float ora_norm_f64(const double* x, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
  return (float)sqrt(s);
}

float ora_norm_f32(const float* x, int64_t n) {
  float s = 0.0f;
  for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
  return sqrtf(s);
}

void ora_tile_norms_f64(const double* pool, const int64_t* offsets,
                        const int64_t* volumes, int64_t ntiles, float* out) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(ora_num_threads())
  for (int64_t t = 0; t < ntiles; ++t)
    out[t] = offsets[t] < 0 ? 0.0f : ora_norm_f64(pool + offsets[t], volumes[t]);
}

void ora_tile_norms_f32(const float* pool, const int64_t* offsets,
                        const int64_t* volumes, int64_t ntiles, float* out) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(ora_num_threads())
  for (int64_t t = 0; t < ntiles; ++t)
    out[t] = offsets[t] < 0 ? 0.0f : ora_norm_f32(pool + offsets[t], volumes[t]);
}

/* Tensor<Tensor<T>>::squared_norm (tensor.h:3768-3776): the outer reduce
 * walks the cells in order and adds each cell's own squared norm (an inner
 * sequential reduce, as above) to a scalar_type accumulator; the outer
 * tile's norm is its sqrt, stored as float.  Tile t holds cells
 * [tile_cells[t], tile_cells[t+1]) of the cell table (offset < 0: empty). */
void ora_tot_tile_norms_f64(const double* pool, const int64_t* tile_cells,
                            const int64_t* cell_off, const int64_t* cell_vol,
                            int64_t ntiles, float* out) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(ora_num_threads())
  for (int64_t t = 0; t < ntiles; ++t) {
    double res = 0.0;
    for (int64_t c = tile_cells[t]; c < tile_cells[t + 1]; ++c) {
      if (cell_off[c] < 0) continue;
      double cs = 0.0;
      const double* x = pool + cell_off[c];
      for (int64_t i = 0; i < cell_vol[c]; ++i) cs += x[i] * x[i];
      res += cs;
    }
    out[t] = (float)sqrt(res);
  }
}

void ora_tot_tile_norms_f32(const float* pool, const int64_t* tile_cells,
                            const int64_t* cell_off, const int64_t* cell_vol,
                            int64_t ntiles, float* out) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(ora_num_threads())
  for (int64_t t = 0; t < ntiles; ++t) {
    float res = 0.0f;
    for (int64_t c = tile_cells[t]; c < tile_cells[t + 1]; ++c) {
      if (cell_off[c] < 0) continue;
      float cs = 0.0f;
      const float* x = pool + cell_off[c];
      for (int64_t i = 0; i < cell_vol[c]; ++i) cs += x[i] * x[i];
      res += cs;
    }
    out[t] = sqrtf(res);
  }
}

/* ----------------------------------------------------- shape scaling ---- */

/* recursive_outer_product (sparse_shape.h:106-131) with the inverse-volume op
 * (:177-181): dim == 1 -> {1/s}; otherwise split at middle = dim/2 + dim%2
 * and form the outer product left[i]*right[j]. */
static float* inv_outer(const float* const* sizes, const int64_t* n,
                        unsigned dim, int64_t* out_len) {
  if (dim == 1u) {
    float* r = (float*)malloc(sizeof(float) * (size_t)(n[0] > 0 ? n[0] : 1));
    for (int64_t i = 0; i < n[0]; ++i) r[i] = 1.0f / sizes[0][i];
    *out_len = n[0];
    return r;
  }
  const unsigned middle = (dim >> 1u) + (dim & 1u);
  int64_t ln = 0, rn = 0;
  float* l = inv_outer(sizes, n, middle, &ln);
  float* rr = inv_outer(sizes + middle, n + middle, dim - middle, &rn);
  float* r = (float*)malloc(sizeof(float) * (size_t)(ln * rn > 0 ? ln * rn : 1));
  for (int64_t i = 0; i < ln; ++i)
    for (int64_t j = 0; j < rn; ++j) r[i * rn + j] = l[i] * rr[j];
  free(l);
  free(rr);
  *out_len = ln * rn;
  return r;
}

int64_t ora_scale_tile_norms(float* norms, int rank, const int64_t* ntiles,
                             const float* sizes, float threshold) {
  const float* mode_sizes[16];
  int64_t off = 0;
  for (int d = 0; d < rank; ++d) {
    mode_sizes[d] = sizes + off;
    off += ntiles[d];
  }
  int64_t zeros = 0;
  if (rank == 1) {
    /* rank-1 branch divides (sparse_shape.h:160-173) */
    for (int64_t i = 0; i < ntiles[0]; ++i) {
      float v = norms[i];
      v /= mode_sizes[0][i];
      if (v < threshold) {
        v = 0.0f;
        ++zeros;
      }
      norms[i] = v;
    }
    return zeros;
  }
  const unsigned middle = ((unsigned)rank >> 1u) + ((unsigned)rank & 1u);
  int64_t ln = 0, rn = 0;
  float* l = inv_outer(mode_sizes, ntiles, middle, &ln);
  float* r = inv_outer(mode_sizes + middle, ntiles + middle, rank - middle, &rn);
  /* math::outer(...) with op: norm *= x * y; screen (sparse_shape.h:205-213) */
  for (int64_t i = 0; i < ln; ++i)
    for (int64_t j = 0; j < rn; ++j) {
      float v = norms[i * rn + j];
      v *= l[i] * r[j];
      if (v < threshold) {
        v = 0.0f;
        ++zeros;
      }
      norms[i * rn + j] = v;
    }
  free(l);
  free(r);
  return zeros;
}

/* ------------------------------------------------------ shape GEMM ---- */

int64_t ora_shape_gemm(const float* left, const float* right, int op_a,
                       int op_b, int64_t M, int64_t N, int64_t K,
                       const float* k_sizes, float abs_factor, float threshold,
                       float* result) {
  int64_t zeros = 0;
  if (K == 0) {
    /* outer-product branch (sparse_shape.h:1666-1678) */
    for (int64_t i = 0; i < M; ++i)
      for (int64_t j = 0; j < N; ++j) {
        float v = left[i] * right[j] * abs_factor;
        if (v < threshold) {
          v = 0.0f;
          ++zeros;
        }
        result[i * N + j] = v;
      }
    return zeros;
  }
  float* l = (float*)malloc(sizeof(float) * (size_t)(M * K));
  float* r = (float*)malloc(sizeof(float) * (size_t)(K * N));
  /* :1636-1640 -- runs of K contiguous elements times k_sizes */
  for (int64_t s = 0; s < M * K; ++s) l[s] = left[s] * k_sizes[s % K];
  /* :1642-1647 -- runs of N contiguous elements times k_sizes[run] */
  for (int64_t s = 0; s < K * N; ++s) r[s] = right[s] * k_sizes[s / N];
  /* :1649 left.gemm(right, abs_factor, helper); :1652-1660 screening */
  for (int64_t i = 0; i < M; ++i)
    for (int64_t j = 0; j < N; ++j) {
      float acc = 0.0f;
      for (int64_t p = 0; p < K; ++p) {
        const float a = op_a ? l[p * M + i] : l[i * K + p];
        const float b = op_b ? r[j * K + p] : r[p * N + j];
        const float prod = a * b;
        acc = acc + prod;
      }
      float v = abs_factor * acc;
      if (v < threshold) {
        v = 0.0f;
        ++zeros;
      }
      result[i * N + j] = v;
    }
  free(l);
  free(r);
  return zeros;
}

/* --------------------------------------------------------- permute ---- */

#define DEFINE_PERMUTE(NAME, T)                                                \
  void NAME(const T* src, T* dst, int rank, const int64_t* extents,            \
            const int32_t* perm) {                                             \
    int64_t rext[16], rstride[16], dstride[16], idx[16];                       \
    int64_t vol = 1;                                                           \
    for (int d = 0; d < rank; ++d) {                                           \
      rext[perm[d]] = extents[d];                                              \
      vol *= extents[d];                                                       \
    }                                                                          \
    int64_t s = 1;                                                             \
    for (int d = rank - 1; d >= 0; --d) {                                      \
      rstride[d] = s;                                                          \
      s *= rext[d];                                                            \
    }                                                                          \
    for (int d = 0; d < rank; ++d) {                                           \
      dstride[d] = rstride[perm[d]];                                           \
      idx[d] = 0;                                                              \
    }                                                                          \
    int64_t o = 0;                                                             \
    for (int64_t i = 0; i < vol; ++i) {                                        \
      dst[o] = src[i];                                                         \
      for (int d = rank - 1; d >= 0; --d) {                                    \
        o += dstride[d];                                                       \
        if (++idx[d] < extents[d]) break;                                      \
        o -= dstride[d] * extents[d];                                          \
        idx[d] = 0;                                                            \
      }                                                                        \
    }                                                                          \
  }

DEFINE_PERMUTE(ora_permute_f64, double)
DEFINE_PERMUTE(ora_permute_f32, float)

/* ----------------------------------------------------- contraction ---- */

#define DEFINE_CONTRACT(NAME, T, GEMM)                                         \
  int64_t NAME(int op_a, int op_b, int64_t Mt, int64_t Nt, int64_t Kt,         \
               const int32_t* mext, const int32_t* next,                       \
               const int32_t* kext, const T* A, const int64_t* a_off,          \
               const T* B, const int64_t* b_off, T* C, const int64_t* c_off,   \
               T alpha, const float* a_norms, const float* b_norms,            \
               float pair_filter_threshold) {                                  \
    int64_t npairs = 0;                                                        \
    const float thr_k =                                                        \
        pair_filter_threshold > 0.0f ? pair_filter_threshold / (float)Kt       \
                                     : 0.0f;                                   \
    _Pragma("omp parallel for schedule(dynamic, 1) reduction(+ : npairs) num_threads(ora_num_threads())") \
    for (int64_t ij = 0; ij < Mt * Nt; ++ij) {                                 \
      const int64_t i = ij / Nt, j = ij % Nt;                                  \
      if (c_off[ij] < 0) continue; /* zero result tile: no task */             \
      const int64_t m = mext[i], n = next[j];                                  \
      T* c = C + c_off[ij];                                                    \
      int first = 1;                                                           \
      for (int64_t k = 0; k < Kt; ++k) {                                       \
        const int64_t ai = op_a ? k * Mt + i : i * Kt + k;                     \
        const int64_t bi = op_b ? j * Kt + k : k * Nt + j;                     \
        if (a_off[ai] < 0 || b_off[bi] < 0) continue;                          \
        if (thr_k > 0.0f && a_norms[ai] * b_norms[bi] < thr_k) continue;       \
        const int64_t kk = kext[k];                                            \
        GEMM(op_a, op_b, m, n, kk, alpha, A + a_off[ai], op_a ? m : kk,        \
             B + b_off[bi], op_b ? kk : n, first ? (T)0 : (T)1, c, n);         \
        first = 0;                                                             \
        ++npairs;                                                              \
      }                                                                        \
      if (first) memset(c, 0, sizeof(T) * (size_t)(m * n));                    \
    }                                                                          \
    return npairs;                                                             \
  }

DEFINE_CONTRACT(ora_contract_f64, double, ora_dgemm)
DEFINE_CONTRACT(ora_contract_f32, float, ora_sgemm)
