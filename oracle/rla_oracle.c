/*
 * C restatement of RandLinearAlgebra.jl's hot loops — TEST INFRASTRUCTURE
 * ONLY (the checker and the CPU-baseline timing; never linked into the
 * product library).  OpenMP over independent columns; the per-column loops
 * follow the reference line by line:
 *   fwht            src/Compressors/helpers/fwht.jl:21-70
 *   srht_left       src/Compressors/srht.jl:220-286 (block loop collapsed:
 *                   each column is padded, transformed and gathered)
 *   sparse_left     the SparseArrays `_spmatmul!` loop behind
 *                   src/Compressors/sparse_sign.jl:347 (CSC scatter order)
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int sbit(const uint64_t* s, int64_t i) { return (int)((s[i >> 6] >> (i & 63)) & 1u); }

#define FWHT_BODY(T)                                                        \
  {                                                                         \
    int64_t h = 1;                                                          \
    for (int64_t i = 0; i < n; i += 2) {                                    \
      double sc0 = (!signs || sbit(signs, i)) ? scaling : -scaling;         \
      double sc1 = (!signs || sbit(signs, i + 1)) ? scaling : -scaling;     \
      double z = (double)x[i] * sc0;                                        \
      double y = (double)x[i + 1] * sc1;                                    \
      x[i] = (T)(z + y);                                                    \
      x[i + 1] = (T)(z - y);                                                \
    }                                                                       \
    for (h = 2; h < n; h <<= 1) {                                           \
      int64_t inc = h << 1;                                                 \
      for (int64_t i = 0; i < n; i += inc)                                  \
        for (int64_t j = i; j < i + h; ++j) {                               \
          T z = x[j], y = x[j + h];                                         \
          x[j] = z + y;                                                     \
          x[j + h] = z - y;                                                 \
        }                                                                   \
    }                                                                       \
  }

static void fwht_f32(float* x, int64_t n, const uint64_t* signs, double scaling) FWHT_BODY(float)
static void fwht_f64(double* x, int64_t n, const uint64_t* signs, double scaling) FWHT_BODY(double)

void oracle_fwht_f32(float* X, int64_t n, int64_t ncols, int64_t ldx, const uint64_t* signs,
                     double scaling, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t k = 0; k < ncols; ++k) fwht_f32(X + k * ldx, n, signs, scaling);
}

void oracle_fwht_f64(double* X, int64_t n, int64_t ncols, int64_t ldx, const uint64_t* signs,
                     double scaling, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t k = 0; k < ncols; ++k) fwht_f64(X + k * ldx, n, signs, scaling);
}

/* C[:, k] = beta*C[:, k] + alpha*(scale*H*D*pad(A[:, k]))[op] (Float64 broadcast). */
void oracle_srht_left_f32(const float* A, int64_t m, int64_t ncols, int64_t lda, int64_t n_pad,
                          int64_t s, const int32_t* op, const uint64_t* signs, double scale,
                          float* C, int64_t ldc, double alpha, double beta, int nthreads) {
#pragma omp parallel num_threads(nthreads)
  {
    float* pad = (float*)malloc(sizeof(float) * (size_t)n_pad);
#pragma omp for schedule(static)
    for (int64_t k = 0; k < ncols; ++k) {
      memset(pad, 0, sizeof(float) * (size_t)n_pad);
      memcpy(pad, A + k * lda, sizeof(float) * (size_t)m);
      fwht_f32(pad, n_pad, signs, scale);
      float* c = C + k * ldc;
      for (int64_t r = 0; r < s; ++r) c[r] = (float)(beta * (double)c[r] + alpha * (double)pad[op[r]]);
    }
    free(pad);
  }
}

/* C = alpha*M*A + beta*C, M s x m CSC (colptr, rowidx, val), A m x ncols. */
void oracle_sparse_left_f64(const int32_t* colptr, const int32_t* rowidx, const double* val,
                            int64_t s, int64_t m, const double* A, int64_t ncols, int64_t lda,
                            double* C, int64_t ldc, double alpha, double beta, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t k = 0; k < ncols; ++k) {
    double* c = C + k * ldc;
    if (beta == 0.0) memset(c, 0, sizeof(double) * (size_t)s);
    else if (beta != 1.0) for (int64_t i = 0; i < s; ++i) c[i] *= beta;
    const double* a = A + k * lda;
    for (int64_t col = 0; col < m; ++col) {
      double axj = a[col] * alpha;
      for (int32_t p = colptr[col]; p < colptr[col + 1]; ++p) c[rowidx[p]] += val[p] * axj;
    }
  }
}

int oracle_max_threads(void) { return omp_get_max_threads(); }
