// Sparse compressor application (SparseSign, CountSketch, FJLT's K).
//
// Every orientation of the reference's sparse mul! is C = alpha*op(M)*X +
// beta*C on (possibly transposed) strided views, and op(M)*X is a row gather:
// output row i sums v_p * X[idx_p, :] over the nonzeros p of row i of op(M),
// i.e. over the CSR of M (op = M) or the CSC of M (op = M').  Each output
// entry is owned by one thread and accumulated sequentially in ascending
// index order, as the SparseArrays loop does, so results are deterministic
// and independent of the launch shape.  No floating-point atomics.
//
// Arithmetic follows Julia's promotion in that loop: alpha and beta are
// Float64, so `alpha*x` and `acc + v*(alpha*x)` are evaluated in double and
// the sum is rounded back to the element type after every step
// (C[rv[j],k] += nzv[j]*αxj with αxj = B[col,k]*α).
//
// Two gather kernels:
//  * rowgather_lds: X column-major (unit row stride).  A workgroup stages a
//    block of rows of KT adjacent columns of X in LDS (coalesced reads,
//    interleaved [row][KT]) and every thread gathers its output rows from
//    LDS.  Rows of X beyond one LDS block are handled block by block, with
//    the partial sums carried in C (the workgroup owns its C tile).  The
//    index/value reads come either straight from the CSR/CSC arrays (each
//    thread walks its own row: a wavefront's load touches up to 64 cache
//    lines) or from a row-interleaved copy built once per recipe (entry q of
//    row i at q * rows + i: a wavefront's load is 2-4 consecutive lines).
//  * rowgather_direct: X row-major view (unit column stride).  One wavefront
//    per output row, lanes over the right-hand sides, gathered rows read
//    straight from global memory (each gathered row is contiguous).
#pragma clang fp contract(off)

#include <algorithm>

#include <mutex>
#include <vector>

#include "rla_internal.hpp"

namespace rla {

constexpr int kGatherThreads = 512;
constexpr size_t kGatherLds = 64 * 1024;

// ---- CSC -> CSR ----------------------------------------------------------

__global__ void zero_i32_kernel(int32_t* p, int64_t n) {
  int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void count_rows_kernel(const int32_t* __restrict__ rowidx, int64_t nnz,
                                  int32_t* __restrict__ cnt /* rowptr + 1 */) {
  int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (p < nnz) atomicAdd(&cnt[rowidx[p]], 1);
}

// In-place inclusive scan of a[0..n) by one workgroup of 1024 threads:
// each thread scans a contiguous segment, segment totals are scanned in LDS.
__global__ __launch_bounds__(1024) void scan_inclusive_kernel(int32_t* a, int64_t n) {
  __shared__ int32_t part[1024];
  const int T = blockDim.x;
  int64_t seg = (n + T - 1) / T;
  int64_t b = threadIdx.x * seg, e = b + seg < n ? b + seg : n;
  int32_t s = 0;
  for (int64_t i = b; i < e; ++i) s += a[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    int32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
  for (int64_t i = b; i < e; ++i) {
    run += a[i];
    a[i] = run;
  }
}

int scan_inclusive_i32(rla_ctx* ctx, int32_t* a, int64_t n) {
  if (n <= 0) return RLA_OK;
  scan_inclusive_kernel<<<1, 1024, 0, ctx->stream>>>(a, n);
  RLA_LAUNCH_CHECK("scan_inclusive");
  return RLA_OK;
}

template <class T>
__global__ void fill_csr_kernel(int64_t ncols, const int32_t* __restrict__ colptr,
                                const int32_t* __restrict__ rowidx, const T* __restrict__ val,
                                int32_t* __restrict__ cursor, int32_t* __restrict__ colidx,
                                T* __restrict__ val_t) {
  int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (j >= ncols) return;
  for (int32_t p = colptr[j]; p < colptr[j + 1]; ++p) {
    int32_t pos = atomicAdd(&cursor[rowidx[p]], 1);
    colidx[pos] = static_cast<int32_t>(j);
    val_t[pos] = val[p];
  }
}

// Sorts each CSR row by column index (shell sort with Ciura's gaps; keys are
// distinct, so the result does not depend on the fill order above).
template <class T>
__global__ void sort_rows_kernel(int64_t nrows, const int32_t* __restrict__ rowptr,
                                 int32_t* __restrict__ colidx, T* __restrict__ val_t) {
  int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= nrows) return;
  int32_t b = rowptr[i], n = rowptr[i + 1] - b;
  int32_t* c = colidx + b;
  T* v = val_t + b;
  const int gaps[8] = {701, 301, 132, 57, 23, 10, 4, 1};
  for (int g = 0; g < 8; ++g) {
    int gap = gaps[g];
    for (int32_t q = gap; q < n; ++q) {
      int32_t kc = c[q];
      T kv = v[q];
      int32_t r = q;
      while (r >= gap && c[r - gap] > kc) {
        c[r] = c[r - gap];
        v[r] = v[r - gap];
        r -= gap;
      }
      c[r] = kc;
      v[r] = kv;
    }
  }
}

template <class T>
static int sparse_transpose(rla_ctx* ctx, int64_t nrows, int64_t ncols, int64_t nnz,
                            const int32_t* colptr, const int32_t* rowidx, const T* val,
                            int32_t* rowptr, int32_t* colidx, T* val_t) {
  RLA_CHECK_CTX(ctx);
  if (nrows < 0 || ncols < 0 || nnz < 0) return fail(RLA_EARG, "negative sparse dimensions");
  if (nrows >= INT32_MAX || nnz >= INT32_MAX)
    return fail(RLA_ENOTSUP, "sparse operator larger than int32 indexing");
  int rc = use_device(ctx);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  zero_i32_kernel<<<dim3(unsigned(ceil_div(nrows + 1, 256))), 256, 0, st>>>(rowptr, nrows + 1);
  if (nnz > 0)
    count_rows_kernel<<<dim3(unsigned(ceil_div(nnz, 256))), 256, 0, st>>>(rowidx, nnz,
                                                                          rowptr + 1);
  scan_inclusive_kernel<<<1, 1024, 0, st>>>(rowptr + 1, nrows);
  RLA_LAUNCH_CHECK("sparse_transpose/scan");
  if (nnz == 0 || ncols == 0) return RLA_OK;
  void* ws = nullptr;
  rc = workspace(ctx, static_cast<size_t>(nrows + 1) * sizeof(int32_t), &ws);
  if (rc) return rc;
  int32_t* cursor = static_cast<int32_t*>(ws);
  RLA_HIP(hipMemcpyAsync(cursor, rowptr, static_cast<size_t>(nrows) * sizeof(int32_t),
                         hipMemcpyDeviceToDevice, st));
  fill_csr_kernel<T><<<dim3(unsigned(ceil_div(ncols, 256))), 256, 0, st>>>(
      ncols, colptr, rowidx, val, cursor, colidx, val_t);
  sort_rows_kernel<T><<<dim3(unsigned(ceil_div(nrows, 256))), 256, 0, st>>>(nrows, rowptr,
                                                                           colidx, val_t);
  RLA_LAUNCH_CHECK("sparse_transpose/fill");
  return RLA_OK;
}

// ---- gather kernels --------------------------------------------------------

// first q in [lo, hi) with a[q * es] >= key (a sorted at stride es)
__device__ inline int32_t lower_bound_i32(const int32_t* __restrict__ a, int64_t es, int32_t lo,
                                          int32_t hi, int64_t key) {
  while (lo < hi) {
    int32_t mid = (lo + hi) >> 1;
    if (a[mid * es] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// ELL = false: the entries of row i are idx/val[ptr[i] .. ptr[i+1]) (CSR or
// CSC).  ELL = true: entry q of row i is at idx/val[q * ell_ld + i] (the
// row-interleaved copy from rla_sparse_ell_build), so the lanes of a
// wavefront, which own consecutive rows, read consecutive words; ptr still
// gives the row lengths.
template <class T, int KT, bool ELL>
__global__ __launch_bounds__(kGatherThreads) void rowgather_lds_kernel(
    int64_t out_rows, int64_t in_rows, int64_t n_rhs, const int32_t* __restrict__ ptr,
    const int32_t* __restrict__ idx, const T* __restrict__ val, int64_t ell_ld,
    const T* __restrict__ X, int64_t x_cs, T* __restrict__ C, int64_t c_rs, int64_t c_cs,
    double alpha, double beta, int64_t rb_rows, int nk_tiles, int rsplit,
    int64_t rows_per_part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  // Block -> (column tile, output-row part).  The parts of one column tile
  // are 8 block ids apart, so they land on the same XCD (blocks are dealt
  // round-robin over the 8 XCDs) and share that XCD's L2 for the X tile.
  const int b = blockIdx.x;
  const int grp = b / (8 * rsplit), within = b % (8 * rsplit);
  const int ktile = grp * 8 + (within % 8);
  const int part = within / 8;
  if (ktile >= nk_tiles) return;
  const int64_t k0 = int64_t(ktile) * KT;
  const int kt = static_cast<int>(n_rhs - k0 < KT ? n_rhs - k0 : KT);
  const int64_t r_begin = part * rows_per_part;
  const int64_t r_end = r_begin + rows_per_part < out_rows ? r_begin + rows_per_part : out_rows;
  const int64_t nblocks = in_rows > 0 ? (in_rows + rb_rows - 1) / rb_rows : 1;

  for (int64_t blk = 0; blk < nblocks; ++blk) {
    const int64_t j0 = blk * rb_rows;
    const int64_t jn = in_rows - j0 < rb_rows ? in_rows - j0 : rb_rows;
    __syncthreads();
    // stage X[j0:j0+jn, k0:k0+kt] as lds[j*KT + kk], 8 loads in flight per thread
    for (int kk = 0; kk < KT; ++kk) {
      const T* xc = X + (k0 + kk) * x_cs + j0;
      for (int64_t jb = 0; jb < jn; jb += 8 * int64_t(blockDim.x)) {
        T tmp[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t j = jb + u * int64_t(blockDim.x) + threadIdx.x;
          tmp[u] = (kk < kt && j < jn) ? xc[j] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t j = jb + u * int64_t(blockDim.x) + threadIdx.x;
          if (j < jn) lds[j * KT + kk] = tmp[u];
        }
      }
    }
    __syncthreads();
    for (int64_t i = r_begin + threadIdx.x; i < r_end; i += blockDim.x) {
      const int32_t row_p = ptr[i], row_n = ptr[i + 1] - row_p;
      const int64_t es = ELL ? ell_ld : 1;
      const int32_t* __restrict__ ix = ELL ? idx + i : idx + row_p;
      const T* __restrict__ vx = ELL ? val + i : val + row_p;
      int32_t p = 0, pe = row_n;
      if (nblocks > 1) {
        p = lower_bound_i32(ix, es, 0, row_n, j0);
        pe = lower_bound_i32(ix, es, p, row_n, j0 + jn);
      }
      T acc[KT];
#pragma unroll
      for (int kk = 0; kk < KT; ++kk) {
        if (kk < kt) {
          T* cp = C + i * c_rs + (k0 + kk) * c_cs;
          if (blk > 0) acc[kk] = *cp;
          else acc[kk] = (beta == 0.0) ? T(0) : static_cast<T>(beta * double(*cp));
        } else {
          acc[kk] = T(0);
        }
      }
      // 8 nonzeros per batch: their index/value loads and LDS reads are issued
      // together; the sums stay in ascending order.
      for (; p + 8 <= pe; p += 8) {
        int32_t jj[8];
        T vv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          jj[u] = ix[(p + u) * es];
          vv[u] = vx[(p + u) * es];
        }
        T xv[8][KT];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int kk = 0; kk < KT; ++kk) xv[u][kk] = lds[(jj[u] - j0) * KT + kk];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int kk = 0; kk < KT; ++kk)
            acc[kk] = static_cast<T>(double(acc[kk]) + double(vv[u]) * (double(xv[u][kk]) * alpha));
      }
      for (; p < pe; ++p) {
        const int64_t j = ix[p * es] - j0;
        const T v = vx[p * es];
#pragma unroll
        for (int kk = 0; kk < KT; ++kk)
          acc[kk] = static_cast<T>(double(acc[kk]) + double(v) * (double(lds[j * KT + kk]) * alpha));
      }
#pragma unroll
      for (int kk = 0; kk < KT; ++kk)
        if (kk < kt) C[i * c_rs + (k0 + kk) * c_cs] = acc[kk];
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void rowgather_direct_kernel(
    int64_t out_rows, int64_t n_rhs, const int32_t* __restrict__ ptr,
    const int32_t* __restrict__ idx, const T* __restrict__ val, const T* __restrict__ X,
    int64_t x_rs, int64_t x_cs, T* __restrict__ C, int64_t c_rs, int64_t c_cs, double alpha,
    double beta) {
  const int lane = threadIdx.x & 63;
  const int64_t i = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (i >= out_rows) return;
  const int32_t p0 = ptr[i], pe = ptr[i + 1];
  for (int64_t k = int64_t(blockIdx.y) * 64 + lane; k < n_rhs; k += int64_t(gridDim.y) * 64) {
    T* cp = C + i * c_rs + k * c_cs;
    T acc = (beta == 0.0) ? T(0) : static_cast<T>(beta * double(*cp));
    int32_t p = p0;
    for (; p + 8 <= pe; p += 8) {
      T xv[8], vv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        vv[u] = val[p + u];
        xv[u] = X[idx[p + u] * x_rs + k * x_cs];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        acc = static_cast<T>(double(acc) + double(vv[u]) * (double(xv[u]) * alpha));
    }
    for (; p < pe; ++p)
      acc = static_cast<T>(double(acc) + double(val[p]) * (double(X[idx[p] * x_rs + k * x_cs]) * alpha));
    *cp = acc;
  }
}

template <class T, int KT, bool ELL>
static int launch_rowgather_lds(rla_ctx* ctx, int64_t out_rows, int64_t in_rows, int64_t n_rhs,
                                const int32_t* ptr, const int32_t* idx, const T* val,
                                const T* X, int64_t x_cs, T* C, int64_t c_rs, int64_t c_cs,
                                double alpha, double beta) {
  int64_t rb_rows = static_cast<int64_t>(kGatherLds / (KT * sizeof(T)));
  if (in_rows < rb_rows) rb_rows = in_rows > 0 ? in_rows : 1;
  size_t lds = static_cast<size_t>(rb_rows) * KT * sizeof(T);
  int nk = static_cast<int>(ceil_div(n_rhs, KT));
  // Split output rows so that the launch has about two workgroups per CU.
  int64_t max_parts = ceil_div(out_rows, kGatherThreads);
  int64_t want = ceil_div(512, nk);
  int rsplit = static_cast<int>(std::max<int64_t>(1, std::min(max_parts, want)));
  int64_t rows_per_part = ceil_div(out_rows, rsplit);
  int64_t groups = ceil_div(nk, 8);
  int64_t blocks = groups * 8 * rsplit;
  rowgather_lds_kernel<T, KT, ELL><<<dim3(static_cast<unsigned>(blocks)), kGatherThreads, lds,
                                     ctx->stream>>>(out_rows, in_rows, n_rhs, ptr, idx, val,
                                                    out_rows, X, x_cs, C, c_rs, c_cs, alpha,
                                                    beta, rb_rows, nk, rsplit, rows_per_part);
  RLA_LAUNCH_CHECK("rowgather_lds");
  return RLA_OK;
}

// X column-major: wider column tiles amortise the index reads; narrower ones
// keep enough workgroups in flight.
template <class T, bool ELL>
static int rowgather_lds(rla_ctx* ctx, int64_t out_rows, int64_t in_rows, int64_t n_rhs,
                         const int32_t* ptr, const int32_t* idx, const T* v, const T* X,
                         int64_t x_cs, T* C, int64_t c_rs, int64_t c_cs, double a, double bt) {
  const int64_t kt_target = n_rhs / 512;
  if (sizeof(T) == 4 && kt_target >= 4)
    return launch_rowgather_lds<T, 4, ELL>(ctx, out_rows, in_rows, n_rhs, ptr, idx, v, X, x_cs,
                                           C, c_rs, c_cs, a, bt);
  if (kt_target >= 2)
    return launch_rowgather_lds<T, 2, ELL>(ctx, out_rows, in_rows, n_rhs, ptr, idx, v, X, x_cs,
                                           C, c_rs, c_cs, a, bt);
  return launch_rowgather_lds<T, 1, ELL>(ctx, out_rows, in_rows, n_rhs, ptr, idx, v, X, x_cs, C,
                                         c_rs, c_cs, a, bt);
}

template <class T>
static int sparse_apply(rla_ctx* ctx, int trans_m, int64_t nrows, int64_t ncols,
                        const int32_t* colptr, const int32_t* rowidx, const T* val,
                        const int32_t* rowptr, const int32_t* colidx, const T* val_t,
                        int64_t n_rhs, const T* X, int64_t x_rs, int64_t x_cs, T* C,
                        int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  RLA_CHECK_CTX(ctx);
  if (nrows < 0 || ncols < 0 || n_rhs < 0) return fail(RLA_EARG, "negative dimensions");
  const int64_t out_rows = trans_m ? ncols : nrows;
  const int64_t in_rows = trans_m ? nrows : ncols;
  const int32_t* ptr = trans_m ? colptr : rowptr;
  const int32_t* idx = trans_m ? rowidx : colidx;
  const T* v = trans_m ? val : val_t;
  if (out_rows == 0 || n_rhs == 0) return RLA_OK;
  if (!ptr || !idx || !v || !C || (in_rows > 0 && !X))
    return fail(RLA_EARG, trans_m ? "sparse apply with op = M' needs the CSC arrays"
                                  : "sparse apply with op = M needs the CSR arrays "
                                    "(rla_sparse_transpose)");
  int rc = use_device(ctx);
  if (rc) return rc;
  const double a = alpha, bt = beta;
  if (x_rs == 1 && in_rows > 0)
    return rowgather_lds<T, false>(ctx, out_rows, in_rows, n_rhs, ptr, idx, v, X, x_cs, C, c_rs,
                                   c_cs, a, bt);
  int64_t gy = std::min<int64_t>(ceil_div(n_rhs, 64), 1024);
  rowgather_direct_kernel<T><<<dim3(static_cast<unsigned>(ceil_div(out_rows, 4)),
                                    static_cast<unsigned>(gy)),
                               256, 0, ctx->stream>>>(out_rows, n_rhs, ptr, idx, v, X, x_rs,
                                                      x_cs, C, c_rs, c_cs, a, bt);
  RLA_LAUNCH_CHECK("rowgather_direct");
  return RLA_OK;
}


// ---- row-interleaved gather structure ----------------------------------------

__global__ void ell_width_kernel(int64_t rows, const int32_t* __restrict__ ptr,
                                 int32_t* __restrict__ width) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < rows) atomicMax(width, ptr[i + 1] - ptr[i]);
}

__global__ void set_i32_kernel(int32_t* p, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ell[q * rows + i] = entry q of row i for q < its length, index 0 and value
// 0 past it (the gathers stop at the row length and never read those).
template <class T>
__global__ void ell_build_kernel(int64_t rows, int32_t width, const int32_t* __restrict__ ptr,
                                 const int32_t* __restrict__ idx, const T* __restrict__ val,
                                 int32_t* __restrict__ ell_idx, T* __restrict__ ell_val) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= rows) return;
  const int32_t p = ptr[i], n = ptr[i + 1] - p;
  for (int32_t q = 0; q < width; ++q) {
    ell_idx[int64_t(q) * rows + i] = q < n ? idx[p + q] : 0;
    ell_val[int64_t(q) * rows + i] = q < n ? val[p + q] : T(0);
  }
}

static int ell_width(rla_ctx* ctx, int64_t rows, const int32_t* ptr, int32_t* width) {
  RLA_CHECK_CTX(ctx);
  if (rows < 0) return fail(RLA_EARG, "negative dimensions");
  if (!width) return fail(RLA_EARG, "width is NULL");
  *width = 0;
  if (rows == 0) return RLA_OK;
  if (!ptr) return fail(RLA_EARG, "ptr is NULL");
  int rc = use_device(ctx);
  if (rc) return rc;
  void* ws = nullptr;
  rc = workspace(ctx, sizeof(int32_t), &ws);
  if (rc) return rc;
  int32_t* w = static_cast<int32_t*>(ws);
  set_i32_kernel<<<1, 64, 0, ctx->stream>>>(w, 0);
  ell_width_kernel<<<dim3(unsigned(ceil_div(rows, 256))), 256, 0, ctx->stream>>>(rows, ptr, w);
  RLA_LAUNCH_CHECK("ell_width");
  RLA_HIP(hipMemcpyAsync(width, w, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  return RLA_OK;
}

template <class T>
static int ell_build(rla_ctx* ctx, int64_t rows, int32_t width, const int32_t* ptr,
                     const int32_t* idx, const T* val, int32_t* ell_idx, T* ell_val) {
  RLA_CHECK_CTX(ctx);
  if (rows < 0 || width < 0) return fail(RLA_EARG, "negative dimensions");
  if (rows == 0 || width == 0) return RLA_OK;
  if (!ptr || !idx || !val || !ell_idx || !ell_val) return fail(RLA_EARG, "ELL build: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  ell_build_kernel<T><<<dim3(unsigned(ceil_div(rows, 256))), 256, 0, ctx->stream>>>(
      rows, width, ptr, idx, val, ell_idx, ell_val);
  RLA_LAUNCH_CHECK("ell_build");
  return RLA_OK;
}

template <class T>
static int sparse_apply_ell(rla_ctx* ctx, int64_t out_rows, int64_t in_rows, const int32_t* ptr,
                            const int32_t* ell_idx, const T* ell_val, int64_t n_rhs, const T* X,
                            int64_t x_rs, int64_t x_cs, T* C, int64_t c_rs, int64_t c_cs,
                            double alpha, double beta) {
  RLA_CHECK_CTX(ctx);
  if (out_rows < 0 || in_rows < 0 || n_rhs < 0) return fail(RLA_EARG, "negative dimensions");
  if (out_rows == 0 || n_rhs == 0) return RLA_OK;
  if (!ptr || !ell_idx || !ell_val || !C || (in_rows > 0 && !X))
    return fail(RLA_EARG, "interleaved sparse apply: NULL pointer");
  if (x_rs != 1) return fail(RLA_EARG, "interleaved sparse apply needs X column-major (x_rs == 1)");
  int rc = use_device(ctx);
  if (rc) return rc;
  if (in_rows == 0) {
    // no nonzeros: C = beta * C
    rowgather_direct_kernel<T><<<dim3(static_cast<unsigned>(ceil_div(out_rows, 4)),
                                      static_cast<unsigned>(std::min<int64_t>(ceil_div(n_rhs, 64), 1024))),
                                 256, 0, ctx->stream>>>(out_rows, n_rhs, ptr, ell_idx, ell_val, X,
                                                        x_rs, x_cs, C, c_rs, c_cs, alpha, beta);
    RLA_LAUNCH_CHECK("rowgather_direct");
    return RLA_OK;
  }
  return rowgather_lds<T, true>(ctx, out_rows, in_rows, n_rhs, ptr, ell_idx, ell_val, X, x_cs, C,
                                c_rs, c_cs, alpha, beta);
}

// ---- sign-packed interleaved gather --------------------------------------------
// For operators whose nonzeros all share one magnitude (SparseSign +-1/sqrt(nnz),
// CountSketch +-1, Sampling 1): entry q of row i is one 32-bit word,
// column | (value < 0) << 31, at sidx[q * rows + i], and the magnitude is a
// scalar.  A third of the metadata of the index + value arrays.

template <class T>
__global__ void ell_sign_build_kernel(int64_t rows, int32_t width,
                                      const int32_t* __restrict__ ptr,
                                      const int32_t* __restrict__ idx, const T* __restrict__ val,
                                      uint32_t* __restrict__ sidx, int32_t* __restrict__ mixed) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= rows) return;
  const T mag = val[0] < T(0) ? -val[0] : val[0];
  const int32_t p = ptr[i], n = ptr[i + 1] - p;
  bool bad = false;
  for (int32_t q = 0; q < width; ++q) {
    uint32_t e = 0;
    if (q < n) {
      const T v = val[p + q];
      bad |= !(v == mag || v == -mag) || idx[p + q] < 0;
      e = static_cast<uint32_t>(idx[p + q]) | (v < T(0) ? 0x80000000u : 0u);
    }
    sidx[int64_t(q) * rows + i] = e;
  }
  if (bad) atomicOr(mixed, 1);
}

constexpr int kSignThreads = 1024;
constexpr size_t kSignLdsMax = 160 * 1024;

// One workgroup per column of X: the column is staged in LDS, then each
// thread sums its output rows' entries in ascending column order with the
// products of the index/value kernel (bit-identical results).
template <class T>
__global__ __launch_bounds__(kSignThreads) void rowgather_sign_kernel(
    int64_t out_rows, int64_t in_rows, const int32_t* __restrict__ ptr,
    const uint32_t* __restrict__ sidx, T mag, const T* __restrict__ X, int64_t x_cs,
    T* __restrict__ C, int64_t c_rs, int64_t c_cs, double alpha, double beta, int vec16) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const int64_t k = blockIdx.x;
  const T* xc = X + k * x_cs;
  if (vec16) {
    typedef T v16 __attribute__((ext_vector_type(16 / sizeof(T))));
    constexpr int VW = 16 / sizeof(T);
    const int64_t nv = in_rows / VW;
    for (int64_t vb = 0; vb < nv; vb += 4 * int64_t(kSignThreads)) {
      v16 tmp[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t v = vb + u * int64_t(kSignThreads) + threadIdx.x;
        if (v < nv) tmp[u] = *reinterpret_cast<const v16*>(xc + v * VW);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t v = vb + u * int64_t(kSignThreads) + threadIdx.x;
        if (v < nv) *reinterpret_cast<v16*>(lds + v * VW) = tmp[u];
      }
    }
    for (int64_t j = nv * VW + threadIdx.x; j < in_rows; j += kSignThreads) lds[j] = xc[j];
  } else {
    for (int64_t j = threadIdx.x; j < in_rows; j += kSignThreads) lds[j] = xc[j];
  }
  __syncthreads();
  const double mpos = double(mag), mneg = -double(mag);
  for (int64_t i = threadIdx.x; i < out_rows; i += kSignThreads) {
    const int32_t n = ptr[i + 1] - ptr[i];
    T* cp = C + i * c_rs + k * c_cs;
    T acc = (beta == 0.0) ? T(0) : static_cast<T>(beta * double(*cp));
    const uint32_t* sx = sidx + i;
    int32_t p = 0;
    for (; p + 8 <= n; p += 8) {
      uint32_t e[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) e[u] = sx[int64_t(p + u) * out_rows];
      T xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) xv[u] = lds[e[u] & 0x7fffffffu];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        acc = static_cast<T>(double(acc) +
                             ((e[u] >> 31) ? mneg : mpos) * (double(xv[u]) * alpha));
    }
    for (; p < n; ++p) {
      const uint32_t e = sx[int64_t(p) * out_rows];
      acc = static_cast<T>(double(acc) +
                           ((e >> 31) ? mneg : mpos) * (double(lds[e & 0x7fffffffu]) * alpha));
    }
    *cp = acc;
  }
}

template <class T>
static int ell_sign_build(rla_ctx* ctx, int64_t rows, int32_t width, const int32_t* ptr,
                          const int32_t* idx, const T* val, uint32_t* sidx, double* magnitude) {
  RLA_CHECK_CTX(ctx);
  if (rows < 0 || width < 0) return fail(RLA_EARG, "negative dimensions");
  if (!magnitude) return fail(RLA_EARG, "magnitude is NULL");
  *magnitude = 0.0;
  if (rows == 0 || width == 0) return RLA_OK;
  if (!ptr || !idx || !val || !sidx) return fail(RLA_EARG, "sign-packed ELL build: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  void* ws = nullptr;
  rc = workspace(ctx, 16, &ws);
  if (rc) return rc;
  int32_t* mixed = static_cast<int32_t*>(ws);
  set_i32_kernel<<<1, 64, 0, ctx->stream>>>(mixed, 0);
  ell_sign_build_kernel<T><<<dim3(unsigned(ceil_div(rows, 256))), 256, 0, ctx->stream>>>(
      rows, width, ptr, idx, val, sidx, mixed);
  RLA_LAUNCH_CHECK("ell_sign_build");
  int32_t h_mixed = 0;
  T h_v0 = T(0);
  RLA_HIP(hipMemcpyAsync(&h_mixed, mixed, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipMemcpyAsync(&h_v0, val, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (h_mixed)
    return fail(RLA_EDOMAIN, "sign-packed ELL: the nonzeros do not share one magnitude");
  *magnitude = double(h_v0 < T(0) ? -h_v0 : h_v0);
  return RLA_OK;
}

template <class T>
static int sparse_apply_ell_sign(rla_ctx* ctx, int64_t out_rows, int64_t in_rows,
                                 const int32_t* ptr, const uint32_t* sidx, double magnitude,
                                 int64_t n_rhs, const T* X, int64_t x_rs, int64_t x_cs, T* C,
                                 int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  RLA_CHECK_CTX(ctx);
  if (out_rows < 0 || in_rows < 0 || n_rhs < 0) return fail(RLA_EARG, "negative dimensions");
  if (out_rows == 0 || n_rhs == 0) return RLA_OK;
  if (!ptr || !sidx || !C || (in_rows > 0 && !X))
    return fail(RLA_EARG, "sign-packed sparse apply: NULL pointer");
  if (x_rs != 1) return fail(RLA_EARG, "sign-packed sparse apply needs X column-major (x_rs == 1)");
  if (in_rows < 1 || size_t(in_rows) * sizeof(T) > kSignLdsMax)
    return fail(RLA_ENOTSUP, "sign-packed sparse apply stages a whole column of X: "
                             "1 <= in_rows <= 160 KiB / element size");
  if (n_rhs > INT32_MAX) return fail(RLA_ENOTSUP, "sign-packed sparse apply: too many columns");
  int rc = use_device(ctx);
  if (rc) return rc;
  const size_t lds = size_t(in_rows) * sizeof(T);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(&rowgather_sign_kernel<T>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(kSignLdsMax));
  });
  RLA_HIP(attr_err);
  const int vw = 16 / static_cast<int>(sizeof(T));
  const int vec16 = (reinterpret_cast<uintptr_t>(X) % 16 == 0) && (x_cs % vw == 0);
  rowgather_sign_kernel<T><<<dim3(static_cast<unsigned>(n_rhs)), kSignThreads, lds,
                             ctx->stream>>>(out_rows, in_rows, ptr, sidx, static_cast<T>(magnitude),
                                            X, x_cs, C, c_rs, c_cs, alpha, beta, vec16);
  RLA_LAUNCH_CHECK("rowgather_sign");
  return RLA_OK;
}


// ---- block-interleaved gather -------------------------------------------------
// For X taller than one LDS block: the rows of X are cut into blocks of rb,
// and for each block b the entries of op(M) whose column falls in b are kept
// slot-major, entry s of row i (within the block) at
// bidx/bval[off_b + s * rows + i], with off_b = rows * (w_0 + ... + w_{b-1})
// and w_b the most entries any row has in block b; -1 marks the slots past a
// row's entries.  The lanes of a wavefront own consecutive rows and read the
// same slot of every block, so each load reads consecutive words.  (In the
// row-interleaved copy the lanes' positions drift apart after the first block,
// and one load then touches up to 64 cache lines.)

constexpr int kBellRoundSlots = 4;  // slots per row and round

// w_b by atomic maxima; *bad set if a row's columns are out of range or not
// strictly ascending.
__global__ void bell_count_kernel(int64_t rows, int64_t in_rows, int64_t rb,
                                  const int32_t* __restrict__ ptr,
                                  const int32_t* __restrict__ idx, int32_t* __restrict__ widths,
                                  int32_t* __restrict__ bad) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= rows) return;
  int64_t cur = -1, prev = -1;
  int32_t cnt = 0;
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) {
    const int64_t c = idx[p];
    if (c <= prev || c >= in_rows) {
      atomicOr(bad, 1);
      return;
    }
    prev = c;
    const int64_t b = c / rb;
    if (b != cur) {
      if (cur >= 0) atomicMax(&widths[cur], cnt);
      cur = b;
      cnt = 0;
    }
    ++cnt;
  }
  if (cur >= 0) atomicMax(&widths[cur], cnt);
}

template <class T>
__global__ void bell_fill_kernel(int64_t rows, int64_t rb, int64_t nblocks,
                                 const int32_t* __restrict__ ptr,
                                 const int32_t* __restrict__ idx, const T* __restrict__ val,
                                 const int32_t* __restrict__ widths, int32_t* __restrict__ bidx,
                                 T* __restrict__ bval) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= rows) return;
  int32_t p = ptr[i];
  const int32_t pe = ptr[i + 1];
  int64_t off = i;
  for (int64_t b = 0; b < nblocks; ++b) {
    const int32_t w = widths[b];
    int32_t sl = 0;
    for (; p < pe && idx[p] / rb == b && sl < w; ++p, ++sl) {
      bidx[off + int64_t(sl) * rows] = idx[p];
      bval[off + int64_t(sl) * rows] = val[p];
    }
    for (; sl < w; ++sl) {
      bidx[off + int64_t(sl) * rows] = -1;
      bval[off + int64_t(sl) * rows] = T(0);
    }
    off += int64_t(w) * rows;
  }
}

// One column of X per workgroup (and a part of the output rows), RP rows per
// thread with their sums in registers for the whole launch; per block, stage
// the block of X in LDS, then rounds of kBellRoundSlots slots of all the
// thread's rows, issued together.  The entries of each row are added in
// ascending column order, block after block, as rla_sparse_apply does, so
// the results are identical to it.
template <class T, int RP>
__global__ __launch_bounds__(kGatherThreads) void rowgather_bell_kernel(
    int64_t rows, int64_t in_rows, int64_t rb, int64_t n_rhs,
    const int32_t* __restrict__ widths, const int32_t* __restrict__ bidx,
    const T* __restrict__ bval, const T* __restrict__ X, int64_t x_cs, T* __restrict__ C,
    int64_t c_rs, int64_t c_cs, double alpha, double beta, int rsplit, int64_t rows_per_part,
    int vec_ok) {
  constexpr int B = kBellRoundSlots;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  // Block -> (column, part); the parts of one column are 8 block ids apart,
  // so they land on the same XCD and share its L2 for the column of X.
  const int bid = blockIdx.x;
  const int grp = bid / (8 * rsplit), within = bid % (8 * rsplit);
  const int64_t k = int64_t(grp) * 8 + (within % 8);
  if (k >= n_rhs) return;
  const int64_t r_begin = int64_t(within / 8) * rows_per_part;
  const int64_t r_end = r_begin + rows_per_part < rows ? r_begin + rows_per_part : rows;
  const uint32_t rows32 = static_cast<uint32_t>(rows);
  bool valid[RP];
  T acc[RP];
#pragma unroll
  for (int q = 0; q < RP; ++q) {
    const int64_t i = r_begin + threadIdx.x + int64_t(q) * blockDim.x;
    valid[q] = i < r_end;
    acc[q] = (!valid[q] || beta == 0.0) ? T(0)
                                        : static_cast<T>(beta * double(C[i * c_rs + k * c_cs]));
  }
  const T* xc = X + k * x_cs;
  const int64_t nblocks = (in_rows + rb - 1) / rb;
  uint32_t off = static_cast<uint32_t>(r_begin + threadIdx.x);  // element of slot 0, row q = 0
  for (int64_t blk = 0; blk < nblocks; ++blk) {
    const int64_t j0 = blk * rb;
    const int64_t jn = in_rows - j0 < rb ? in_rows - j0 : rb;
    const int32_t w = widths[blk];
    __syncthreads();
    if (vec_ok) {
      constexpr int VW = 16 / sizeof(T);
      const int64_t nv = jn / VW;
      for (int64_t vb = 0; vb < nv; vb += 8 * int64_t(blockDim.x)) {
        uint4 tmp[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t v = vb + u * int64_t(blockDim.x) + threadIdx.x;
          tmp[u] = v < nv ? reinterpret_cast<const uint4*>(xc + j0)[v] : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t v = vb + u * int64_t(blockDim.x) + threadIdx.x;
          if (v < nv) reinterpret_cast<uint4*>(lds)[v] = tmp[u];
        }
      }
      for (int64_t j = nv * VW + threadIdx.x; j < jn; j += blockDim.x) lds[j] = xc[j0 + j];
    } else {
      for (int64_t jb = 0; jb < jn; jb += 8 * int64_t(blockDim.x)) {
        T tmp[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t j = jb + u * int64_t(blockDim.x) + threadIdx.x;
          tmp[u] = j < jn ? xc[j0 + j] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t j = jb + u * int64_t(blockDim.x) + threadIdx.x;
          if (j < jn) lds[j] = tmp[u];
        }
      }
    }
    __syncthreads();
    bool act[RP];
    bool any = false;
#pragma unroll
    for (int q = 0; q < RP; ++q) {
      act[q] = valid[q] && w > 0;
      any = any || act[q];
    }
    for (int32_t s0 = 0; any; s0 += B) {
      int32_t jj[RP][B];
      T vv[RP][B];
#pragma unroll
      for (int q = 0; q < RP; ++q)
#pragma unroll
        for (int u = 0; u < B; ++u) {
          const bool in = act[q] && s0 + u < w;
          const uint32_t e = off + static_cast<uint32_t>(q) * blockDim.x +
                             static_cast<uint32_t>(s0 + u) * rows32;
          jj[q][u] = in ? bidx[e] : -1;
          vv[q][u] = in ? bval[e] : T(0);
        }
      T xv[RP][B];
#pragma unroll
      for (int q = 0; q < RP; ++q)
#pragma unroll
        for (int u = 0; u < B; ++u) {
          const int64_t d = int64_t(jj[q][u]) - j0;
          xv[q][u] = lds[(d >= 0 && d < jn) ? d : 0];
        }
      any = false;
#pragma unroll
      for (int q = 0; q < RP; ++q) {
        bool more = act[q];
#pragma unroll
        for (int u = 0; u < B; ++u) {
          if (more && jj[q][u] >= 0) {
            acc[q] = static_cast<T>(double(acc[q]) + double(vv[q][u]) * (double(xv[q][u]) * alpha));
          } else {
            more = false;  // first empty slot (or the block's width): the row is done
          }
        }
        act[q] = more && s0 + B < w;
        any = any || act[q];
      }
    }
    off += static_cast<uint32_t>(w) * rows32;
  }
#pragma unroll
  for (int q = 0; q < RP; ++q) {
    const int64_t i = r_begin + threadIdx.x + int64_t(q) * blockDim.x;
    if (valid[q]) C[i * c_rs + k * c_cs] = acc[q];
  }
}

// Block-split form of rowgather_bell_kernel for tall X.  A workgroup takes NC
// columns of X (the entries of M it reads serve all of them) and a run of G
// consecutive rb-row blocks: per block it stages the NC column blocks in LDS
// and walks each output row's slots, adding entry by entry in ascending
// order into FP64 sums; with CARRY (at most 2 rows per thread) the sums run
// on across the G blocks and are written once, otherwise G = 1 and each
// block's sums are written.  bell_combine_kernel adds the runs in ascending
// order (part[(k * nruns + run) * rows + i]).  Workgroup id: XCD id % 8 takes
// runs x, x + 8, ..., all column groups of one run before the next, so a
// run's slots are read from that XCD's L2.  Rows are summed run by run rather
// than entry by entry in one accumulator, so the result is within rounding
// of rla_sparse_apply, not identical to it.
template <class T, int NC, bool CARRY>
__global__ __launch_bounds__(kGatherThreads) void rowgather_bell_split_kernel(
    int64_t rows, int64_t in_rows, int64_t rb, int64_t nblocks, int64_t n_rhs, int64_t G,
    int64_t nruns, int64_t ncg, const int32_t* __restrict__ widths,
    const int32_t* __restrict__ bidx, const T* __restrict__ bval, const T* __restrict__ X,
    int64_t x_cs, int vec_ok, double* __restrict__ part) {
  constexpr int B = kBellRoundSlots;
  constexpr int RP = CARRY ? 2 : 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  __shared__ int64_t slot0;
  const int64_t id = blockIdx.x;
  const int64_t xcd = id % 8, seq = id / 8;
  const int64_t run = (seq / ncg) * 8 + xcd;
  const int64_t k0 = (seq % ncg) * NC;
  if (run >= nruns) return;
  const int64_t b_first = run * G;
  const int64_t b_end = b_first + G < nblocks ? b_first + G : nblocks;
  const int64_t ls = rb < in_rows ? rb : in_rows;  // LDS stride between the columns
  if (threadIdx.x == 0) {
    int64_t w = 0;
    for (int64_t b = 0; b < b_first; ++b) w += widths[b];
    slot0 = w;
  }
  double acc[RP][NC];
#pragma unroll
  for (int q = 0; q < RP; ++q)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[q][c] = 0.0;
  int64_t sl = -1;  // first slot of the current block (read from slot0 after a barrier)
  for (int64_t blk = b_first; blk < b_end; ++blk) {
    const int64_t j0 = blk * rb;
    const int64_t jn = in_rows - j0 < rb ? in_rows - j0 : rb;
    __syncthreads();  // the previous block's reads of LDS are done; slot0 is visible
    if (sl < 0) sl = slot0;
    // the NC column blocks as one flat range, 8 loads per thread in flight
    const int64_t ncv = n_rhs - k0 < NC ? n_rhs - k0 : NC;  // columns present
    if (vec_ok) {
      constexpr int VW = 16 / sizeof(T);
      const int64_t nv = jn / VW, tot = ncv * nv;
      for (int64_t vb = 0; vb < tot; vb += 8 * int64_t(blockDim.x)) {
        uint4 tmp[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t f = vb + u * int64_t(blockDim.x) + threadIdx.x;
          const int64_t c = f / nv, v = f - c * nv;
          tmp[u] = f < tot ? reinterpret_cast<const uint4*>(X + (k0 + c) * x_cs + j0)[v]
                           : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t f = vb + u * int64_t(blockDim.x) + threadIdx.x;
          const int64_t c = f / nv, v = f - c * nv;
          if (f < tot) reinterpret_cast<uint4*>(lds + c * ls)[v] = tmp[u];
        }
      }
      for (int64_t c = 0; c < ncv; ++c)
        for (int64_t j = nv * VW + threadIdx.x; j < jn; j += blockDim.x)
          lds[c * ls + j] = X[(k0 + c) * x_cs + j0 + j];
    } else {
      const int64_t tot = ncv * jn;
      for (int64_t fb = 0; fb < tot; fb += 8 * int64_t(blockDim.x)) {
        T tmp[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t f = fb + u * int64_t(blockDim.x) + threadIdx.x;
          const int64_t c = f / jn, j = f - c * jn;
          tmp[u] = f < tot ? X[(k0 + c) * x_cs + j0 + j] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t f = fb + u * int64_t(blockDim.x) + threadIdx.x;
          const int64_t c = f / jn, j = f - c * jn;
          if (f < tot) lds[c * ls + j] = tmp[u];
        }
      }
    }
    __syncthreads();
    const int32_t w = widths[blk];
    const int64_t base = sl * rows;
    // one output row's slots in this block, added to a[0..NC)
    auto walk = [&](int64_t i, double (&a)[NC]) {
      for (int32_t s0 = 0; s0 < w; s0 += B) {
        int32_t jj[B];
        T vv[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
          const bool in = s0 + u < w;
          const int64_t e = base + int64_t(s0 + u) * rows + i;
          jj[u] = in ? bidx[e] : -1;
          vv[u] = in ? bval[e] : T(0);
        }
        bool done = false;
#pragma unroll
        for (int u = 0; u < B; ++u) {
          const int64_t d = int64_t(jj[u]) - j0;
          const int64_t dd = (d >= 0 && d < jn) ? d : 0;
          if (jj[u] < 0) done = true;  // first empty slot: the row is done
          if (!done) {
#pragma unroll
            for (int c = 0; c < NC; ++c) a[c] += double(vv[u]) * double(lds[c * ls + dd]);
          }
        }
        if (done) break;
      }
    };
    if (CARRY) {
      // the thread's rows walk together: per round, BS slots of each of them
      // are loaded at once, then added in slot order per row
      constexpr int BS = 2 * B;
      bool act[RP];
      bool any = false;
#pragma unroll
      for (int q = 0; q < RP; ++q) {
        act[q] = threadIdx.x + int64_t(q) * blockDim.x < rows && w > 0;
        any = any || act[q];
      }
      for (int32_t s0 = 0; any; s0 += BS) {
        int32_t jj[RP][BS];
        T vv[RP][BS];
#pragma unroll
        for (int q = 0; q < RP; ++q)
#pragma unroll
          for (int u = 0; u < BS; ++u) {
            const bool in = act[q] && s0 + u < w;
            const int64_t e =
                base + int64_t(s0 + u) * rows + threadIdx.x + int64_t(q) * blockDim.x;
            jj[q][u] = in ? bidx[e] : -1;
            vv[q][u] = in ? bval[e] : T(0);
          }
        any = false;
#pragma unroll
        for (int q = 0; q < RP; ++q) {
          bool more = act[q];
#pragma unroll
          for (int u = 0; u < BS; ++u) {
            if (more && jj[q][u] >= 0) {
              const int64_t d = int64_t(jj[q][u]) - j0;
              const int64_t dd = (d >= 0 && d < jn) ? d : 0;
#pragma unroll
              for (int c = 0; c < NC; ++c)
                acc[q][c] += double(vv[q][u]) * double(lds[c * ls + dd]);
            } else {
              more = false;  // first empty slot (or the block's width): the row is done
            }
          }
          act[q] = more && s0 + BS < w;
          any = any || act[q];
        }
      }
    } else {
      for (int64_t i = threadIdx.x; i < rows; i += blockDim.x) {
        double a[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) a[c] = 0.0;
        walk(i, a);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (k0 + c < n_rhs) part[((k0 + c) * nruns + run) * rows + i] = a[c];
      }
    }
    sl += w;
  }
  if (CARRY) {
#pragma unroll
    for (int q = 0; q < RP; ++q) {
      const int64_t i = threadIdx.x + int64_t(q) * blockDim.x;
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (i < rows && k0 + c < n_rhs) part[((k0 + c) * nruns + run) * rows + i] = acc[q][c];
    }
  }
}

template <class T>
__global__ void bell_combine_kernel(int64_t rows, int64_t nblocks, int64_t n_rhs,
                                    const double* __restrict__ part, T* __restrict__ C,
                                    int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= rows * n_rhs) return;
  const int64_t i = e % rows, k = e / rows;
  const double* p = part + k * nblocks * rows + i;
  double s = 0.0;
  for (int64_t b = 0; b < nblocks; ++b) s += p[b * rows];
  T* c = C + i * c_rs + k * c_cs;
  *c = static_cast<T>(beta == 0.0 ? alpha * s : alpha * s + beta * double(*c));
}

// Shape of the block-split apply: NC columns per workgroup (as many rb-row
// blocks as fit kSplitLds of LDS, at most 8), runs of G blocks when every
// thread carries at most 2 output rows (about 256 workgroups in all).
constexpr size_t kSplitLds = 128 * 1024;

struct SplitShape {
  int nc;
  bool carry;
  int64_t g, nruns, ncg;
};

static SplitShape split_shape(int64_t rows, int64_t in_rows, int64_t rb, int64_t n_rhs,
                              size_t elem) {
  SplitShape sh;
  const int64_t nblocks = ceil_div(in_rows, rb);
  const size_t col_bytes = size_t(std::min(rb, in_rows)) * elem;
  sh.nc = 1;
  while (sh.nc < 8 && sh.nc < n_rhs && size_t(2 * sh.nc) * col_bytes <= kSplitLds) sh.nc *= 2;
  sh.ncg = ceil_div(n_rhs, sh.nc);
  sh.carry = rows <= 2 * int64_t(kGatherThreads);
  sh.g = sh.carry ? std::max<int64_t>(1, ceil_div(nblocks * sh.ncg, 256)) : 1;
  sh.nruns = ceil_div(nblocks, sh.g);
  return sh;
}

// partial sums of the block-split apply (one per run, column and row)
static size_t bell_split_bytes(int64_t rows, int64_t in_rows, int64_t rb, int64_t n_rhs,
                               size_t elem) {
  const SplitShape sh = split_shape(rows, in_rows, rb, n_rhs, elem);
  return sizeof(double) * size_t(sh.nruns) * size_t(n_rhs) * size_t(rows);
}

template <class T, int NC, bool CARRY>
static int launch_split(rla_ctx* ctx, const SplitShape& sh, unsigned ids, size_t lds,
                        int64_t rows, int64_t in_rows, int64_t rb, int64_t nblocks,
                        int64_t n_rhs, const int32_t* widths, const int32_t* bidx, const T* bval,
                        const T* X, int64_t x_cs, int vec_ok, double* part) {
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&rowgather_bell_split_kernel<T, NC, CARRY>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(kSplitLds));
  });
  RLA_HIP(attr_err);
  rowgather_bell_split_kernel<T, NC, CARRY><<<dim3(ids), kGatherThreads, lds, ctx->stream>>>(
      rows, in_rows, rb, nblocks, n_rhs, sh.g, sh.nruns, sh.ncg, widths, bidx, bval, X, x_cs,
      vec_ok, part);
  return RLA_OK;
}

template <class T>
static int sparse_apply_bell_split(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                                   const int32_t* widths, const int32_t* bidx, const T* bval,
                                   int64_t n_rhs, const T* X, int64_t x_rs, int64_t x_cs, T* C,
                                   int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  RLA_CHECK_CTX(ctx);
  if (rows < 0 || n_rhs < 0) return fail(RLA_EARG, "negative dimensions");
  if (in_rows < 1 || rb < 1)
    return fail(RLA_EARG, "block-interleaved apply: need in_rows >= 1 and rb >= 1");
  if (rows == 0 || n_rhs == 0) return RLA_OK;
  if (!widths || !bidx || !bval || !X || !C)
    return fail(RLA_EARG, "block-interleaved apply: NULL pointer");
  if (x_rs != 1)
    return fail(RLA_EARG, "block-interleaved apply needs X column-major (x_rs == 1)");
  if (size_t(rb) * sizeof(T) > kGatherLds)
    return fail(RLA_ENOTSUP, "block-interleaved apply: rb * element size must be <= 64 KiB");
  const int64_t nblocks = ceil_div(in_rows, rb);
  const SplitShape sh = split_shape(rows, in_rows, rb, n_rhs, sizeof(T));
  const int64_t ids = ceil_div(sh.nruns, 8) * 8 * sh.ncg;
  if (ids > INT32_MAX || rows * n_rhs > INT32_MAX * int64_t(256))
    return fail(RLA_ENOTSUP, "block-interleaved apply: grid too large");
  int rc = use_device(ctx);
  if (rc) return rc;
  void* ws = nullptr;
  rc = workspace(ctx, bell_split_bytes(rows, in_rows, rb, n_rhs, sizeof(T)), &ws);
  if (rc) return rc;
  double* part = static_cast<double*>(ws);
  const int64_t ls = std::min(rb, in_rows);
  const size_t lds = size_t(sh.nc) * size_t(ls) * sizeof(T);
  const int vec_ok = reinterpret_cast<uintptr_t>(X) % 16 == 0 &&
                     (x_cs * int64_t(sizeof(T))) % 16 == 0 &&
                     (rb * int64_t(sizeof(T))) % 16 == 0 && (ls * int64_t(sizeof(T))) % 16 == 0;
  const unsigned g = static_cast<unsigned>(ids);
#define RLA_SPLIT(NC)                                                                          \
  rc = sh.carry ? launch_split<T, NC, true>(ctx, sh, g, lds, rows, in_rows, rb, nblocks, n_rhs, \
                                            widths, bidx, bval, X, x_cs, vec_ok, part)         \
                : launch_split<T, NC, false>(ctx, sh, g, lds, rows, in_rows, rb, nblocks,      \
                                             n_rhs, widths, bidx, bval, X, x_cs, vec_ok, part)
  switch (sh.nc) {
    case 8: RLA_SPLIT(8); break;
    case 4: RLA_SPLIT(4); break;
    case 2: RLA_SPLIT(2); break;
    default: RLA_SPLIT(1); break;
  }
#undef RLA_SPLIT
  if (rc) return rc;
  RLA_LAUNCH_CHECK("rowgather_bell_split");
  bell_combine_kernel<T><<<dim3(unsigned(ceil_div(rows * n_rhs, 256))), 256, 0, ctx->stream>>>(
      rows, sh.nruns, n_rhs, part, C, c_rs, c_cs, alpha, beta);
  RLA_LAUNCH_CHECK("bell_combine");
  return RLA_OK;
}

static int bell_widths(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                       const int32_t* ptr, const int32_t* idx, int32_t* widths,
                       int64_t* total) {
  RLA_CHECK_CTX(ctx);
  if (!total) return fail(RLA_EARG, "block-interleaved widths: total is NULL");
  *total = 0;
  if (rows < 0 || in_rows < 1 || rb < 1)
    return fail(RLA_EARG, "block-interleaved widths: need rows >= 0, in_rows >= 1, rb >= 1");
  if (rows > INT32_MAX) return fail(RLA_ENOTSUP, "block-interleaved copy: too many rows");
  if (!ptr || !widths || (rows > 0 && !idx))
    return fail(RLA_EARG, "block-interleaved widths: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  const int64_t nblocks = ceil_div(in_rows, rb);
  void* ws = nullptr;
  rc = workspace(ctx, 16, &ws);
  if (rc) return rc;
  int32_t* bad = static_cast<int32_t*>(ws);
  set_i32_kernel<<<1, 64, 0, ctx->stream>>>(bad, 0);
  zero_i32_kernel<<<dim3(unsigned(ceil_div(nblocks, 256))), 256, 0, ctx->stream>>>(widths,
                                                                                nblocks);
  if (rows > 0)
    bell_count_kernel<<<dim3(unsigned(ceil_div(rows, 256))), 256, 0, ctx->stream>>>(
        rows, in_rows, rb, ptr, idx, widths, bad);
  RLA_LAUNCH_CHECK("bell_count");
  std::vector<int32_t> hw(static_cast<size_t>(nblocks));
  int32_t h_bad = 0;
  RLA_HIP(hipMemcpyAsync(hw.data(), widths, hw.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                         ctx->stream));
  RLA_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (h_bad)
    return fail(RLA_EDOMAIN, "block-interleaved copy: each row's columns must lie in "
                             "[0, in_rows) and be strictly ascending");
  int64_t t = 0;
  for (int32_t v : hw) t += v;
  if (t * rows >= (int64_t(1) << 31))
    return fail(RLA_ENOTSUP, "block-interleaved copy: more than 2^31 slots");
  *total = t;
  return RLA_OK;
}

template <class T>
static int bell_build(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb, const int32_t* ptr,
                      const int32_t* idx, const T* val, const int32_t* widths, int32_t* bidx,
                      T* bval) {
  RLA_CHECK_CTX(ctx);
  if (rows < 0 || in_rows < 1 || rb < 1)
    return fail(RLA_EARG, "block-interleaved build: need rows >= 0, in_rows >= 1, rb >= 1");
  if (rows == 0) return RLA_OK;
  if (!ptr || !idx || !val || !widths || !bidx || !bval)
    return fail(RLA_EARG, "block-interleaved build: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  bell_fill_kernel<T><<<dim3(unsigned(ceil_div(rows, 256))), 256, 0, ctx->stream>>>(
      rows, rb, ceil_div(in_rows, rb), ptr, idx, val, widths, bidx, bval);
  RLA_LAUNCH_CHECK("bell_fill");
  return RLA_OK;
}

template <class T, int RP>
static void launch_bell(rla_ctx* ctx, dim3 grid, size_t lds, int64_t rows, int64_t in_rows,
                        int64_t rb, int64_t n_rhs, const int32_t* widths, const int32_t* bidx,
                        const T* bval, const T* X, int64_t x_cs, T* C, int64_t c_rs,
                        int64_t c_cs, double alpha, double beta, int rsplit,
                        int64_t rows_per_part, int vec_ok) {
  rowgather_bell_kernel<T, RP><<<grid, kGatherThreads, lds, ctx->stream>>>(
      rows, in_rows, rb, n_rhs, widths, bidx, bval, X, x_cs, C, c_rs, c_cs, alpha, beta, rsplit,
      rows_per_part, vec_ok);
}

template <class T>
static int sparse_apply_bell(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                             const int32_t* widths, const int32_t* bidx, const T* bval,
                             int64_t n_rhs, const T* X, int64_t x_rs, int64_t x_cs, T* C,
                             int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  RLA_CHECK_CTX(ctx);
  if (rows < 0 || n_rhs < 0) return fail(RLA_EARG, "negative dimensions");
  if (in_rows < 1 || rb < 1)
    return fail(RLA_EARG, "block-interleaved apply: need in_rows >= 1 and rb >= 1");
  if (rows == 0 || n_rhs == 0) return RLA_OK;
  if (!widths || !bidx || !bval || !X || !C)
    return fail(RLA_EARG, "block-interleaved apply: NULL pointer");
  if (x_rs != 1)
    return fail(RLA_EARG, "block-interleaved apply needs X column-major (x_rs == 1)");
  if (size_t(rb) * sizeof(T) > kGatherLds)
    return fail(RLA_ENOTSUP, "block-interleaved apply: rb * element size must be <= 64 KiB");
  if (rows > INT32_MAX) return fail(RLA_ENOTSUP, "block-interleaved apply: too many rows");
  int rc = use_device(ctx);
  if (rc) return rc;
  // Parts of at most 4 * kGatherThreads rows and at least one wavefront's
  // worth, and about two workgroups per CU.  The parts of a column read it
  // from the L2 of their XCD.
  const int64_t min_parts = ceil_div(rows, 4 * int64_t(kGatherThreads));
  const int64_t max_parts = std::max(min_parts, ceil_div(rows, 64));
  const int64_t want = ceil_div(512, n_rhs);
  const int64_t rsplit = std::max(min_parts, std::min(max_parts, want));
  const int64_t rows_per_part = ceil_div(rows, rsplit);
  const int64_t blocks = ceil_div(n_rhs, 8) * 8 * rsplit;
  if (blocks > INT32_MAX) return fail(RLA_ENOTSUP, "block-interleaved apply: grid too large");
  const size_t lds = size_t(std::min(rb, in_rows)) * sizeof(T);
  const int vec_ok = reinterpret_cast<uintptr_t>(X) % 16 == 0 &&
                     (x_cs * int64_t(sizeof(T))) % 16 == 0 && (rb * int64_t(sizeof(T))) % 16 == 0;
  const dim3 grid(static_cast<unsigned>(blocks));
  const int rs = static_cast<int>(rsplit);
  if (rows_per_part <= kGatherThreads)
    launch_bell<T, 1>(ctx, grid, lds, rows, in_rows, rb, n_rhs, widths, bidx, bval, X, x_cs, C,
                      c_rs, c_cs, alpha, beta, rs, rows_per_part, vec_ok);
  else if (rows_per_part <= 2 * kGatherThreads)
    launch_bell<T, 2>(ctx, grid, lds, rows, in_rows, rb, n_rhs, widths, bidx, bval, X, x_cs, C,
                      c_rs, c_cs, alpha, beta, rs, rows_per_part, vec_ok);
  else
    launch_bell<T, 4>(ctx, grid, lds, rows, in_rows, rb, n_rhs, widths, bidx, bval, X, x_cs, C,
                      c_rs, c_cs, alpha, beta, rs, rows_per_part, vec_ok);
  RLA_LAUNCH_CHECK("rowgather_bell");
  return RLA_OK;
}

// ---- row-group tiles: tall X, one magnitude, at most 1024 output rows ------
//
// The lanes of a wavefront are 64 columns of X and the wavefront owns RW
// output rows, whose FP32 sums sit in registers for the whole launch.  X is
// read in blocks of kTgJ rows: a block of 64 columns is staged in LDS column
// by column (column stride kTgJ + 1 words, so a wavefront reading row j of all
// 64 columns hits 64 distinct banks, and word kTgJ of every column holds 0).
// Each entry of M is then one wave-uniform 32-bit word: the block-local row
// of X times 4 (the LDS byte offset) in the low half and its sign as the
// half-precision number +-1 in the high half, read once for all 64 columns:
// per entry one LDS read, one address add and one mixed-precision FMA
// (+-1 * x + sum, exact in the multiply) per lane.
//
// Record (block b, row group g) of a tiled copy, in 32-bit words:
//   header: (RWG + 1) 16-bit step offsets, row i of the group owning steps
//           [ptr[i], ptr[i+1]) (rounded up to whole words, then to 4 words);
//   steps:  two entry words each, a row's entries in ascending column order,
//           padded to an even count with the word of the zero row kTgJ;
// records are l32 words apart (l32 a multiple of 4).  The workgroup of
// (range q, row group g, column group) stages records and X blocks of its
// range in turn (the next block's loads are in flight while the current one
// is gathered), writes its FP32 sums, and a combine kernel adds the ranges'
// sums in ascending order in FP64 and applies alpha*magnitude and beta.
constexpr int kTgJ = 384;
constexpr int kTgThreads = 1024;
constexpr int kTgWaves = kTgThreads / 64;
constexpr int kTgStride = kTgJ + 1;
constexpr uint32_t kTgPlus = 0x3c000000u, kTgMinus = 0xbc000000u;  // +-1 in the high half
constexpr uint32_t kTgDummy = uint32_t(kTgJ) * 4u | kTgPlus;
constexpr size_t kTgLdsX = size_t(64) * kTgStride * sizeof(float);
constexpr size_t kTgLdsMax = 160 * 1024;
constexpr size_t kTgSlack = 16;  // LDS past a record: the last step reads two words ahead
constexpr int64_t kTgMaxRows = 4096;  // four row groups of 1024
constexpr int kTgWorkgroups = 256;  // one per CU

struct TileShape {
  int rw, rwg, g, h32, ncg, q;
  int64_t nblocks;
};

static TileShape tile_shape(int64_t rows, int64_t in_rows, int64_t n_rhs) {
  TileShape s;
  s.rw = 4;
  while (s.rw < 64 && int64_t(kTgWaves) * s.rw < rows) s.rw *= 2;
  s.rwg = kTgWaves * s.rw;
  s.g = int(ceil_div(std::max<int64_t>(rows, 1), s.rwg));
  s.h32 = int(ceil_div((s.rwg + 2) / 2, 4) * 4);
  s.nblocks = ceil_div(in_rows, kTgJ);
  s.ncg = int(ceil_div(std::max<int64_t>(n_rhs, 1), 64));
  const int64_t q = kTgWorkgroups / (int64_t(s.g) * s.ncg);
  s.q = int(std::max<int64_t>(1, std::min<int64_t>(q, s.nblocks)));
  return s;
}

static size_t tile_part_bytes(const TileShape& s) {
  return sizeof(float) * size_t(s.ncg) * size_t(s.g) * size_t(s.q) * size_t(s.rwg) * 64;
}

// pc[b * rows + i] = steps of row i in block b (pc zeroed before): entries,
// or entry pairs without values; *bad set
// when a row's columns are out of range or not strictly ascending.
__global__ void tile_count_kernel(int64_t rows, int64_t in_rows, const int32_t* __restrict__ ptr,
                                  const int32_t* __restrict__ idx, int vals,
                                  int32_t* __restrict__ pc, int32_t* __restrict__ bad) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= rows) return;
  int64_t prev = -1, cur = -1;
  int32_t cnt = 0;
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) {
    const int64_t c = idx[p];
    if (c <= prev || c >= in_rows) {
      atomicOr(bad, 1);
      return;
    }
    prev = c;
    const int64_t b = c / kTgJ;
    if (b != cur) {
      if (cur >= 0) pc[cur * rows + i] = vals ? cnt : (cnt + 1) / 2;
      cur = b;
      cnt = 0;
    }
    ++cnt;
  }
  if (cur >= 0) pc[cur * rows + i] = vals ? cnt : (cnt + 1) / 2;
}

// One workgroup per record: the maximum of the records' pair counts in *mx.
__global__ __launch_bounds__(kTgThreads) void tile_stats_kernel(int64_t rows, int rwg,
                                                                int g_count,
                                                                const int32_t* __restrict__ pc,
                                                                int32_t* __restrict__ mx) {
  __shared__ int32_t s[kTgThreads];
  const int64_t b = blockIdx.x / g_count;
  const int g = blockIdx.x % g_count;
  const int t = threadIdx.x;
  const int64_t i = int64_t(g) * rwg + t;
  s[t] = (t < rwg && i < rows) ? pc[b * rows + i] : 0;
  __syncthreads();
  if (t == 0) {
    int32_t tot = 0;
    for (int k = 0; k < rwg; ++k) tot += s[k];
    atomicMax(mx, tot);
  }
}

// Record headers: exclusive prefix sums of each row group's pair counts.
__global__ __launch_bounds__(kTgThreads) void tile_header_kernel(int64_t rows, int rwg,
                                                                 int g_count, int64_t l32,
                                                                 const int32_t* __restrict__ pc,
                                                                 uint32_t* __restrict__ tiles) {
  __shared__ int32_t s[kTgThreads];
  const int64_t b = blockIdx.x / g_count;
  const int g = blockIdx.x % g_count;
  const int t = threadIdx.x;
  const int64_t i = int64_t(g) * rwg + t;
  s[t] = (t < rwg && i < rows) ? pc[b * rows + i] : 0;
  __syncthreads();
  if (t == 0) {
    uint16_t* h = reinterpret_cast<uint16_t*>(tiles + int64_t(blockIdx.x) * l32);
    int32_t run = 0;
    for (int k = 0; k < rwg; ++k) {
      h[k] = static_cast<uint16_t>(run);
      run += s[k];
    }
    h[rwg] = static_cast<uint16_t>(run);
  }
}

// Entries: thread per output row, walking its CSR row once.  VAL: each step
// is one entry, its offset word then its value; otherwise two sign words.
template <class T, bool VAL>
__global__ void tile_fill_kernel(int64_t rows, int rwg, int g_count, int h32, int64_t l32,
                                 const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                 const T* __restrict__ val, uint32_t* __restrict__ tiles,
                                 int32_t* __restrict__ mixed) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= rows) return;
  const int32_t p0 = ptr[i], p1 = ptr[i + 1];
  if (p0 == p1) return;
  const T mag = val[0] < T(0) ? -val[0] : val[0];
  const int g = int(i / rwg), r = int(i % rwg);
  bool bad = false;
  int64_t cur = -1;
  uint32_t* out = nullptr;
  int32_t k = 0;
  for (int32_t p = p0; p < p1; ++p) {
    const int64_t c = idx[p];
    const int64_t b = c / kTgJ;
    if (b != cur) {
      if (!VAL && cur >= 0 && (k & 1)) out[k] = kTgDummy;
      cur = b;
      uint32_t* rec = tiles + (b * g_count + g) * l32;
      const uint16_t* h = reinterpret_cast<const uint16_t*>(rec);
      out = rec + h32 + 2 * int64_t(h[r]);
      k = 0;
    }
    const T v = val[p];
    const uint32_t off = uint32_t(c - b * kTgJ) * 4u;
    if (VAL) {
      out[k++] = off;
      out[k++] = __float_as_uint(static_cast<float>(v));
    } else {
      bad |= !(v == mag || v == -mag);
      out[k++] = off | (v < T(0) ? kTgMinus : kTgPlus);
    }
  }
  if (!VAL && (k & 1)) out[k] = kTgDummy;
  if (bad) atomicOr(mixed, 1);
}

__global__ void tile_zero_kernel(uint32_t* p, int64_t n) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * blockDim.x)
    p[e] = 0u;
}

template <int RW, bool VAL, bool LONG>
__global__ __launch_bounds__(kTgThreads) void tilegather_kernel(
    int64_t in_rows, int64_t nblocks, int G, int Q, int NCG, int64_t n_rhs, int l32, int h32,
    const uint32_t* __restrict__ tiles, const float* __restrict__ X, int64_t x_cs,
    float* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* xl = reinterpret_cast<float*>(smem_raw);
  uint32_t* rec = reinterpret_cast<uint32_t*>(smem_raw + kTgLdsX);
  // the G row groups of one range and column group are ids 8 apart (one XCD)
  const int id = blockIdx.x;
  const int xcd = id % 8, seq = id / 8;
  const int g = seq % G;
  const int64_t qc = int64_t(seq / G) * 8 + xcd;
  if (qc >= int64_t(Q) * NCG) return;
  const int q = int(qc % Q), cg = int(qc / Q);
  const int64_t b0 = int64_t(q) * nblocks / Q, b1 = int64_t(q + 1) * nblocks / Q;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t k0 = int64_t(cg) * 64;
  const int n4 = l32 / 4;
  if (threadIdx.x < 64) xl[threadIdx.x * kTgStride + kTgJ] = 0.f;

  // staging registers: wavefront w holds columns w + 16 t, rows lane + 64 i
  constexpr int NI = kTgJ / 64;
  float sx[4 * NI];
  uint4 sr[2];
  auto load = [&](int64_t b) {
    const int64_t j0 = b * kTgJ;
    const bool full = j0 + kTgJ <= in_rows;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int64_t col = k0 + w + 16 * t;
      const bool colok = col < n_rhs;
      const float* xc = X + (colok ? col : 0) * x_cs + j0 + lane;
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const bool ok = colok && (full || j0 + lane + 64 * i < in_rows);
        sx[t * NI + i] = ok ? xc[64 * i] : 0.f;
      }
    }
    const uint4* src = reinterpret_cast<const uint4*>(tiles + (b * G + g) * int64_t(l32));
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int k = threadIdx.x + u * kTgThreads;
      sr[u] = k < n4 ? src[k] : make_uint4(0u, 0u, 0u, 0u);
    }
  };
  auto store = [&](int64_t b) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < NI; ++i) xl[(w + 16 * t) * kTgStride + lane + 64 * i] = sx[t * NI + i];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int k = threadIdx.x + u * kTgThreads;
      if (k < n4) reinterpret_cast<uint4*>(rec)[k] = sr[u];
    }
    // LONG: a record longer than the staging registers hold; the rest goes
    // straight from global memory to LDS
    if (LONG) {
      const uint4* src = reinterpret_cast<const uint4*>(tiles + (b * G + g) * int64_t(l32));
      for (int k = 2 * kTgThreads + threadIdx.x; k < n4; k += kTgThreads)
        reinterpret_cast<uint4*>(rec)[k] = src[k];
    }
  };

  float acc[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) acc[r] = 0.f;
  const unsigned char* colb = smem_raw + size_t(lane) * kTgStride * sizeof(float);
  if (b0 < b1) load(b0);
  for (int64_t b = b0; b < b1; ++b) {
    __syncthreads();  // the previous block's gathers are done
    store(b);
    __syncthreads();
    if (b + 1 < b1) load(b + 1);
    const uint16_t* rp = reinterpret_cast<const uint16_t*>(rec) + w * RW;
    const int wbase = __builtin_amdgcn_readfirstlane(int(rp[0]));
    const int nsteps = __builtin_amdgcn_readfirstlane(int(rp[RW])) - wbase;
    const int vp = int(rp[lane < RW ? lane : RW]) - wbase;  // lane r: row r's first step
    const uint2* ps = reinterpret_cast<const uint2*>(rec + h32) + wbase;
    auto xat = [&](uint32_t word) -> float {
      return *reinterpret_cast<const float*>(colb + (word & 0xffffu));
    };
    auto sgn = [](uint32_t word) -> float {
      return float(__builtin_bit_cast(_Float16, static_cast<uint16_t>(word >> 16)));
    };
    // the next step's two words are read while this one gathers
    uint2 nx = ps[0];
    int p = 0;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int p1 = r + 1 < RW ? __builtin_amdgcn_readlane(vp, r + 1) : nsteps;
      float a = acc[r];
      if constexpr (VAL) {  // a step: one entry's offset and value
        for (; p + 2 <= p1; p += 2) {
          const uint2 nb = ps[p + 1];
          const float x0 = xat(nx.x);
          a = __builtin_fmaf(__uint_as_float(nx.y), x0, a);
          nx = ps[p + 2];
          const float x1 = xat(nb.x);
          a = __builtin_fmaf(__uint_as_float(nb.y), x1, a);
        }
        if (p < p1) {
          const uint2 cu = nx;
          nx = ps[p + 1];
          a = __builtin_fmaf(__uint_as_float(cu.y), xat(cu.x), a);
          ++p;
        }
      } else {
        for (; p + 2 <= p1; p += 2) {  // two steps, the reads alternating registers
          const uint2 nb = ps[p + 1];
          const float x0 = xat(nx.x), x1 = xat(nx.y);
          a = __builtin_fmaf(sgn(nx.x), x0, a);
          a = __builtin_fmaf(sgn(nx.y), x1, a);
          nx = ps[p + 2];
          const float x2 = xat(nb.x), x3 = xat(nb.y);
          a = __builtin_fmaf(sgn(nb.x), x2, a);
          a = __builtin_fmaf(sgn(nb.y), x3, a);
        }
        if (p < p1) {
          const uint2 cu = nx;
          nx = ps[p + 1];
          const float x0 = xat(cu.x), x1 = xat(cu.y);
          a = __builtin_fmaf(sgn(cu.x), x0, a);
          a = __builtin_fmaf(sgn(cu.y), x1, a);
          ++p;
        }
      }
      acc[r] = a;
    }
  }
  float* pp = part + ((int64_t(cg) * G + g) * Q + q) * int64_t(kTgWaves * RW) * 64 +
              int64_t(w) * RW * 64 + lane;
#pragma unroll
  for (int r = 0; r < RW; ++r) pp[r * 64] = acc[r];
}

// C[i, k] = alpha*magnitude*(sum over ranges, ascending, in FP64) + beta*C.
__global__ void tile_combine_kernel(int64_t rows, int64_t n_rhs, int G, int Q, int rwg,
                                    const float* __restrict__ part, double scale,
                                    float* __restrict__ C, int64_t c_rs, int64_t c_cs,
                                    double beta) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int lane = int(e & 63);
  const int64_t rc = e >> 6;
  const int64_t i = rc % rows, cg = rc / rows;
  const int64_t col = cg * 64 + lane;
  if (col >= n_rhs) return;
  const int g = int(i / rwg), r = int(i % rwg);
  const int64_t stride = int64_t(rwg) * 64;
  const float* p = part + (cg * G + g) * int64_t(Q) * stride + int64_t(r) * 64 + lane;
  double s = 0.0;
  int k = 0;
  for (; k + 16 <= Q; k += 16) {  // 16 loads in flight, added in range order
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = p[(k + u) * stride];
#pragma unroll
    for (int u = 0; u < 16; ++u) s += double(v[u]);
  }
  for (; k < Q; ++k) s += double(p[k * stride]);
  float* c = C + i * c_rs + col * c_cs;
  *c = static_cast<float>(beta == 0.0 ? scale * s : scale * s + beta * double(*c));
}

static int tile_shape_query(rla_ctx* ctx, int64_t rows, int64_t in_rows, const int32_t* ptr,
                            const int32_t* idx, int vals, int64_t* rec_words, int64_t* nrec) {
  RLA_CHECK_CTX(ctx);
  if (!rec_words || !nrec) return fail(RLA_EARG, "row-group tiles: NULL output");
  *rec_words = 0;
  *nrec = 0;
  if (rows < 1 || in_rows < 1)
    return fail(RLA_EARG, "row-group tiles: need rows >= 1 and in_rows >= 1");
  if (rows > kTgMaxRows) return fail(RLA_ENOTSUP, "row-group tiles: more than 4096 rows");
  if (!ptr || !idx) return fail(RLA_EARG, "row-group tiles: NULL pointer");
  const TileShape sh = tile_shape(rows, in_rows, 1);
  if (sh.nblocks * int64_t(sh.g) > INT32_MAX)
    return fail(RLA_ENOTSUP, "row-group tiles: too many records");
  int rc = use_device(ctx);
  if (rc) return rc;
  const size_t pc_bytes = sizeof(int32_t) * size_t(sh.nblocks) * size_t(rows);
  void* ws = nullptr;
  rc = workspace(ctx, 16 + pc_bytes, &ws);
  if (rc) return rc;
  int32_t* flags = static_cast<int32_t*>(ws);  // bad columns, largest record
  int32_t* pc = reinterpret_cast<int32_t*>(static_cast<char*>(ws) + 16);
  zero_i32_kernel<<<1, 64, 0, ctx->stream>>>(flags, 4);
  tile_zero_kernel<<<dim3(unsigned(std::min<int64_t>(4096, ceil_div(sh.nblocks * rows, 256)))),
                    256, 0, ctx->stream>>>(reinterpret_cast<uint32_t*>(pc), sh.nblocks * rows);
  tile_count_kernel<<<dim3(unsigned(ceil_div(rows, 256))), 256, 0, ctx->stream>>>(
      rows, in_rows, ptr, idx, vals ? 1 : 0, pc, flags);
  tile_stats_kernel<<<dim3(unsigned(sh.nblocks * sh.g)), kTgThreads, 0, ctx->stream>>>(
      rows, sh.rwg, sh.g, pc, flags + 1);
  RLA_LAUNCH_CHECK("tile_stats");
  int32_t h[2] = {0, 0};
  RLA_HIP(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (h[0])
    return fail(RLA_EDOMAIN,
                "row-group tiles: columns out of range or not strictly ascending in a row");
  const int64_t words = sh.h32 + ceil_div(2 * int64_t(std::max<int32_t>(h[1], 1)), 4) * 4;
  if (kTgLdsX + size_t(words) * 4 + kTgSlack > kTgLdsMax)
    return fail(RLA_ENOTSUP, "row-group tiles: a record does not fit the LDS");
  *rec_words = words;
  *nrec = sh.nblocks * sh.g;
  return RLA_OK;
}

template <class T>
static int tile_build(rla_ctx* ctx, int64_t rows, int64_t in_rows, const int32_t* ptr,
                      const int32_t* idx, const T* val, int vals, int64_t rec_words,
                      uint32_t* tiles, double* magnitude) {
  RLA_CHECK_CTX(ctx);
  if (!magnitude) return fail(RLA_EARG, "row-group tiles: magnitude is NULL");
  *magnitude = 0.0;
  if (rows < 1 || in_rows < 1 || rows > kTgMaxRows)
    return fail(RLA_EARG, "row-group tiles: need 1 <= rows <= 4096 and in_rows >= 1");
  if (!ptr || !idx || !val || !tiles) return fail(RLA_EARG, "row-group tiles: NULL pointer");
  const TileShape sh = tile_shape(rows, in_rows, 1);
  if (rec_words < sh.h32 + 4 || rec_words % 4 ||
      kTgLdsX + size_t(rec_words) * 4 + kTgSlack > kTgLdsMax)
    return fail(RLA_EARG, "row-group tiles: bad record length");
  if (reinterpret_cast<uintptr_t>(tiles) % 16)
    return fail(RLA_EARG, "row-group tiles: the copy must be 16-byte aligned");
  int rc = use_device(ctx);
  if (rc) return rc;
  int32_t nnz = 0;
  RLA_HIP(hipMemcpyAsync(&nnz, ptr + rows, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  const size_t pc_bytes = sizeof(int32_t) * size_t(sh.nblocks) * size_t(rows);
  void* ws = nullptr;
  rc = workspace(ctx, 16 + pc_bytes, &ws);
  if (rc) return rc;
  int32_t* flags = static_cast<int32_t*>(ws);  // bad columns, mixed magnitudes
  int32_t* pc = reinterpret_cast<int32_t*>(static_cast<char*>(ws) + 16);
  const int64_t nrec = sh.nblocks * sh.g;
  zero_i32_kernel<<<1, 64, 0, ctx->stream>>>(flags, 4);
  tile_zero_kernel<<<dim3(unsigned(std::min<int64_t>(4096, ceil_div(sh.nblocks * rows, 256)))),
                    256, 0, ctx->stream>>>(reinterpret_cast<uint32_t*>(pc), sh.nblocks * rows);
  tile_zero_kernel<<<dim3(unsigned(std::min<int64_t>(4096, ceil_div(nrec * rec_words, 256)))),
                    256, 0, ctx->stream>>>(tiles, nrec * rec_words);
  tile_count_kernel<<<dim3(unsigned(ceil_div(rows, 256))), 256, 0, ctx->stream>>>(
      rows, in_rows, ptr, idx, vals ? 1 : 0, pc, flags);
  tile_header_kernel<<<dim3(unsigned(nrec)), kTgThreads, 0, ctx->stream>>>(rows, sh.rwg, sh.g,
                                                                           rec_words, pc, tiles);
  RLA_LAUNCH_CHECK("tile_header");
  int32_t h[2] = {0, 0};
  RLA_HIP(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (h[0])
    return fail(RLA_EDOMAIN,
                "row-group tiles: columns out of range or not strictly ascending in a row");
  if (vals) *magnitude = 1.0;
  if (nnz == 0) return RLA_OK;
  if (vals)
    tile_fill_kernel<T, true><<<dim3(unsigned(ceil_div(rows, 256))), 256, 0, ctx->stream>>>(
        rows, sh.rwg, sh.g, sh.h32, rec_words, ptr, idx, val, tiles, flags + 1);
  else
    tile_fill_kernel<T, false><<<dim3(unsigned(ceil_div(rows, 256))), 256, 0, ctx->stream>>>(
        rows, sh.rwg, sh.g, sh.h32, rec_words, ptr, idx, val, tiles, flags + 1);
  RLA_LAUNCH_CHECK("tile_fill");
  if (vals) return RLA_OK;
  T v0 = T(0);
  RLA_HIP(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipMemcpyAsync(&v0, val, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (h[1]) return fail(RLA_EDOMAIN, "row-group tiles: the nonzeros do not share one magnitude");
  *magnitude = double(v0 < T(0) ? -v0 : v0);
  return RLA_OK;
}

template <int RW, bool VAL, bool LONG>
static int launch_tile(rla_ctx* ctx, const TileShape& sh, unsigned ids, int64_t in_rows,
                       int64_t n_rhs, int l32, const uint32_t* tiles, const float* X,
                       int64_t x_cs, float* part) {
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&tilegather_kernel<RW, VAL, LONG>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(kTgLdsMax));
  });
  RLA_HIP(attr_err);
  const size_t lds = kTgLdsX + size_t(l32) * 4 + kTgSlack;
  tilegather_kernel<RW, VAL, LONG><<<dim3(ids), kTgThreads, lds, ctx->stream>>>(
      in_rows, sh.nblocks, sh.g, sh.q, sh.ncg, n_rhs, l32, sh.h32, tiles, X, x_cs, part);
  return RLA_OK;
}

template <bool VAL, bool LONG>
static int launch_tile_rw(rla_ctx* ctx, const TileShape& sh, unsigned g, int64_t in_rows,
                          int64_t n_rhs, int l32, const uint32_t* tiles, const float* X,
                          int64_t x_cs, float* part) {
  switch (sh.rw) {
    case 64:
      return launch_tile<64, VAL, LONG>(ctx, sh, g, in_rows, n_rhs, l32, tiles, X, x_cs, part);
    case 32:
      return launch_tile<32, VAL, LONG>(ctx, sh, g, in_rows, n_rhs, l32, tiles, X, x_cs, part);
    case 16:
      return launch_tile<16, VAL, LONG>(ctx, sh, g, in_rows, n_rhs, l32, tiles, X, x_cs, part);
    case 8:
      return launch_tile<8, VAL, LONG>(ctx, sh, g, in_rows, n_rhs, l32, tiles, X, x_cs, part);
    default:
      return launch_tile<4, VAL, LONG>(ctx, sh, g, in_rows, n_rhs, l32, tiles, X, x_cs, part);
  }
}

// part: the ranges' sums (sparse_tile_part_bytes bytes), or NULL for the
// context workspace.
int sparse_apply_tile_part(rla_ctx* ctx, int64_t rows, int64_t in_rows, int vals,
                                int64_t rec_words, const uint32_t* tiles, double magnitude,
                                int64_t n_rhs, const float* X, int64_t x_rs, int64_t x_cs,
                                float* C, int64_t c_rs, int64_t c_cs, double alpha, double beta,
                                float* part) {
  RLA_CHECK_CTX(ctx);
  if (rows < 0 || n_rhs < 0) return fail(RLA_EARG, "negative dimensions");
  if (in_rows < 1) return fail(RLA_EARG, "row-group tile apply: need in_rows >= 1");
  if (rows == 0 || n_rhs == 0) return RLA_OK;
  if (rows > kTgMaxRows) return fail(RLA_ENOTSUP, "row-group tile apply: more than 4096 rows");
  if (!tiles || !X || !C) return fail(RLA_EARG, "row-group tile apply: NULL pointer");
  if (x_rs != 1) return fail(RLA_EARG, "row-group tile apply needs X column-major (x_rs == 1)");
  const TileShape sh = tile_shape(rows, in_rows, n_rhs);
  if (rec_words < sh.h32 + 4 || rec_words % 4 ||
      kTgLdsX + size_t(rec_words) * 4 + kTgSlack > kTgLdsMax)
    return fail(RLA_EARG, "row-group tile apply: bad record length");
  if (reinterpret_cast<uintptr_t>(tiles) % 16)
    return fail(RLA_EARG, "row-group tile apply: the copy must be 16-byte aligned");
  const int64_t ids = ceil_div(int64_t(sh.q) * sh.ncg, 8) * 8 * sh.g;
  if (ids > INT32_MAX || rows * sh.ncg * 64 > INT32_MAX * int64_t(256))
    return fail(RLA_ENOTSUP, "row-group tile apply: grid too large");
  int rc = use_device(ctx);
  if (rc) return rc;
  if (!part) {
    void* ws = nullptr;
    rc = workspace(ctx, tile_part_bytes(sh), &ws);
    if (rc) return rc;
    part = static_cast<float*>(ws);
  }
  const unsigned g = static_cast<unsigned>(ids);
  const int l32 = static_cast<int>(rec_words);
  const bool lng = l32 > 8 * kTgThreads;  // more than 2 uint4 per thread to stage
  if (vals)
    rc = lng ? launch_tile_rw<true, true>(ctx, sh, g, in_rows, n_rhs, l32, tiles, X, x_cs, part)
             : launch_tile_rw<true, false>(ctx, sh, g, in_rows, n_rhs, l32, tiles, X, x_cs, part);
  else
    rc = lng ? launch_tile_rw<false, true>(ctx, sh, g, in_rows, n_rhs, l32, tiles, X, x_cs, part)
             : launch_tile_rw<false, false>(ctx, sh, g, in_rows, n_rhs, l32, tiles, X, x_cs,
                                            part);
  if (rc) return rc;
  RLA_LAUNCH_CHECK("tilegather");
  tile_combine_kernel<<<dim3(unsigned(ceil_div(rows * sh.ncg * 64, 256))), 256, 0,
                        ctx->stream>>>(rows, n_rhs, sh.g, sh.q, sh.rwg, part, alpha * magnitude,
                                       C, c_rs, c_cs, beta);
  RLA_LAUNCH_CHECK("tile_combine");
  return RLA_OK;
}

size_t sparse_tile_part_bytes(int64_t rows, int64_t in_rows, int64_t n_rhs) {
  if (rows <= 0 || n_rhs <= 0 || in_rows < 1) return 0;
  return tile_part_bytes(tile_shape(rows, in_rows, n_rhs));
}
}  // namespace rla

using namespace rla;

extern "C" {

int rla_sparse_transpose_f32(rla_ctx* ctx, int64_t nrows, int64_t ncols, int64_t nnz,
                             const int32_t* colptr, const int32_t* rowidx, const float* val,
                             int32_t* rowptr, int32_t* colidx, float* val_t) {
  return sparse_transpose<float>(ctx, nrows, ncols, nnz, colptr, rowidx, val, rowptr, colidx,
                                 val_t);
}

int rla_sparse_transpose_f64(rla_ctx* ctx, int64_t nrows, int64_t ncols, int64_t nnz,
                             const int32_t* colptr, const int32_t* rowidx, const double* val,
                             int32_t* rowptr, int32_t* colidx, double* val_t) {
  return sparse_transpose<double>(ctx, nrows, ncols, nnz, colptr, rowidx, val, rowptr, colidx,
                                  val_t);
}

int rla_sparse_apply_f32(rla_ctx* ctx, int trans_m, int64_t nrows, int64_t ncols,
                         const int32_t* colptr, const int32_t* rowidx, const float* val,
                         const int32_t* rowptr, const int32_t* colidx, const float* val_t,
                         int64_t n_rhs, const float* X, int64_t x_rs, int64_t x_cs, float* C,
                         int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  return sparse_apply<float>(ctx, trans_m, nrows, ncols, colptr, rowidx, val, rowptr, colidx,
                             val_t, n_rhs, X, x_rs, x_cs, C, c_rs, c_cs, alpha, beta);
}

int rla_sparse_apply_f64(rla_ctx* ctx, int trans_m, int64_t nrows, int64_t ncols,
                         const int32_t* colptr, const int32_t* rowidx, const double* val,
                         const int32_t* rowptr, const int32_t* colidx, const double* val_t,
                         int64_t n_rhs, const double* X, int64_t x_rs, int64_t x_cs, double* C,
                         int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  return sparse_apply<double>(ctx, trans_m, nrows, ncols, colptr, rowidx, val, rowptr, colidx,
                              val_t, n_rhs, X, x_rs, x_cs, C, c_rs, c_cs, alpha, beta);
}

int rla_sparse_ell_width(rla_ctx* ctx, int64_t rows, const int32_t* ptr, int32_t* width) {
  return ell_width(ctx, rows, ptr, width);
}

int rla_sparse_ell_build_f32(rla_ctx* ctx, int64_t rows, int32_t width, const int32_t* ptr,
                             const int32_t* idx, const float* val, int32_t* ell_idx,
                             float* ell_val) {
  return ell_build<float>(ctx, rows, width, ptr, idx, val, ell_idx, ell_val);
}

int rla_sparse_ell_build_f64(rla_ctx* ctx, int64_t rows, int32_t width, const int32_t* ptr,
                             const int32_t* idx, const double* val, int32_t* ell_idx,
                             double* ell_val) {
  return ell_build<double>(ctx, rows, width, ptr, idx, val, ell_idx, ell_val);
}

int rla_sparse_apply_ell_f32(rla_ctx* ctx, int64_t out_rows, int64_t in_rows, const int32_t* ptr,
                             const int32_t* ell_idx, const float* ell_val, int64_t n_rhs,
                             const float* X, int64_t x_rs, int64_t x_cs, float* C, int64_t c_rs,
                             int64_t c_cs, double alpha, double beta) {
  return sparse_apply_ell<float>(ctx, out_rows, in_rows, ptr, ell_idx, ell_val, n_rhs, X, x_rs,
                                 x_cs, C, c_rs, c_cs, alpha, beta);
}

int rla_sparse_apply_ell_f64(rla_ctx* ctx, int64_t out_rows, int64_t in_rows, const int32_t* ptr,
                             const int32_t* ell_idx, const double* ell_val, int64_t n_rhs,
                             const double* X, int64_t x_rs, int64_t x_cs, double* C,
                             int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  return sparse_apply_ell<double>(ctx, out_rows, in_rows, ptr, ell_idx, ell_val, n_rhs, X, x_rs,
                                  x_cs, C, c_rs, c_cs, alpha, beta);
}

int rla_sparse_ell_sign_build_f32(rla_ctx* ctx, int64_t rows, int32_t width,
                                  const int32_t* ptr, const int32_t* idx, const float* val,
                                  uint32_t* ell_sidx, double* magnitude) {
  return ell_sign_build<float>(ctx, rows, width, ptr, idx, val, ell_sidx, magnitude);
}
int rla_sparse_ell_sign_build_f64(rla_ctx* ctx, int64_t rows, int32_t width,
                                  const int32_t* ptr, const int32_t* idx, const double* val,
                                  uint32_t* ell_sidx, double* magnitude) {
  return ell_sign_build<double>(ctx, rows, width, ptr, idx, val, ell_sidx, magnitude);
}
int rla_sparse_apply_ell_sign_f32(rla_ctx* ctx, int64_t out_rows, int64_t in_rows,
                                  const int32_t* ptr, const uint32_t* ell_sidx, double magnitude,
                                  int64_t n_rhs, const float* X, int64_t x_rs, int64_t x_cs,
                                  float* C, int64_t c_rs, int64_t c_cs, double alpha,
                                  double beta) {
  return sparse_apply_ell_sign<float>(ctx, out_rows, in_rows, ptr, ell_sidx, magnitude, n_rhs, X,
                                      x_rs, x_cs, C, c_rs, c_cs, alpha, beta);
}
int rla_sparse_apply_ell_sign_f64(rla_ctx* ctx, int64_t out_rows, int64_t in_rows,
                                  const int32_t* ptr, const uint32_t* ell_sidx, double magnitude,
                                  int64_t n_rhs, const double* X, int64_t x_rs, int64_t x_cs,
                                  double* C, int64_t c_rs, int64_t c_cs, double alpha,
                                  double beta) {
  return sparse_apply_ell_sign<double>(ctx, out_rows, in_rows, ptr, ell_sidx, magnitude, n_rhs,
                                       X, x_rs, x_cs, C, c_rs, c_cs, alpha, beta);
}

int rla_sparse_bell_widths(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                           const int32_t* ptr, const int32_t* idx, int32_t* widths,
                           int64_t* total_slots) {
  return bell_widths(ctx, rows, in_rows, rb, ptr, idx, widths, total_slots);
}

int rla_sparse_bell_build_f32(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                              const int32_t* ptr, const int32_t* idx, const float* val,
                              const int32_t* widths, int32_t* bidx, float* bval) {
  return bell_build<float>(ctx, rows, in_rows, rb, ptr, idx, val, widths, bidx, bval);
}

int rla_sparse_bell_build_f64(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                              const int32_t* ptr, const int32_t* idx, const double* val,
                              const int32_t* widths, int32_t* bidx, double* bval) {
  return bell_build<double>(ctx, rows, in_rows, rb, ptr, idx, val, widths, bidx, bval);
}

int rla_sparse_apply_bell_f32(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                              const int32_t* widths, const int32_t* bidx, const float* bval,
                              int64_t n_rhs, const float* X, int64_t x_rs, int64_t x_cs, float* C,
                              int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  return sparse_apply_bell<float>(ctx, rows, in_rows, rb, widths, bidx, bval, n_rhs, X, x_rs,
                                  x_cs, C, c_rs, c_cs, alpha, beta);
}

int rla_sparse_apply_bell_split_f32(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                                    const int32_t* widths, const int32_t* bidx,
                                    const float* bval, int64_t n_rhs, const float* X,
                                    int64_t x_rs, int64_t x_cs, float* C, int64_t c_rs,
                                    int64_t c_cs, double alpha, double beta) {
  return sparse_apply_bell_split<float>(ctx, rows, in_rows, rb, widths, bidx, bval, n_rhs, X,
                                        x_rs, x_cs, C, c_rs, c_cs, alpha, beta);
}

int rla_sparse_apply_bell_split_f64(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                                    const int32_t* widths, const int32_t* bidx,
                                    const double* bval, int64_t n_rhs, const double* X,
                                    int64_t x_rs, int64_t x_cs, double* C, int64_t c_rs,
                                    int64_t c_cs, double alpha, double beta) {
  return sparse_apply_bell_split<double>(ctx, rows, in_rows, rb, widths, bidx, bval, n_rhs, X,
                                         x_rs, x_cs, C, c_rs, c_cs, alpha, beta);
}

int rla_sparse_apply_bell_split_bytes(int64_t rows, int64_t in_rows, int64_t rb, int64_t n_rhs,
                                      size_t* bytes) {
  if (!bytes || rows < 0 || in_rows < 1 || rb < 1 || n_rhs < 0)
    return fail(RLA_EARG, "rla_sparse_apply_bell_split_bytes: bad arguments");
  // the shape depends on the element size: the larger of the two
  *bytes = std::max(bell_split_bytes(rows, in_rows, rb, n_rhs, sizeof(float)),
                    bell_split_bytes(rows, in_rows, rb, n_rhs, sizeof(double)));
  return RLA_OK;
}

int rla_sparse_apply_bell_f64(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                              const int32_t* widths, const int32_t* bidx, const double* bval,
                              int64_t n_rhs, const double* X, int64_t x_rs, int64_t x_cs,
                              double* C, int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  return sparse_apply_bell<double>(ctx, rows, in_rows, rb, widths, bidx, bval, n_rhs, X, x_rs,
                                   x_cs, C, c_rs, c_cs, alpha, beta);
}

int rla_sparse_tile_shape(rla_ctx* ctx, int64_t rows, int64_t in_rows, const int32_t* ptr,
                          const int32_t* idx, int with_values, int64_t* rec_words, int64_t* nrec) {
  return tile_shape_query(ctx, rows, in_rows, ptr, idx, with_values, rec_words, nrec);
}

int rla_sparse_tile_build_f32(rla_ctx* ctx, int64_t rows, int64_t in_rows, const int32_t* ptr,
                              const int32_t* idx, const float* val, int with_values,
                              int64_t rec_words, uint32_t* tiles, double* magnitude) {
  return tile_build<float>(ctx, rows, in_rows, ptr, idx, val, with_values, rec_words, tiles,
                           magnitude);
}

int rla_sparse_apply_tile_bytes(int64_t rows, int64_t in_rows, int64_t n_rhs, size_t* bytes) {
  if (!bytes || rows < 0 || in_rows < 1 || n_rhs < 0)
    return fail(RLA_EARG, "rla_sparse_apply_tile_bytes: bad arguments");
  *bytes = sparse_tile_part_bytes(rows, in_rows, n_rhs);
  return RLA_OK;
}

int rla_sparse_apply_tile_f32(rla_ctx* ctx, int64_t rows, int64_t in_rows, int with_values,
                              int64_t rec_words, const uint32_t* tiles, double magnitude,
                              int64_t n_rhs, const float* X, int64_t x_rs, int64_t x_cs, float* C,
                              int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  return sparse_apply_tile_part(ctx, rows, in_rows, with_values, rec_words, tiles, magnitude,
                                n_rhs, X, x_rs, x_cs, C, c_rs, c_cs, alpha, beta, nullptr);
}

}  // extern "C"
