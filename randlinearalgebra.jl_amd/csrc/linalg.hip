// Dense drivers: FP64 CholeskyQR2, the RangeFinder power / orthogonalised
// iterations (src/Approximators/RangeApproximators/helpers/power_its.jl) and
// the RandSVD finish (randsvd.jl:174-189), on top of the BF16 MFMA GEMM
// (mixed precision) or vendor FP32/FP64 GEMMs (plain precision).
//
// QR: CholeskyQR2 in FP64 — G = Y'Y, R = chol(G), Y <- Y R^-1, twice.  For
// k <= 288 columns each pass is a few launches and no host round trip: the
// one-read Gram kernel, a one-workgroup Cholesky, and an in-place blocked
// triangular solve that is skipped on the device when the factorisation
// broke down.  Wider k uses the tiled FP64 GEMM with rocSOLVER's potrf/trtri.
// If the Cholesky factorisation breaks down (Y numerically rank deficient in
// FP64) the QR falls back to Householder (GEQRF + ORGQR) on the current Y,
// which spans the same space.  Q is returned with the column signs of a
// positive-diagonal R.  Inside the RangeFinder loop the intermediate iterates
// take one CholeskyQR pass and the returned Q two (rf_run).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <vector>

#include "rla_internal.hpp"

namespace rla {

int gemm_bf16_tn(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, float alpha, const uint16_t* A,
                 int64_t lda, const uint16_t* B, int64_t ldb, float beta, float* C, int64_t c_rs,
                 int64_t c_cs);
int gemm_bf16_tn_tiled(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, float alpha,
                       const uint16_t* At, const uint16_t* B, int64_t ldb, float beta, float* C,
                       int64_t c_rs, int64_t c_cs);
template <class T>
int to_bf16_rows(rla_ctx* ctx, int64_t rows, int64_t cols, const T* src, int64_t s_rs,
                 int64_t s_cs, uint16_t* dst, int64_t ldd);
int gemm_f64(rla_ctx* ctx, bool ta, bool tb, int64_t M, int64_t N, int64_t K, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
             int64_t ldc, double* part, size_t part_bytes);
size_t gemm_f64_split_bytes(int64_t M, int64_t N);
bool cholqr_small_k(int64_t k);
size_t gram_f64_part_bytes(int64_t m, int64_t k);
template <class T>
int gram_f64(rla_ctx* ctx, int64_t m, int64_t k, const T* Y, int64_t ldy, double* G,
             int64_t ldg, double* part, size_t part_bytes);
int chol_blocks(rla_ctx* ctx, int64_t k, const double* G, int64_t ldg, double* R, int64_t ldr,
                double* Rt, int64_t ldt, double* Dinv, int* info_dev, bool g_symmetric);
int trsm_ru_f64(rla_ctx* ctx, int64_t m, int64_t k, double* Y, int64_t ldy, const double* Rt,
                int64_t ldt, const double* Dinv, const int* info);
template <class Tin>
int trsm_ru(rla_ctx* ctx, int64_t m, int64_t k, const Tin* Yin, int64_t ldy, double* Yout,
            int64_t ldo, uint16_t* Ybf, int64_t ld_bf, const double* Rt, int64_t ldt,
            const double* Dinv, const int* info, int ninfo);

#define RLA_BLAS(call)                                                                       \
  do {                                                                                       \
    rocblas_status _s = (call);                                                              \
    if (_s != rocblas_status_success)                                                        \
      return ::rla::fail(RLA_EHIP, std::string(#call " failed: ") + rocblas_status_to_string(_s)); \
  } while (0)

int blas(rla_ctx* ctx, rocblas_handle* h) {
  if (!ctx->blas) {
    rocblas_handle nh = nullptr;
    RLA_BLAS(rocblas_create_handle(&nh));
    ctx->blas = nh;
  }
  *h = static_cast<rocblas_handle>(ctx->blas);
  RLA_BLAS(rocblas_set_stream(*h, ctx->stream));
  return RLA_OK;
}

void destroy_blas(rla_ctx* ctx) {
  if (ctx->blas) rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->blas));
  ctx->blas = nullptr;
}

// bump allocator over a caller-provided workspace
struct Bump {
  char* p;
  size_t left;
  size_t need = 0;
  template <class T>
  T* take(size_t n) {
    const size_t b = (n * sizeof(T) + 255) & ~size_t(255);
    need += b;
    if (!p || b > left) {
      left = 0;
      return nullptr;
    }
    T* r = reinterpret_cast<T*>(p);
    p += b;
    left -= b;
    return r;
  }
};

// dst = src * diag(d)
__global__ void scale_cols_kernel(int64_t rows, int64_t cols, const double* __restrict__ src,
                                  int64_t lds, const double* __restrict__ d,
                                  double* __restrict__ dst, int64_t ldd) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= rows * cols) return;
  const int64_t i = e % rows, j = e / rows;
  dst[i + j * ldd] = src[i + j * lds] * d[j];
}

template <class S, class D>
__global__ void convert_kernel(int64_t rows, int64_t cols, const S* __restrict__ src, int64_t lds,
                               D* __restrict__ dst, int64_t ldd) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= rows * cols) return;
  const int64_t i = e % rows, j = e / rows;
  dst[i + j * ldd] = static_cast<D>(src[i + j * lds]);
}

template <class S, class D>
static int convert(rla_ctx* ctx, int64_t rows, int64_t cols, const S* src, int64_t lds, D* dst,
                   int64_t ldd) {
  if (rows * cols == 0) return RLA_OK;
  convert_kernel<S, D><<<dim3(static_cast<unsigned>(ceil_div(rows * cols, 256))), 256, 0,
                         ctx->stream>>>(rows, cols, src, lds, dst, ldd);
  RLA_LAUNCH_CHECK("convert");
  return RLA_OK;
}

// Flip the signs of columns j with sg[j] < 0 (the Householder R's diagonal
// signs) so Q matches the positive-diagonal convention of the Cholesky path.
__global__ void fix_signs_kernel(int64_t m, int64_t k, double* Y, int64_t ldy,
                                 const double* sg) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= m * k) return;
  const int64_t i = e % m, j = e / m;
  if (sg[j] < 0) Y[i + j * ldy] = -Y[i + j * ldy];
}

__global__ void diag_sign_kernel(int64_t k, const double* R, int64_t ldr, double* sg) {
  const int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (j < k) sg[j] = R[j + j * ldr] < 0 ? -1.0 : 1.0;
}

__global__ void copy_upper_kernel(int64_t k, const double* src, int64_t lds, double* dst,
                                  int64_t ldd) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= k * k) return;
  const int64_t i = e % k, j = e / k;
  dst[i + j * ldd] = i <= j ? src[i + j * lds] : 0.0;
}

// dst = upper(R) with row i negated where R[i,i] < 0 (pairs with fix_signs)
__global__ void signed_upper_kernel(int64_t k, const double* R, int64_t ldr, double* dst,
                                    int64_t ldd) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= k * k) return;
  const int64_t i = e % k, j = e / k;
  const double v = i <= j ? R[i + j * ldr] : 0.0;
  dst[i + j * ldd] = R[i + i * ldr] < 0 ? -v : v;
}

struct QrBufs {
  double* G;     // k x k Gram, then the Cholesky factor (in place)
  double* Racc;  // k x k accumulated R
  double* X;     // k x k: R row-major (small k) or R^-1 (wide k)
  double* Dinv;  // inverses of R's 16 x 16 diagonal blocks (small k)
  double* tau;
  double* sg;
  int* info;     // 2 slots, one per pass
  double* T;     // m x k product temporary (wide-k path only)
  double* part;  // split-K / Gram partials
  size_t part_bytes;
};

static size_t qr_layout(int64_t m, int64_t k, void* work, size_t bytes, QrBufs* out) {
  Bump b{static_cast<char*>(work), bytes};
  QrBufs q{};
  q.G = b.take<double>(k * k);
  q.Racc = b.take<double>(k * k);
  q.X = b.take<double>(k * k);
  q.Dinv = b.take<double>(ceil_div(k, 16) * 256);
  q.tau = b.take<double>(k);
  q.sg = b.take<double>(k);
  q.info = b.take<int>(2);
  const bool small = cholqr_small_k(k);
  q.T = small ? nullptr : b.take<double>(m * k);
  q.part_bytes = small ? gram_f64_part_bytes(m, k) : gemm_f64_split_bytes(k, k);
  q.part = reinterpret_cast<double*>(b.take<char>(q.part_bytes));
  if (out) *out = q;
  return b.need;
}

size_t qr_work_bytes(int64_t m, int64_t k) { return qr_layout(m, k, nullptr, 0, nullptr); }

__global__ void zero_lower_kernel(int64_t k, double* R, int64_t ldr) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= k * k) return;
  const int64_t i = e % k, j = e / k;
  if (i > j) R[i + j * ldr] = 0.0;
}

// One CholeskyQR pass on Y: G = Y'Y, G <- chol(G) (upper, in place), Y <- Y G^-1.
// The small-k path leaves Y untouched on the device when *info != 0 (the
// solve is skipped there); the wide path checks on the host.
static int cholqr_pass(rla_ctx* ctx, int64_t m, int64_t k, double* Y, int64_t ldy, QrBufs& q,
                       int* info, bool* broke) {
  if (cholqr_small_k(k)) {
    int rc = gram_f64<double>(ctx, m, k, Y, ldy, q.G, k, q.part, q.part_bytes);
    if (rc) return rc;
    // gram_f64 writes both triangles of G
    rc = chol_blocks(ctx, k, q.G, k, q.G, k, q.X, k, q.Dinv, info, true);
    if (rc) return rc;
    return trsm_ru_f64(ctx, m, k, Y, ldy, q.X, k, q.Dinv, info);
  }
  rocblas_handle h;
  int rc = blas(ctx, &h);
  if (rc) return rc;
  const int ki = static_cast<int>(k);
  rc = gemm_f64(ctx, true, false, k, k, m, 1.0, Y, ldy, Y, ldy, 0.0, q.G, k, q.part,
                q.part_bytes);
  if (rc) return rc;
  RLA_BLAS(rocsolver_dpotrf(h, rocblas_fill_upper, ki, q.G, ki, info));
  int hinfo = 0;
  RLA_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (hinfo != 0) {
    *broke = true;
    return RLA_OK;
  }
  zero_lower_kernel<<<dim3(unsigned(ceil_div(k * k, 256))), 256, 0, ctx->stream>>>(k, q.G, k);
  RLA_HIP(hipMemcpyAsync(q.X, q.G, sizeof(double) * k * k, hipMemcpyDeviceToDevice,
                         ctx->stream));
  RLA_BLAS(rocsolver_dtrtri(h, rocblas_fill_upper, rocblas_diagonal_non_unit, ki, q.X, ki, info));
  rc = gemm_f64(ctx, false, false, m, k, k, 1.0, Y, ldy, q.X, k, 0.0, q.T, m, nullptr, 0);
  if (rc) return rc;
  return convert<double, double>(ctx, m, k, q.T, m, Y, ldy);
}

// One CholeskyQR pass with separate input and output, for the mixed-precision
// RangeFinder: Yin (FP64, or an FP32 product widened exactly on load) gives G,
// R and Yin R^-1, which goes to Yout (FP64) or to Ybf (its BF16 rounding, the
// next product's operand).  This pass's breakdown flag is info[nflags - 1];
// the solve writes nothing unless all of info[0 .. nflags) are 0.
template <class Tin>
static int cholqr_to(rla_ctx* ctx, int64_t m, int64_t k, const Tin* Yin, int64_t ldy,
                     double* Yout, int64_t ldo, uint16_t* Ybf, int64_t ld_bf, QrBufs& q,
                     int* info, int nflags) {
  int rc = gram_f64<Tin>(ctx, m, k, Yin, ldy, q.G, k, q.part, q.part_bytes);
  if (rc) return rc;
  rc = chol_blocks(ctx, k, q.G, k, q.G, k, q.X, k, q.Dinv, info + nflags - 1, true);
  if (rc) return rc;
  return trsm_ru<Tin>(ctx, m, k, Yin, ldy, Yout, ldo, Ybf, ld_bf, q.X, k, q.Dinv, info, nflags);
}

// Y (m x k, ldy) <- Q; optionally R (k x k upper, ldr) with Y_in = Q R.
// info_async == nullptr: the call checks each pass on the host and falls back
// to Householder QR itself.  Otherwise (small k, R == nullptr) nothing is
// synchronised: the passes' breakdown flags land in info_async[0..passes-1]
// and a failed pass leaves Y as it was before that pass; the caller checks
// later.  passes = 2 is CholeskyQR2; passes = 1 gives a basis of the same
// range whose orthogonality error is about cond(Y)^2 * eps.
int qr_f64(rla_ctx* ctx, int64_t m, int64_t k, double* Y, int64_t ldy, double* R, int64_t ldr,
           void* work, size_t work_bytes, bool* used_householder, int* info_async,
           int passes = 2) {
  if (used_householder) *used_householder = false;
  if (k == 0 || m == 0) return RLA_OK;
  if (m < k) return fail(RLA_EARG, "QR of a wide matrix (m < k) is not supported");
  if (m > INT32_MAX || k > INT32_MAX) return fail(RLA_ENOTSUP, "QR dimensions exceed int32");
  QrBufs q;
  if (qr_layout(m, k, work, work_bytes, &q) > work_bytes) return fail(RLA_EARG, "QR workspace too small");
  const bool small = cholqr_small_k(k);
  if (passes != 1 && passes != 2) return fail(RLA_EARG, "CholeskyQR takes 1 or 2 passes");
  if (info_async && small && !R) {
    for (int pass = 0; pass < passes; ++pass) {
      bool broke = false;
      int rc = cholqr_pass(ctx, m, k, Y, ldy, q, info_async + pass, &broke);
      if (rc) return rc;
    }
    return RLA_OK;
  }
  rocblas_handle h;
  int rc = blas(ctx, &h);
  if (rc) return rc;
  const double one = 1.0;
  const int mi = static_cast<int>(m), ki = static_cast<int>(k), li = static_cast<int>(ldy);
  int passes_done = 0;
  bool breakdown = false;
  for (int pass = 0; pass < passes && !breakdown; ++pass) {
    rc = cholqr_pass(ctx, m, k, Y, ldy, q, q.info + pass, &breakdown);
    if (rc) return rc;
    if (small) {
      int hinfo = 0;
      RLA_HIP(hipMemcpyAsync(&hinfo, q.info + pass, sizeof(int), hipMemcpyDeviceToHost,
                             ctx->stream));
      RLA_HIP(hipStreamSynchronize(ctx->stream));
      breakdown = hinfo != 0;  // Y was left as it was before this pass
    }
    if (breakdown) break;
    if (R) {
      if (pass == 0) {
        copy_upper_kernel<<<dim3(unsigned(ceil_div(k * k, 256))), 256, 0, ctx->stream>>>(
            k, q.G, k, q.Racc, k);
      } else {
        // Racc <- R2 * Racc
        RLA_BLAS(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                               rocblas_diagonal_non_unit, ki, ki, &one, q.G, ki, q.Racc, ki,
                               q.Racc, ki));
      }
    }
    ++passes_done;
  }
  if (breakdown) {
    if (used_householder) *used_householder = true;
    RLA_BLAS(rocsolver_dgeqrf(h, mi, ki, Y, li, q.tau));
    // D = sign(diag(R_h)); keep D*R_h (positive diagonal) in G, D in sg
    diag_sign_kernel<<<dim3(unsigned(ceil_div(k, 256))), 256, 0, ctx->stream>>>(k, Y, ldy, q.sg);
    signed_upper_kernel<<<dim3(unsigned(ceil_div(k * k, 256))), 256, 0, ctx->stream>>>(
        k, Y, ldy, q.G, k);
    RLA_BLAS(rocsolver_dorgqr(h, mi, ki, ki, Y, li, q.tau));
    // Q = Q_h * D, so Q * (D * R_h) = Q_h * R_h
    fix_signs_kernel<<<dim3(unsigned(ceil_div(m * k, 256))), 256, 0, ctx->stream>>>(m, k, Y, ldy,
                                                                                  q.sg);
    if (R) {
      if (passes_done == 0) {
        copy_upper_kernel<<<dim3(unsigned(ceil_div(k * k, 256))), 256, 0, ctx->stream>>>(
            k, q.G, k, q.Racc, k);
      } else {
        RLA_BLAS(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                               rocblas_diagonal_non_unit, ki, ki, &one, q.G, ki, q.Racc, ki,
                               q.Racc, ki));
      }
    }
  }
  if (R) {
    copy_upper_kernel<<<dim3(unsigned(ceil_div(k * k, 256))), 256, 0, ctx->stream>>>(
        k, q.Racc, k, R, ldr);
  }
  RLA_LAUNCH_CHECK("qr_f64");
  return RLA_OK;
}

// ---- RangeFinder -------------------------------------------------------------

// BF16 products read the resident BF16 copies of A: rows-contiguous
// (PREC_BF16) or in the GEMM's tiled layout (PREC_BF16_TILED, the same
// products in the same order)
enum { PREC_PLAIN = 0, PREC_BF16 = 1, PREC_BF16_TILED = 2 };

static bool is_bf16(int precision) {
  return precision == PREC_BF16 || precision == PREC_BF16_TILED;
}

// C (M x N, FP32) = Acopy (M x K, a BF16 copy of A or A') * X, X BF16 with
// K-contiguous columns (ld ldx)
static int bf16_product(rla_ctx* ctx, int precision, int64_t M, int64_t N, int64_t K,
                        const void* Acopy, int64_t ld_a, const uint16_t* X, int64_t ldx, float* C,
                        int64_t c_rs, int64_t c_cs) {
  const uint16_t* a = static_cast<const uint16_t*>(Acopy);
  if (precision == PREC_BF16_TILED)
    return gemm_bf16_tn_tiled(ctx, M, N, K, 1.f, a, X, ldx, 0.f, C, c_rs, c_cs);
  return gemm_bf16_tn(ctx, M, N, K, 1.f, a, ld_a, X, ldx, 0.f, C, c_rs, c_cs);
}

static int64_t ld8(int64_t n) { return (n + 7) & ~int64_t(7); }

struct RfBufs {
  double* Z64;    // n x k
  float* Y32;     // m x k  (plain f32 / mixed GEMM output)
  float* Z32;     // n x k
  uint16_t* Ybf;  // m x k, ld8(m)
  uint16_t* Zbf;  // n x k, ld8(n)
  double* Ybak;   // m x k copy of the sketch, for a rerun after a QR breakdown
  int* info;      // Cholesky breakdown flags, 2 per QR of the run
  int ninfo;
  void* qr;
  size_t qr_bytes;
};

static size_t rf_layout(int precision, int a_elem, int64_t m, int64_t n, int64_t k, void* work,
                        size_t bytes, RfBufs* out) {
  Bump b{static_cast<char*>(work), bytes};
  RfBufs r{};
  r.Z64 = b.take<double>(n * k);
  if (is_bf16(precision) || a_elem == 4) {
    r.Y32 = b.take<float>(m * k);
    r.Z32 = b.take<float>(n * k);
  }
  if (is_bf16(precision)) {
    r.Ybf = b.take<uint16_t>(ld8(m) * k);
    r.Zbf = b.take<uint16_t>(ld8(n) * k);
  }
  r.Ybak = b.take<double>(m * k);
  r.ninfo = 2 * 64;
  r.info = b.take<int>(r.ninfo);
  r.qr_bytes = qr_work_bytes(std::max(m, n), k);
  r.qr = b.take<char>(r.qr_bytes);
  if (out) *out = r;
  return b.need;
}

// Z (n x k, f64) = A' * Y (Y m x k, f64).  At: for bf16 the columns-
// contiguous BF16 copy of A; for plain, NULL or a copy of A in A's type with
// rows contiguous (row stride ld_t), i.e. A' column-major, which makes A'Y a
// plain NN product instead of a transposed read of A.
static int rf_at_y(rla_ctx* ctx, int precision, int a_elem, int64_t m, int64_t n, int64_t k,
                   const void* A, int64_t lda, const void* At, int64_t ld_t,
                   const double* Y, int64_t ldy, RfBufs& w) {
  rocblas_handle h;
  int rc;
  if (is_bf16(precision)) {
    rc = to_bf16_rows<double>(ctx, k, m, Y, ldy, 1, w.Ybf, ld8(m));
    if (rc) return rc;
    rc = bf16_product(ctx, precision, n, k, m, At, ld_t, w.Ybf, ld8(m), w.Z32, 1, n);
    if (rc) return rc;
    return convert<float, double>(ctx, n, k, w.Z32, n, w.Z64, n);
  }
  rc = blas(ctx, &h);
  if (rc) return rc;
  const rocblas_operation op_a = At ? rocblas_operation_none : rocblas_operation_transpose;
  const void* Aop = At ? At : A;
  const int64_t ld_op = At ? ld_t : lda;
  if (a_elem == 8) {
    const double one = 1.0, zero = 0.0;
    RLA_BLAS(rocblas_dgemm(h, op_a, rocblas_operation_none, int(n), int(k), int(m), &one,
                           static_cast<const double*>(Aop), int(ld_op), Y, int(ldy), &zero,
                           w.Z64, int(n)));
    return RLA_OK;
  }
  const float one = 1.f, zero = 0.f;
  rc = convert<double, float>(ctx, m, k, Y, ldy, w.Y32, m);
  if (rc) return rc;
  RLA_BLAS(rocblas_sgemm(h, op_a, rocblas_operation_none, int(n), int(k), int(m), &one,
                         static_cast<const float*>(Aop), int(ld_op), w.Y32, int(m), &zero,
                         w.Z32, int(n)));
  return convert<float, double>(ctx, n, k, w.Z32, n, w.Z64, n);
}

// Y (m x k, f64) = A * Z (Z n x k, f64 in w.Z64)
static int rf_a_z(rla_ctx* ctx, int precision, int a_elem, int64_t m, int64_t n, int64_t k,
                  const void* A, int64_t lda, const void* A_rm, int64_t ld_rm, double* Y,
                  int64_t ldy, RfBufs& w) {
  rocblas_handle h;
  int rc;
  if (is_bf16(precision)) {
    rc = to_bf16_rows<double>(ctx, k, n, w.Z64, n, 1, w.Zbf, ld8(n));
    if (rc) return rc;
    rc = bf16_product(ctx, precision, m, k, n, A_rm, ld_rm, w.Zbf, ld8(n), w.Y32, 1, m);
    if (rc) return rc;
    return convert<float, double>(ctx, m, k, w.Y32, m, Y, ldy);
  }
  rc = blas(ctx, &h);
  if (rc) return rc;
  if (a_elem == 8) {
    const double one = 1.0, zero = 0.0;
    RLA_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, int(m), int(k),
                           int(n), &one, static_cast<const double*>(A), int(lda), w.Z64, int(n),
                           &zero, Y, int(ldy)));
    return RLA_OK;
  }
  const float one = 1.f, zero = 0.f;
  rc = convert<double, float>(ctx, n, k, w.Z64, n, w.Z32, n);
  if (rc) return rc;
  RLA_BLAS(rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, int(m), int(k),
                         int(n), &one, static_cast<const float*>(A), int(lda), w.Z32, int(n),
                         &zero, w.Y32, int(m)));
  return convert<float, double>(ctx, m, k, w.Y32, m, Y, ldy);
}

// rand_ortho_it (power_its.jl:53-80) on the BF16 path with device-side
// breakdown flags (power_its >= 1, k <= 288).  Each CholeskyQR pass reads the
// product's FP32 output directly and writes only what comes next: the BF16
// operand of the next product, or, for the returned Q, the first pass of
// CholeskyQR2 into Ybak and the second into Y.  The arithmetic is that of
// rf_run's passes (the same Gram, factor and solve on exactly widened
// values, the same BF16 rounding), so Q is bit-identical; the conversion
// kernels between the passes and the products are gone.  Y holds the sketch
// until the last solve, which writes nothing if any pass broke down, so a
// rerun needs no saved copy.
static int rf_run_bf16_fused(rla_ctx* ctx, int precision, int64_t m, int64_t n, int64_t k,
                             int64_t power_its,
                             const void* A_rm, int64_t ld_rm, const void* A_cm, int64_t ld_cm,
                             double* Y, int64_t ldy, RfBufs& w, int* info) {
  QrBufs q;
  if (qr_layout(std::max(m, n), k, w.qr, w.qr_bytes, &q) > w.qr_bytes)
    return fail(RLA_EARG, "RangeFinder workspace too small");
  int ns = 0;
  int rc = cholqr_to<double>(ctx, m, k, Y, ldy, nullptr, 0, w.Ybf, ld8(m), q, info, ++ns);
  if (rc) return rc;
  for (int64_t it = 0; it < power_its; ++it) {
    rc = bf16_product(ctx, precision, n, k, m, A_cm, ld_cm, w.Ybf, ld8(m), w.Z32, 1, n);
    if (rc) return rc;
    rc = cholqr_to<float>(ctx, n, k, w.Z32, n, nullptr, 0, w.Zbf, ld8(n), q, info, ++ns);
    if (rc) return rc;
    rc = bf16_product(ctx, precision, m, k, n, A_rm, ld_rm, w.Zbf, ld8(n), w.Y32, 1, m);
    if (rc) return rc;
    if (it + 1 < power_its) {
      rc = cholqr_to<float>(ctx, m, k, w.Y32, m, nullptr, 0, w.Ybf, ld8(m), q, info, ++ns);
      if (rc) return rc;
      continue;
    }
    rc = cholqr_to<float>(ctx, m, k, w.Y32, m, w.Ybak, m, nullptr, 0, q, info, ++ns);
    if (rc) return rc;
    rc = cholqr_to<double>(ctx, m, k, w.Ybak, m, Y, ldy, nullptr, 0, q, info, ++ns);
    if (rc) return rc;
  }
  return RLA_OK;
}

static int check_rf(int precision, int a_elem, int64_t m, int64_t n, int64_t k, const void* A,
                    const void* A_rm, const void* A_cm) {
  if (precision != PREC_PLAIN && !is_bf16(precision))
    return fail(RLA_EARG,
                "precision must be 0 (plain), 1 (bf16 products) or 2 (bf16, tiled copies)");
  if (a_elem != 4 && a_elem != 8) return fail(RLA_EARG, "A element size must be 4 or 8");
  if (m < 1 || n < 1 || k < 1) return fail(RLA_EARG, "RangeFinder dimensions must be positive");
  if (k > m) return fail(RLA_EARG, "compression dimension larger than the row count");
  if (m > INT32_MAX || n > INT32_MAX) return fail(RLA_ENOTSUP, "dimensions exceed int32");
  if (precision == PREC_PLAIN && !A) return fail(RLA_EARG, "plain precision needs A");
  if (is_bf16(precision) && (!A_rm || !A_cm))
    return fail(RLA_EARG, "bf16 precision needs both BF16 copies of A");
  return RLA_OK;
}

// the plain path's optional rows-contiguous copy of A (row stride ld >= n)
static int check_rows_copy(int precision, const void* At, int64_t n, int64_t ld) {
  if (precision == PREC_PLAIN && At && ld < n)
    return fail(RLA_EARG, "the row copy of A needs a row stride >= n");
  return RLA_OK;
}

// Breakdown flag slots of one solve: one per CholeskyQR pass (rf_run).
static int64_t rf_info_slots(int orthogonalize, int64_t power_its) {
  return orthogonalize ? 2 * power_its + 2 : 2;
}

// The iterations of power_its.jl.  info == nullptr: QRs check breakdown on
// the host; otherwise every CholeskyQR pass records its flag in the next
// slot of info (rf_info_slots in all).
//
// rand_ortho_it orthonormalises after every product so that the next product
// starts from a well-conditioned basis of the current range; the range of
// Y R^-1 is that of Y whatever the accuracy of R, so the intermediate
// iterates take one CholeskyQR pass (orthogonality error ~ cond(Y)^2 * eps;
// a pivot below the breakdown floor reruns the solve with checked QRs) and
// only the returned Q takes the second pass of CholeskyQR2.
static int rf_run(rla_ctx* ctx, int precision, int orthogonalize, int64_t m, int64_t n,
                  int64_t k, int64_t power_its, int a_elem, const void* A, int64_t lda,
                  const void* A_rm, int64_t ld_rm, const void* A_cm, int64_t ld_cm,
                  double* Y, int64_t ldy, RfBufs& w, int* info) {
  int rc;
  // the copy that A'X reads: bf16 columns-contiguous, or plain rows-contiguous
  const void* At = is_bf16(precision) ? A_cm : A_rm;
  const int64_t ld_t = is_bf16(precision) ? ld_cm : ld_rm;
  int64_t ns = 0;
  auto qr = [&](int64_t rows, double* M, int64_t ld, bool last) {
    const int passes = last ? 2 : 1;
    int* slot = info ? info + ns : nullptr;
    ns += passes;
    return qr_f64(ctx, rows, k, M, ld, nullptr, 0, w.qr, w.qr_bytes, nullptr, slot, passes);
  };
  if (orthogonalize) {
    // rand_ortho_it: QR after every product (power_its.jl:53-80)
    rc = qr(m, Y, ldy, power_its == 0);
    if (rc) return rc;
    for (int64_t it = 0; it < power_its; ++it) {
      rc = rf_at_y(ctx, precision, a_elem, m, n, k, A, lda, At, ld_t, Y, ldy, w);
      if (rc) return rc;
      rc = qr(n, w.Z64, n, false);
      if (rc) return rc;
      rc = rf_a_z(ctx, precision, a_elem, m, n, k, A, lda, A_rm, ld_rm, Y, ldy, w);
      if (rc) return rc;
      rc = qr(m, Y, ldy, it + 1 == power_its);
      if (rc) return rc;
    }
    return RLA_OK;
  }
  // rand_power_it: products only, one QR at the end (power_its.jl:15-36)
  for (int64_t it = 0; it < power_its; ++it) {
    rc = rf_at_y(ctx, precision, a_elem, m, n, k, A, lda, At, ld_t, Y, ldy, w);
    if (rc) return rc;
    rc = rf_a_z(ctx, precision, a_elem, m, n, k, A, lda, A_rm, ld_rm, Y, ldy, w);
    if (rc) return rc;
  }
  return qr(m, Y, ldy, true);
}

}  // namespace rla

using namespace rla;

extern "C" {

int rla_gemm_f64(rla_ctx* ctx, int trans_a, int trans_b, int64_t m, int64_t n, int64_t k,
                 double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                 double beta, double* C, int64_t ldc) {
  RLA_CHECK_CTX(ctx);
  if (m == 0 || n == 0) return RLA_OK;
  int rc = use_device(ctx);
  if (rc) return rc;
  rocblas_handle h;
  rc = blas(ctx, &h);
  if (rc) return rc;
  RLA_BLAS(rocblas_dgemm(h, trans_a ? rocblas_operation_transpose : rocblas_operation_none,
                         trans_b ? rocblas_operation_transpose : rocblas_operation_none, int(m),
                         int(n), int(k), &alpha, A, int(lda), B, int(ldb), &beta, C, int(ldc)));
  return RLA_OK;
}

int rla_gemm_f32(rla_ctx* ctx, int trans_a, int trans_b, int64_t m, int64_t n, int64_t k,
                 double alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
                 double beta, float* C, int64_t ldc) {
  RLA_CHECK_CTX(ctx);
  if (m == 0 || n == 0) return RLA_OK;
  int rc = use_device(ctx);
  if (rc) return rc;
  rocblas_handle h;
  rc = blas(ctx, &h);
  if (rc) return rc;
  const float a = float(alpha), b = float(beta);
  RLA_BLAS(rocblas_sgemm(h, trans_a ? rocblas_operation_transpose : rocblas_operation_none,
                         trans_b ? rocblas_operation_transpose : rocblas_operation_none, int(m),
                         int(n), int(k), &a, A, int(lda), B, int(ldb), &b, C, int(ldc)));
  return RLA_OK;
}

int rla_qr_workspace_bytes(int64_t m, int64_t k, size_t* bytes) {
  if (!bytes || k < 0 || m < 0) return fail(RLA_EARG, "rla_qr_workspace_bytes: bad arguments");
  *bytes = qr_work_bytes(m, k);
  return RLA_OK;
}

// x <- op(R)^-1 x for an n x n upper triangular R (vendor trsv)
int rla_trsv_upper_f64(rla_ctx* ctx, int trans, int64_t n, const double* R, int64_t ldr,
                       double* x) {
  RLA_CHECK_CTX(ctx);
  if (n < 0 || ldr < std::max<int64_t>(1, n)) return fail(RLA_EARG, "rla_trsv_upper_f64: bad sizes");
  if (n == 0) return RLA_OK;
  if (!R || !x) return fail(RLA_EARG, "rla_trsv_upper_f64: NULL pointer");
  if (n > INT32_MAX || ldr > INT32_MAX) return fail(RLA_ENOTSUP, "dimensions exceed int32");
  int rc = use_device(ctx);
  if (rc) return rc;
  rocblas_handle h;
  rc = blas(ctx, &h);
  if (rc) return rc;
  RLA_BLAS(rocblas_dtrsv(h, rocblas_fill_upper,
                         trans ? rocblas_operation_transpose : rocblas_operation_none,
                         rocblas_diagonal_non_unit, static_cast<int>(n), R,
                         static_cast<int>(ldr), x, 1));
  return RLA_OK;
}

int rla_qr_f64(rla_ctx* ctx, int64_t m, int64_t k, double* Y, int64_t ldy, double* R,
               int64_t ldr, void* work, size_t work_bytes, int* used_householder) {
  RLA_CHECK_CTX(ctx);
  int rc = use_device(ctx);
  if (rc) return rc;
  bool hh = false;
  rc = qr_f64(ctx, m, k, Y, ldy, R, ldr, work, work_bytes, &hh, nullptr);
  if (used_householder) *used_householder = hh ? 1 : 0;
  return rc;
}

int rla_rangefinder_workspace_bytes(int precision, int a_elem, int64_t m, int64_t n, int64_t k,
                                    size_t* bytes) {
  if (!bytes) return fail(RLA_EARG, "bytes is NULL");
  *bytes = rf_layout(precision, a_elem, m, n, k, nullptr, 0, nullptr);
  return RLA_OK;
}

int rla_rangefinder(rla_ctx* ctx, int precision, int orthogonalize, int64_t m, int64_t n,
                    int64_t k, int64_t power_its, int a_elem, const void* A, int64_t lda,
                    const void* A_rm, int64_t ld_rm, const void* A_cm, int64_t ld_cm,
                    double* Y, int64_t ldy, void* work, size_t work_bytes) {
  RLA_CHECK_CTX(ctx);
  int rc = check_rf(precision, a_elem, m, n, k, A, A_rm, A_cm);
  if (rc) return rc;
  rc = check_rows_copy(precision, A_rm, n, ld_rm);
  if (rc) return rc;
  if (power_its < 0) return fail(RLA_EARG, "Field `power_its` must be non-negative.");
  rc = use_device(ctx);
  if (rc) return rc;
  RfBufs w;
  if (rf_layout(precision, a_elem, m, n, k, work, work_bytes, &w) > work_bytes)
    return fail(RLA_EARG, "RangeFinder workspace too small");
  const int64_t nslots = rf_info_slots(orthogonalize, power_its);
  const bool async = cholqr_small_k(k) && nslots <= w.ninfo;
  auto flags_ok = [&](bool* ok) -> int {
    std::vector<int> h(nslots);
    RLA_HIP(hipMemcpyAsync(h.data(), w.info, sizeof(int) * nslots, hipMemcpyDeviceToHost,
                           ctx->stream));
    RLA_HIP(hipStreamSynchronize(ctx->stream));
    *ok = true;
    for (int v : h) *ok = *ok && v == 0;
    return RLA_OK;
  };
  if (async && is_bf16(precision) && orthogonalize && power_its >= 1) {
    // Y is left as it was unless every pass succeeded: a rerun starts from it
    rc = rf_run_bf16_fused(ctx, precision, m, n, k, power_its, A_rm, ld_rm, A_cm, ld_cm, Y, ldy, w,
                           w.info);
    if (rc) return rc;
    bool ok = false;
    rc = flags_ok(&ok);
    if (rc || ok) return rc;
  } else if (async) {
    // no host round trip inside the loop; a breakdown anywhere reruns the
    // whole solve with host-checked QRs (Householder fallback) from the copy
    RLA_HIP(hipMemcpy2DAsync(w.Ybak, sizeof(double) * m, Y, sizeof(double) * ldy,
                             sizeof(double) * m, k, hipMemcpyDeviceToDevice, ctx->stream));
    rc = rf_run(ctx, precision, orthogonalize, m, n, k, power_its, a_elem, A, lda, A_rm, ld_rm,
                A_cm, ld_cm, Y, ldy, w, w.info);
    if (rc) return rc;
    bool ok = false;
    rc = flags_ok(&ok);
    if (rc || ok) return rc;
    RLA_HIP(hipMemcpy2DAsync(Y, sizeof(double) * ldy, w.Ybak, sizeof(double) * m,
                             sizeof(double) * m, k, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return rf_run(ctx, precision, orthogonalize, m, n, k, power_its, a_elem, A, lda, A_rm, ld_rm,
                A_cm, ld_cm, Y, ldy, w, nullptr);
}

// RandSVD finish (randsvd.jl:174-189): with B = Q'A (k x n),
// B' = A'Q = Qb Rb (CholeskyQR2), Rb' = Ur S Vr' (small SVD, FP64),
// so A ~ (Q Ur) S (Qb Vr)'.  Outputs U (m x k), S (k), V (n x k).
int rla_randsvd_workspace_bytes(int precision, int a_elem, int64_t m, int64_t n, int64_t k,
                                size_t* bytes) {
  if (!bytes) return fail(RLA_EARG, "bytes is NULL");
  Bump b{nullptr, 0};
  b.take<double>(k * k);  // Rb
  b.take<double>(k * k);  // Ur
  b.take<double>(k * k);  // Vr
  b.take<char>(svd_small_work_bytes(k));
  *bytes = b.need + rf_layout(precision, a_elem, m, n, k, nullptr, 0, nullptr);
  return RLA_OK;
}

int rla_randsvd_finish(rla_ctx* ctx, int precision, int64_t m, int64_t n, int64_t k, int a_elem,
                       const void* A, int64_t lda, const void* A_cm, int64_t ld_cm,
                       const double* Q, int64_t ldq, double* U, int64_t ldu, double* S, double* V,
                       int64_t ldv, void* work, size_t work_bytes) {
  RLA_CHECK_CTX(ctx);
  int rc = check_rf(precision, a_elem, m, n, k, A, A_cm, A_cm);
  if (rc) return rc;
  rc = check_rows_copy(precision, A_cm, n, ld_cm);
  if (rc) return rc;
  if (k > n) return fail(RLA_EARG, "RandSVD needs k <= n");
  if (k > 288) return fail(RLA_ENOTSUP, "RandSVD: the small SVD takes k <= 288");
  rc = use_device(ctx);
  if (rc) return rc;
  Bump b{static_cast<char*>(work), work_bytes};
  double* Rb = b.take<double>(k * k);
  double* Ur = b.take<double>(k * k);
  double* Vr = b.take<double>(k * k);
  const size_t svd_bytes = svd_small_work_bytes(k);
  void* svd_work = b.take<char>(svd_bytes);
  RfBufs w;
  if (!svd_work || rf_layout(precision, a_elem, m, n, k, b.p, b.left, &w) > b.left)
    return fail(RLA_EARG, "RandSVD workspace too small");
  // B' = A'Q -> w.Z64 (n x k), then V <- B' and QR in place: B' = Qb Rb
  rc = rf_at_y(ctx, precision, a_elem, m, n, k, A, lda, A_cm, ld_cm, Q, ldq, w);
  if (rc) return rc;
  rc = convert<double, double>(ctx, n, k, w.Z64, n, V, ldv);
  if (rc) return rc;
  rc = qr_f64(ctx, n, k, V, ldv, Rb, k, w.qr, w.qr_bytes, nullptr, nullptr);
  if (rc) return rc;
  // Rb' = Ur S Vr' (one-sided Jacobi, FP64), so B = Rb' Qb' = Ur S (Qb Vr)'
  rc = svd_small_f64(ctx, k, Rb, k, 1, Ur, k, S, Vr, k, svd_work, svd_bytes, nullptr);
  if (rc) return rc;
  // U = Q Ur (m x k); V = Qb Vr, through w.Z64 since V holds Qb
  rc = gemm_mfma<double>(ctx, m, k, k, 1.0, Q, 1, ldq, Ur, 1, k, 0.0, U, 1, ldu);
  if (rc) return rc;
  rc = gemm_mfma<double>(ctx, n, k, k, 1.0, V, 1, ldv, Vr, 1, k, 0.0, w.Z64, 1, n);
  if (rc) return rc;
  return convert<double, double>(ctx, n, k, w.Z64, n, V, ldv);
}

}  // extern "C"

extern "C" {

int rla_convert_f32_f64(rla_ctx* ctx, int64_t rows, int64_t cols, const float* src, int64_t lds,
                        double* dst, int64_t ldd) {
  RLA_CHECK_CTX(ctx);
  int rc = use_device(ctx);
  if (rc) return rc;
  return rla::convert<float, double>(ctx, rows, cols, src, lds, dst, ldd);
}

int rla_convert_f64_f32(rla_ctx* ctx, int64_t rows, int64_t cols, const double* src, int64_t lds,
                        float* dst, int64_t ldd) {
  RLA_CHECK_CTX(ctx);
  int rc = use_device(ctx);
  if (rc) return rc;
  return rla::convert<double, float>(ctx, rows, cols, src, lds, dst, ldd);
}

int rla_scale_cols_f64(rla_ctx* ctx, int64_t rows, int64_t cols, const double* src, int64_t lds,
                       const double* d, double* dst, int64_t ldd) {
  RLA_CHECK_CTX(ctx);
  if (rows <= 0 || cols <= 0) return RLA_OK;
  int rc = use_device(ctx);
  if (rc) return rc;
  const int64_t n = rows * cols;
  rla::scale_cols_kernel<<<dim3(unsigned(rla::ceil_div(n, 256))), 256, 0, ctx->stream>>>(
      rows, cols, src, lds, d, dst, ldd);
  RLA_LAUNCH_CHECK("scale_cols");
  return RLA_OK;
}

}  // extern "C"
