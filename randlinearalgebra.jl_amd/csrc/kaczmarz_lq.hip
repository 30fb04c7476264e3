// Block Kaczmarz projection for any block size and Float32 / Float64 data:
// the LQ sub-solve of lq.jl:35-46 inside kaczmarz_update_block!
// (kaczmarz.jl:294-311):
//   skip if At == 0;  r = bt - At x;  At' = Q R;  x += alpha Q (R' \ r),
// the minimum-norm correction of the block system At u = r.  At' is copied
// to Float64 and factored by the library's QR (FP64 CholeskyQR2, Householder
// when it breaks down), then z = R' \ r by forward substitution.  A row
// dependent on earlier rows of the block (a pivot of R not above
// sqrt(s eps) max|R_jj|) is dropped and the block factored again without it,
// so the correction projects onto the independent rows.  The products are
// the FP64 MFMA GEMM.  Used for blocks the one-launch device
// step does not take (Float64 data, more than 64 rows).
//
// Also the FP64 residual ||b - A x|| for Float64 data (full_residual.jl:32-41).
#include <algorithm>
#include <cmath>
#include <vector>

#include "rla_internal.hpp"

namespace rla {

size_t qr_work_bytes(int64_t m, int64_t k);
int qr_f64(rla_ctx* ctx, int64_t m, int64_t k, double* Y, int64_t ldy, double* R, int64_t ldr,
           void* work, size_t work_bytes, bool* used_householder, int* info_async,
           int passes = 2);

namespace {

// Qt(i, j) = At(rows[j], i) in Float64 (n x s, ld n; rows NULL: row j);
// *nonzero set when any entry is
template <class T>
__global__ void lq_pack_kernel(int64_t s, int64_t n, const T* __restrict__ At, int64_t at_rs,
                               int64_t at_cs, const int* __restrict__ rows,
                               double* __restrict__ Qt, int* nonzero) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= s * n) return;
  const int64_t i = e % n, j = e / n;
  const int64_t jr = rows ? rows[j] : j;
  const double v = static_cast<double>(At[jr * at_rs + i * at_cs]);
  Qt[e] = v;
  if (v != 0.0) atomicOr(nonzero, 1);
}

// r2[j] = r[rows[j]]
__global__ void lq_gather_kernel(int64_t s, const int* __restrict__ rows,
                                 const double* __restrict__ r, double* __restrict__ r2) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e < s) r2[e] = r[rows[e]];
}

template <class T>
__global__ void lq_vec_in_kernel(int64_t s, const T* __restrict__ bt, int64_t bt_inc,
                                 double* __restrict__ r, int64_t n, const T* __restrict__ x,
                                 double* __restrict__ x64, int* nonzero) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e == 0) *nonzero = 0;
  if (e < s) r[e] = static_cast<double>(bt[e * bt_inc]);
  if (e < n) x64[e] = static_cast<double>(x[e]);
}

// R' z = r (R upper s x s, ld ldr): column-oriented forward substitution in
// one workgroup; a pivot not above sqrt(s eps) max|R_jj| gives z_j = 0.  r is
// overwritten.
__global__ __launch_bounds__(256) void lq_lower_solve_kernel(int64_t s, const double* __restrict__ R,
                                                             int64_t ldr, double* r,
                                                             double* __restrict__ z) {
  __shared__ double red[256];
  double mx = 0.0;
  for (int64_t j = threadIdx.x; j < s; j += 256) mx = fmax(mx, fabs(R[j + j * ldr]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  // R comes from CholeskyQR when that does not break down: a dependent row
  // then leaves R_jj at about sqrt(eps) of the scale, not at eps
  const double floor = sqrt(double(s) * 2.220446049250313e-16) * red[0];
  for (int64_t j = 0; j < s; ++j) {
    const double d = R[j + j * ldr];
    const double zj = fabs(d) > floor ? r[j] / d : 0.0;
    if (threadIdx.x == 0) z[j] = zj;
    for (int64_t i = j + 1 + threadIdx.x; i < s; i += 256) r[i] -= R[j + i * ldr] * zj;
    __syncthreads();
  }
}

template <class T>
__global__ void lq_axpy_kernel(int64_t n, double alpha, const double* __restrict__ u,
                               T* __restrict__ x) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e < n) x[e] = static_cast<T>(static_cast<double>(x[e]) + alpha * u[e]);
}

// r_i^2 partial sums per workgroup and the final FP64 norm
__global__ __launch_bounds__(256) void sumsq_kernel(int64_t n, const double* __restrict__ v,
                                                    double* __restrict__ part) {
  __shared__ double red[256];
  double a = 0.0;
  for (int64_t e = blockIdx.x * int64_t(256) + threadIdx.x; e < n; e += int64_t(gridDim.x) * 256)
    a = fma(v[e], v[e], a);
  red[threadIdx.x] = a;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void sqrt_sum_kernel(int nparts, const double* __restrict__ part, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < nparts; ++q) s += part[q];
    *out = sqrt(s);
  }
}

struct LqBufs {
  double *Qt, *R, *r, *r2, *z, *u, *x64;
  int *nonzero, *rows;
  void* qr;
  size_t qr_bytes;
};

size_t lq_layout(int64_t s, int64_t n, void* work, size_t bytes, LqBufs* out) {
  auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
  size_t off = 0;
  LqBufs b{};
  char* base = static_cast<char*>(work);
  auto take = [&](size_t nb) {
    char* p = base ? base + off : nullptr;
    off += up(nb);
    return p;
  };
  b.Qt = reinterpret_cast<double*>(take(sizeof(double) * s * n));
  b.R = reinterpret_cast<double*>(take(sizeof(double) * s * s));
  b.r = reinterpret_cast<double*>(take(sizeof(double) * s));
  b.r2 = reinterpret_cast<double*>(take(sizeof(double) * s));
  b.z = reinterpret_cast<double*>(take(sizeof(double) * s));
  b.rows = reinterpret_cast<int*>(take(sizeof(int) * s));
  b.u = reinterpret_cast<double*>(take(sizeof(double) * n));
  b.x64 = reinterpret_cast<double*>(take(sizeof(double) * n));
  b.nonzero = reinterpret_cast<int*>(take(sizeof(int)));
  b.qr_bytes = qr_work_bytes(n, s);
  b.qr = take(b.qr_bytes);
  if (out) *out = b;
  return off;
}

template <class T>
int lq_update(rla_ctx* ctx, int64_t s, int64_t n, const T* At, int64_t at_rs, int64_t at_cs,
              const T* bt, int64_t bt_inc, T* x, double alpha, void* work, size_t work_bytes) {
  RLA_CHECK_CTX(ctx);
  if (s < 1 || n < 1) return fail(RLA_EARG, "rla_kaczmarz_lq_update: empty block");
  if (s > n)
    return fail(RLA_EARG, "rla_kaczmarz_lq_update: a block of more rows than columns is not "
                          "supported");
  if (!At || !bt || !x || !work) return fail(RLA_EARG, "rla_kaczmarz_lq_update: NULL pointer");
  LqBufs b;
  if (lq_layout(s, n, work, work_bytes, &b) > work_bytes)
    return fail(RLA_EARG, "rla_kaczmarz_lq_update: workspace too small");
  int rc = use_device(ctx);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  const int64_t mx = std::max(s, n);
  lq_vec_in_kernel<T><<<dim3(unsigned(ceil_div(mx, 256))), 256, 0, st>>>(s, bt, bt_inc, b.r, n,
                                                                         x, b.x64, b.nonzero);
  lq_pack_kernel<T><<<dim3(unsigned(ceil_div(s * n, 256))), 256, 0, st>>>(
      s, n, At, at_rs, at_cs, nullptr, b.Qt, b.nonzero);
  RLA_LAUNCH_CHECK("kaczmarz_lq/pack");
  int nonzero = 0;
  RLA_HIP(hipMemcpyAsync(&nonzero, b.nonzero, sizeof(int), hipMemcpyDeviceToHost, st));
  RLA_HIP(hipStreamSynchronize(st));
  if (!nonzero) return RLA_OK;  // kaczmarz.jl:300: a zero block changes nothing
  // r = bt - At x: At is Qt' (s x n view: rows of stride n)
  rc = gemm_mfma<double>(ctx, s, 1, n, -1.0, b.Qt, n, 1, b.x64, 1, n, 1.0, b.r, 1, s);
  if (rc) return rc;
  rc = qr_f64(ctx, n, s, b.Qt, n, b.R, s, b.qr, b.qr_bytes, nullptr, nullptr);
  if (rc) return rc;
  // A row of At dependent on the others leaves a pivot of R at rounding level
  // (eps of the scale after Householder, sqrt(eps) after CholeskyQR).  QR
  // without pivoting then mixes an arbitrary direction into the later columns
  // of Q, so such a block is factored again without those rows: the
  // projection onto the independent rows, which for a consistent block is the
  // pseudoinverse step.
  std::vector<double> diag(s);
  RLA_HIP(hipMemcpy2DAsync(diag.data(), sizeof(double), b.R, sizeof(double) * (s + 1),
                           sizeof(double), s, hipMemcpyDeviceToHost, st));
  RLA_HIP(hipStreamSynchronize(st));
  double dmax = 0.0;
  for (double d : diag) dmax = std::max(dmax, std::fabs(d));
  const double floor = std::sqrt(double(s) * 2.220446049250313e-16) * dmax;
  std::vector<int> keep;
  for (int64_t j = 0; j < s; ++j)
    if (std::fabs(diag[j]) > floor) keep.push_back(static_cast<int>(j));
  int64_t sk = s;
  double* r = b.r;
  if (static_cast<int64_t>(keep.size()) < s) {
    sk = static_cast<int64_t>(keep.size());
    if (sk == 0) return RLA_OK;
    RLA_HIP(hipMemcpyAsync(b.rows, keep.data(), sizeof(int) * sk, hipMemcpyHostToDevice, st));
    lq_pack_kernel<T><<<dim3(unsigned(ceil_div(sk * n, 256))), 256, 0, st>>>(
        sk, n, At, at_rs, at_cs, b.rows, b.Qt, b.nonzero);
    lq_gather_kernel<<<dim3(unsigned(ceil_div(sk, 256))), 256, 0, st>>>(sk, b.rows, b.r, b.r2);
    RLA_LAUNCH_CHECK("kaczmarz_lq/repack");
    rc = qr_f64(ctx, n, sk, b.Qt, n, b.R, sk, b.qr, b.qr_bytes, nullptr, nullptr);
    if (rc) return rc;
    r = b.r2;
    // the host copy of keep must outlive the asynchronous upload
    RLA_HIP(hipStreamSynchronize(st));
  }
  lq_lower_solve_kernel<<<1, 256, 0, st>>>(sk, b.R, sk, r, b.z);
  RLA_LAUNCH_CHECK("kaczmarz_lq/solve");
  rc = gemm_mfma<double>(ctx, n, 1, sk, 1.0, b.Qt, 1, n, b.z, 1, sk, 0.0, b.u, 1, n);
  if (rc) return rc;
  lq_axpy_kernel<T><<<dim3(unsigned(ceil_div(n, 256))), 256, 0, st>>>(n, alpha, b.u, x);
  RLA_LAUNCH_CHECK("kaczmarz_lq/axpy");
  return RLA_OK;
}

}  // namespace

}  // namespace rla

using namespace rla;

extern "C" {

int rla_kaczmarz_lq_workspace_bytes(int64_t s, int64_t n, size_t* bytes) {
  if (!bytes || s < 1 || n < 1) return fail(RLA_EARG, "rla_kaczmarz_lq_workspace_bytes: bad arguments");
  *bytes = lq_layout(s, n, nullptr, 0, nullptr);
  return RLA_OK;
}

int rla_kaczmarz_lq_update_f32(rla_ctx* ctx, int64_t s, int64_t n, const float* At,
                               int64_t at_rs, int64_t at_cs, const float* bt, int64_t bt_inc,
                               float* x, double alpha, void* work, size_t work_bytes) {
  return lq_update<float>(ctx, s, n, At, at_rs, at_cs, bt, bt_inc, x, alpha, work, work_bytes);
}

int rla_kaczmarz_lq_update_f64(rla_ctx* ctx, int64_t s, int64_t n, const double* At,
                               int64_t at_rs, int64_t at_cs, const double* bt, int64_t bt_inc,
                               double* x, double alpha, void* work, size_t work_bytes) {
  return lq_update<double>(ctx, s, n, At, at_rs, at_cs, bt, bt_inc, x, alpha, work, work_bytes);
}

// ||b - A x|| in FP64 for Float64 data, A m x n at A[i*a_rs + j*a_cs]
int rla_residual_norm_f64(rla_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t a_rs,
                          int64_t a_cs, const double* x, const double* b, double* out_host) {
  RLA_CHECK_CTX(ctx);
  if (m < 1 || n < 0 || !out_host || !b || (n > 0 && (!A || !x)))
    return fail(RLA_EARG, "bad residual arguments");
  int rc = use_device(ctx);
  if (rc) return rc;
  const int nparts = static_cast<int>(std::min<int64_t>(1024, ceil_div(m, 256)));
  void* ws = nullptr;
  rc = workspace(ctx, sizeof(double) * (m + nparts + 1), &ws);
  if (rc) return rc;
  double* r = static_cast<double*>(ws);
  double* part = r + m;
  double* res = part + nparts;
  RLA_HIP(hipMemcpyAsync(r, b, sizeof(double) * m, hipMemcpyDeviceToDevice, ctx->stream));
  if (n > 0) {
    rc = gemm_mfma<double>(ctx, m, 1, n, -1.0, A, a_rs, a_cs, x, 1, n, 1.0, r, 1, m);
    if (rc) return rc;
  }
  sumsq_kernel<<<dim3(unsigned(nparts)), 256, 0, ctx->stream>>>(m, r, part);
  sqrt_sum_kernel<<<1, 64, 0, ctx->stream>>>(nparts, part, res);
  RLA_LAUNCH_CHECK("residual_f64");
  RLA_HIP(hipMemcpyAsync(out_host, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  return RLA_OK;
}

}  // extern "C"
