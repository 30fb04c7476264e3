// FJLT sparse factor K = sprandn(type, s, n_pad, q) (src/Compressors/fjlt.jl:183,
// SparseArrays.sprandn): every entry independently nonzero with probability
// q, nonzero values N(0, 1), stored column-major (CSC, rows ascending).
//
// Entry (i, j) is nonzero iff word i of the counter-based stream
// (seed, epoch, DOM_FJLT_ROWS, j) is below thr = floor(q * 2^32) (all of
// them for q = 1): an integer test, so the host and device generators give
// the same index sets bit for bit, and P(nonzero) = thr / 2^32 is q to within
// 2^-32.  Word i of a stream is word i%4 of Philox block i/4, so every entry
// is drawn independently of the others.  The r-th nonzero of column j takes
// the Box–Muller pair r/2 of (seed, epoch, DOM_FJLT_VALS, j) (one Philox block
// per pair: two 53-bit uniforms), first member for even r, second for odd.
//
// Device: one wavefront per column; lanes test rows lane, lane + 64, ...,
// ballots give each nonzero its rank in the column.  A count pass, a scan
// into colptr, then a fill pass over the same streams.
#include <cmath>

#include "philox.hpp"
#include "rla_internal.hpp"

namespace rla {

constexpr uint32_t DOM_FJLT_ROWS = 6u;
constexpr uint32_t DOM_FJLT_VALS = 7u;

// Philox block b of stream (seed, epoch, domain, item), as Stream lays it out
__host__ __device__ inline u32x4 fjlt_block(uint64_t seed, uint64_t epoch, uint32_t domain,
                                            uint64_t item, uint32_t b) {
  u32x4 c;
  c.v[0] = b;
  c.v[1] = static_cast<uint32_t>(item);
  c.v[2] = static_cast<uint32_t>(item >> 32) ^ (domain << 24);
  c.v[3] = static_cast<uint32_t>(epoch) ^ static_cast<uint32_t>(epoch >> 32);
  return philox4x32_10(c, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
}

__host__ __device__ inline bool fjlt_nonzero(uint64_t seed, uint64_t epoch, int64_t j, int64_t i,
                                             uint64_t thr) {
  const u32x4 w = fjlt_block(seed, epoch, DOM_FJLT_ROWS, static_cast<uint64_t>(j),
                             static_cast<uint32_t>(i >> 2));
  return static_cast<uint64_t>(w.v[i & 3]) < thr;
}

// N(0, 1) value of the r-th nonzero of column j
__host__ __device__ inline double fjlt_value(uint64_t seed, uint64_t epoch, int64_t j, int64_t r) {
  const u32x4 w = fjlt_block(seed, epoch, DOM_FJLT_VALS, static_cast<uint64_t>(j),
                             static_cast<uint32_t>(r >> 1));
  // two uniforms in (0, 1) with 53 random bits each (as Stream::open01)
  const double u1 = ((static_cast<double>(w.v[0] >> 5) * 67108864.0 +
                      static_cast<double>(w.v[1] >> 6)) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((static_cast<double>(w.v[2] >> 5) * 67108864.0 +
                      static_cast<double>(w.v[3] >> 6)) + 0.5) * (1.0 / 9007199254740992.0);
  const double rad = sqrt(-2.0 * log(u1));
  const double a = 6.283185307179586 * u2;
  return (r & 1) ? rad * sin(a) : rad * cos(a);
}

static uint64_t fjlt_threshold(double q) {
  if (q >= 1.0) return uint64_t(1) << 32;
  return static_cast<uint64_t>(std::floor(q * 4294967296.0));
}

constexpr int kFjltThreads = 256;  // four columns per workgroup

// colcnt[j] = nonzeros of column j
__global__ __launch_bounds__(kFjltThreads) void fjlt_count_kernel(uint64_t seed, uint64_t epoch,
                                                                  int64_t s, int64_t n_pad,
                                                                  uint64_t thr,
                                                                  int32_t* __restrict__ colcnt) {
  const int lane = threadIdx.x & 63;
  const int64_t j = int64_t(blockIdx.x) * (kFjltThreads / 64) + (threadIdx.x >> 6);
  if (j >= n_pad) return;
  int cnt = 0;
  for (int64_t i0 = 0; i0 < s; i0 += 64) {
    const int64_t i = i0 + lane;
    const bool nz = i < s && fjlt_nonzero(seed, epoch, j, i, thr);
    cnt += __popcll(__ballot(nz));
  }
  if (lane == 0) colcnt[j] = cnt;
}

template <class T>
__global__ __launch_bounds__(kFjltThreads) void fjlt_fill_kernel(
    uint64_t seed, uint64_t epoch, int64_t s, int64_t n_pad, uint64_t thr,
    const int32_t* __restrict__ colptr, int32_t* __restrict__ rowidx, T* __restrict__ val) {
  const int lane = threadIdx.x & 63;
  const int64_t j = int64_t(blockIdx.x) * (kFjltThreads / 64) + (threadIdx.x >> 6);
  if (j >= n_pad) return;
  const uint64_t below = (uint64_t(1) << lane) - 1u;
  int64_t r0 = 0;  // nonzeros of this column before the current 64 rows
  const int64_t base = colptr[j];
  for (int64_t i0 = 0; i0 < s; i0 += 64) {
    const int64_t i = i0 + lane;
    const bool nz = i < s && fjlt_nonzero(seed, epoch, j, i, thr);
    const uint64_t m = __ballot(nz);
    if (nz) {
      const int64_t r = r0 + __popcll(m & below);
      rowidx[base + r] = static_cast<int32_t>(i);
      val[base + r] = static_cast<T>(fjlt_value(seed, epoch, j, r));
    }
    r0 += __popcll(m);
  }
}

__global__ void fjlt_colptr0_kernel(int32_t* colptr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) colptr[0] = 0;
}

static int check_fjlt_args(int64_t s, int64_t n_pad, double q) {
  if (s < 1) return fail(RLA_EARG, "Field `compression_dim` must be positive.");
  if (n_pad < 1) return fail(RLA_EARG, "padded dimension must be positive");
  if (!(q >= 0.0 && q <= 1.0)) return fail(RLA_EARG, "Field `sparsity` must be between 0 and 1.");
  // expected nonzeros well inside 32-bit indices (a 2^31 cap with ~10 sigma room)
  if (q * double(s) * double(n_pad) > 1.5e9 || n_pad > INT32_MAX || s > INT32_MAX)
    return fail(RLA_ENOTSUP, "FJLT sparse factor too large for 32-bit indices");
  return RLA_OK;
}

template <class T>
static int fjlt_fill(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t s, int64_t n_pad,
                     double q, const int32_t* colptr, int32_t* rowidx, T* val) {
  RLA_CHECK_CTX(ctx);
  int rc = check_fjlt_args(s, n_pad, q);
  if (rc) return rc;
  if (!colptr || !rowidx || !val) return fail(RLA_EARG, "FJLT fill: NULL pointer");
  rc = use_device(ctx);
  if (rc) return rc;
  fjlt_fill_kernel<T><<<dim3(unsigned(ceil_div(n_pad, kFjltThreads / 64))), kFjltThreads, 0,
                        ctx->stream>>>(seed, epoch, s, n_pad, fjlt_threshold(q), colptr, rowidx,
                                       val);
  RLA_LAUNCH_CHECK("fjlt_fill");
  return RLA_OK;
}

}  // namespace rla

using namespace rla;

extern "C" {

int rla_fjlt_gen_count(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t s, int64_t n_pad,
                       double q, int32_t* colptr, int64_t* nnz) {
  RLA_CHECK_CTX(ctx);
  int rc = check_fjlt_args(s, n_pad, q);
  if (rc) return rc;
  if (!colptr || !nnz) return fail(RLA_EARG, "FJLT count: NULL pointer");
  rc = use_device(ctx);
  if (rc) return rc;
  fjlt_colptr0_kernel<<<1, 1, 0, ctx->stream>>>(colptr);
  RLA_LAUNCH_CHECK("fjlt_colptr0");
  fjlt_count_kernel<<<dim3(unsigned(ceil_div(n_pad, kFjltThreads / 64))), kFjltThreads, 0,
                      ctx->stream>>>(seed, epoch, s, n_pad, fjlt_threshold(q), colptr + 1);
  RLA_LAUNCH_CHECK("fjlt_count");
  rc = scan_inclusive_i32(ctx, colptr + 1, n_pad);
  if (rc) return rc;
  int32_t total = 0;
  RLA_HIP(hipMemcpyAsync(&total, colptr + n_pad, sizeof(int32_t), hipMemcpyDeviceToHost,
                         ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  *nnz = total;
  return RLA_OK;
}

int rla_fjlt_gen_fill_f32(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t s, int64_t n_pad,
                          double q, const int32_t* colptr, int32_t* rowidx, float* val) {
  return fjlt_fill<float>(ctx, seed, epoch, s, n_pad, q, colptr, rowidx, val);
}

int rla_fjlt_gen_fill_f64(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t s, int64_t n_pad,
                          double q, const int32_t* colptr, int32_t* rowidx, double* val) {
  return fjlt_fill<double>(ctx, seed, epoch, s, n_pad, q, colptr, rowidx, val);
}

int rla_host_fjlt_gen_f64(uint64_t seed, uint64_t epoch, int64_t s, int64_t n_pad, double q,
                          int32_t* colptr, int32_t* rowidx, double* val, int64_t cap,
                          int64_t* nnz) {
  int rc = check_fjlt_args(s, n_pad, q);
  if (rc) return rc;
  if (!colptr || !nnz) return fail(RLA_EARG, "FJLT host gen: NULL pointer");
  const uint64_t thr = fjlt_threshold(q);
  int64_t k = 0;
  colptr[0] = 0;
  for (int64_t j = 0; j < n_pad; ++j) {
    int64_t r = 0;
    for (int64_t i = 0; i < s; ++i) {
      if (!fjlt_nonzero(seed, epoch, j, i, thr)) continue;
      if (k < cap && rowidx && val) {
        rowidx[k] = static_cast<int32_t>(i);
        val[k] = fjlt_value(seed, epoch, j, r);
      }
      ++k;
      ++r;
    }
    colptr[j + 1] = static_cast<int32_t>(k);
  }
  *nnz = k;
  return k > cap && rowidx ? fail(RLA_EARG, "FJLT host gen: capacity too small") : RLA_OK;
}

}  // extern "C"
