// Recipe generators: SparseSign / CountSketch indices and signs, SRHT sign
// bits and sampled rows, Gaussian operators.  Each generator has a device
// kernel and a host twin built from the same __host__ __device__ code, so
// integer draws match bit for bit.
#include <cmath>
#include <vector>

#include "philox.hpp"
#include "rla_internal.hpp"

namespace rla {

template <class T>
__host__ __device__ inline void sparse_sign_group(uint64_t seed, uint64_t epoch, int64_t j,
                                                  uint32_t s, int nnz, T sc, int32_t* idx,
                                                  T* val) {
  Stream st(seed, epoch, DOM_INDEX, static_cast<uint64_t>(j));
  floyd_sorted(st, s, nnz, idx + j * nnz);
  for (int q = 0; q < nnz; ++q) {
    uint64_t e = static_cast<uint64_t>(j) * nnz + q;
    uint32_t w = sign_word(seed, epoch, e >> 5);
    val[j * nnz + q] = ((w >> (e & 31)) & 1u) ? sc : -sc;
  }
}

template <class T>
__global__ void sparse_sign_gen_kernel(uint64_t seed, uint64_t epoch, int64_t n_groups,
                                       uint32_t s, int nnz, T sc, int32_t* idx, T* val) {
  int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (j >= n_groups) return;
  sparse_sign_group<T>(seed, epoch, j, s, nnz, sc, idx, val);
}

template <class T>
__host__ __device__ inline void count_sketch_group(uint64_t seed, uint64_t epoch, int64_t j,
                                                   uint32_t s, int32_t* idx, T* val) {
  Stream st(seed, epoch, DOM_INDEX, static_cast<uint64_t>(j));
  idx[j] = static_cast<int32_t>(st.below(s));
  uint32_t w = sign_word(seed, epoch, static_cast<uint64_t>(j) >> 5);
  val[j] = ((w >> (j & 31)) & 1u) ? T(1) : T(-1);
}

template <class T>
__global__ void count_sketch_gen_kernel(uint64_t seed, uint64_t epoch, int64_t n_groups,
                                        uint32_t s, int32_t* idx, T* val) {
  int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (j >= n_groups) return;
  count_sketch_group<T>(seed, epoch, j, s, idx, val);
}

// BitVector chunk c packs sign words 2c (low half) and 2c+1 (high half);
// bits at or beyond n_pad are zero, as in a Julia BitVector.
__host__ __device__ inline uint64_t srht_sign_chunk(uint64_t seed, uint64_t epoch, int64_t c,
                                                    int64_t n_pad) {
  uint64_t lo = sign_word(seed, epoch, static_cast<uint64_t>(2 * c));
  uint64_t hi = sign_word(seed, epoch, static_cast<uint64_t>(2 * c + 1));
  uint64_t ch = lo | (hi << 32);
  int64_t valid = n_pad - 64 * c;
  if (valid < 64) ch &= (uint64_t(1) << valid) - 1;
  return ch;
}

__global__ void srht_sign_kernel(uint64_t seed, uint64_t epoch, int64_t n_pad, int64_t nchunks,
                                 uint64_t* chunks) {
  int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (c >= nchunks) return;
  chunks[c] = srht_sign_chunk(seed, epoch, c, n_pad);
}

// Floyd's algorithm for an s-subset of [0, n_pad) (membership in a bitmap),
// then a Fisher–Yates shuffle so the order is uniform as well.
__host__ __device__ inline void srht_sample(uint64_t seed, uint64_t epoch, uint32_t n_pad,
                                            uint32_t s, int32_t* op, uint32_t* bitmap) {
  Stream st(seed, epoch, DOM_INDEX, 0);
  uint32_t k = 0;
  for (uint32_t t = n_pad - s; t < n_pad; ++t) {
    uint32_t v = st.below(t + 1u);
    if ((bitmap[v >> 5] >> (v & 31)) & 1u) v = t;
    bitmap[v >> 5] |= 1u << (v & 31);
    op[k++] = static_cast<int32_t>(v);
  }
  Stream sp(seed, epoch, DOM_PERM, 0);
  for (uint32_t i = s; i > 1; --i) {
    uint32_t j = sp.below(i);
    int32_t tmp = op[i - 1];
    op[i - 1] = op[j];
    op[j] = tmp;
  }
}

__global__ void srht_sample_kernel(uint64_t seed, uint64_t epoch, uint32_t n_pad, uint32_t s,
                                   int32_t* op, uint32_t* bitmap) {
  if (blockIdx.x == 0 && threadIdx.x == 0) srht_sample(seed, epoch, n_pad, s, op, bitmap);
}

__global__ void zero_u32_kernel(uint32_t* p, int64_t n) {
  int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Box–Muller pair q -> entries 2q, 2q+1 of the column-major rows x cols
// operator.
__host__ __device__ inline void gauss_pair(uint64_t seed, uint64_t epoch, uint64_t q, double& z0,
                                           double& z1) {
  Stream st(seed, epoch, DOM_GAUSS, q);
  double u1 = st.open01();
  double u2 = st.open01();
  double r = sqrt(-2.0 * log(u1));
  double a = 6.283185307179586 * u2;
  z0 = r * cos(a);
  z1 = r * sin(a);
}

template <class T>
__host__ __device__ inline void gauss_store(T* out, int64_t rows, int64_t ld, int64_t e, double z,
                                            T sc) {
  int64_t i = e % rows, j = e / rows;
  out[i + j * ld] = sc * static_cast<T>(z);
}

template <class T>
__global__ void gaussian_gen_kernel(uint64_t seed, uint64_t epoch, int64_t rows, int64_t cols,
                                    T sc, T* out, int64_t ld) {
  int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int64_t total = rows * cols;
  if (2 * q >= total) return;
  double z0, z1;
  gauss_pair(seed, epoch, static_cast<uint64_t>(q), z0, z1);
  gauss_store<T>(out, rows, ld, 2 * q, z0, sc);
  if (2 * q + 1 < total) gauss_store<T>(out, rows, ld, 2 * q + 1, z1, sc);
}

static int check_sparse_sign_args(int64_t n_groups, int64_t s, int64_t nnz) {
  if (s <= 0) return fail(RLA_EARG, "Field `compression_dim` must be positive.");
  if (nnz <= 0) return fail(RLA_EARG, "Field `nnz` must be positive.");
  if (nnz > s)
    return fail(RLA_EARG, "Number of non-zero indices, " + std::to_string(nnz) +
                              ", must be less than or equal to compression dimension, " +
                              std::to_string(s) + ".");
  if (n_groups < 0) return fail(RLA_EARG, "negative number of groups");
  if (s > INT32_MAX || n_groups * nnz > INT32_MAX)
    return fail(RLA_ENOTSUP, "sparse sign recipe larger than int32 indexing");
  return RLA_OK;
}

template <class T>
static int sparse_sign_gen(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_groups,
                           int64_t s, int64_t nnz, int32_t* idx, T* val) {
  RLA_CHECK_CTX(ctx);
  int rc = check_sparse_sign_args(n_groups, s, nnz);
  if (rc) return rc;
  if (n_groups == 0) return RLA_OK;
  rc = use_device(ctx);
  if (rc) return rc;
  T sc = static_cast<T>(1.0 / std::sqrt(static_cast<double>(nnz)));
  sparse_sign_gen_kernel<T><<<dim3(static_cast<unsigned>(ceil_div(n_groups, 256))), 256, 0,
                              ctx->stream>>>(seed, epoch, n_groups, static_cast<uint32_t>(s),
                                             static_cast<int>(nnz), sc, idx, val);
  RLA_LAUNCH_CHECK("sparse_sign_gen");
  return RLA_OK;
}

template <class T>
static int count_sketch_gen(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_groups,
                            int64_t s, int32_t* idx, T* val) {
  RLA_CHECK_CTX(ctx);
  if (s <= 0) return fail(RLA_EARG, "Field 'compression_dim' must be positive.");
  if (n_groups < 0) return fail(RLA_EARG, "negative number of groups");
  if (s > INT32_MAX || n_groups > INT32_MAX)
    return fail(RLA_ENOTSUP, "count sketch recipe larger than int32 indexing");
  if (n_groups == 0) return RLA_OK;
  int rc = use_device(ctx);
  if (rc) return rc;
  count_sketch_gen_kernel<T><<<dim3(static_cast<unsigned>(ceil_div(n_groups, 256))), 256, 0,
                               ctx->stream>>>(seed, epoch, n_groups, static_cast<uint32_t>(s),
                                              idx, val);
  RLA_LAUNCH_CHECK("count_sketch_gen");
  return RLA_OK;
}

static int check_srht_args(int64_t n_pad, int64_t s) {
  if (!is_pow2(n_pad)) return fail(RLA_EDIM, "Size of vector must be power of 2.");
  if (n_pad > (int64_t(1) << 30)) return fail(RLA_ENOTSUP, "n_pad above 2^30 is not supported");
  if (s <= 0) return fail(RLA_EARG, "Field `compression_dim` must be positive.");
  if (s > n_pad)
    return fail(RLA_EARG, "Cannot draw a sample of size " + std::to_string(s) +
                              " without replacement from 1:" + std::to_string(n_pad) + ".");
  return RLA_OK;
}

template <class T>
static int gaussian_gen(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t rows, int64_t cols,
                        double scale, T* out, int64_t ld) {
  RLA_CHECK_CTX(ctx);
  if (rows < 0 || cols < 0) return fail(RLA_EARG, "negative Gaussian operator size");
  if (ld < rows) return fail(RLA_EARG, "leading dimension smaller than the row count");
  if (rows == 0 || cols == 0) return RLA_OK;
  int rc = use_device(ctx);
  if (rc) return rc;
  int64_t pairs = ceil_div(rows * cols, 2);
  gaussian_gen_kernel<T><<<dim3(static_cast<unsigned>(ceil_div(pairs, 256))), 256, 0,
                           ctx->stream>>>(seed, epoch, rows, cols, static_cast<T>(scale), out,
                                          ld);
  RLA_LAUNCH_CHECK("gaussian_gen");
  return RLA_OK;
}

}  // namespace rla

using namespace rla;

extern "C" {

int rla_sparse_sign_gen_f32(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_groups,
                            int64_t s, int64_t nnz, int32_t* idx, float* val) {
  return sparse_sign_gen<float>(ctx, seed, epoch, n_groups, s, nnz, idx, val);
}

int rla_sparse_sign_gen_f64(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_groups,
                            int64_t s, int64_t nnz, int32_t* idx, double* val) {
  return sparse_sign_gen<double>(ctx, seed, epoch, n_groups, s, nnz, idx, val);
}

int rla_host_sparse_sign_gen_f64(uint64_t seed, uint64_t epoch, int64_t n_groups, int64_t s,
                                 int64_t nnz, int32_t* idx, double* val) {
  int rc = check_sparse_sign_args(n_groups, s, nnz);
  if (rc) return rc;
  double sc = 1.0 / std::sqrt(static_cast<double>(nnz));
  for (int64_t j = 0; j < n_groups; ++j)
    sparse_sign_group<double>(seed, epoch, j, static_cast<uint32_t>(s), static_cast<int>(nnz),
                              sc, idx, val);
  return RLA_OK;
}

int rla_count_sketch_gen_f32(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_groups,
                             int64_t s, int32_t* idx, float* val) {
  return count_sketch_gen<float>(ctx, seed, epoch, n_groups, s, idx, val);
}

int rla_count_sketch_gen_f64(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_groups,
                             int64_t s, int32_t* idx, double* val) {
  return count_sketch_gen<double>(ctx, seed, epoch, n_groups, s, idx, val);
}

int rla_host_count_sketch_gen_f64(uint64_t seed, uint64_t epoch, int64_t n_groups, int64_t s,
                                  int32_t* idx, double* val) {
  if (s <= 0) return fail(RLA_EARG, "Field 'compression_dim' must be positive.");
  for (int64_t j = 0; j < n_groups; ++j)
    count_sketch_group<double>(seed, epoch, j, static_cast<uint32_t>(s), idx, val);
  return RLA_OK;
}

int rla_srht_gen(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_pad, int64_t s,
                 int32_t* op, uint64_t* sign_chunks) {
  RLA_CHECK_CTX(ctx);
  int rc = check_srht_args(n_pad, op ? s : 1);
  if (rc) return rc;
  rc = use_device(ctx);
  if (rc) return rc;
  int64_t nchunks = ceil_div(n_pad, 64);
  if (sign_chunks) {
    srht_sign_kernel<<<dim3(static_cast<unsigned>(ceil_div(nchunks, 256))), 256, 0,
                       ctx->stream>>>(seed, epoch, n_pad, nchunks, sign_chunks);
    RLA_LAUNCH_CHECK("srht_sign");
  }
  if (op) {
    int64_t words = ceil_div(n_pad, 32);
    void* ws = nullptr;
    rc = workspace(ctx, static_cast<size_t>(words) * 4, &ws);
    if (rc) return rc;
    uint32_t* bitmap = static_cast<uint32_t*>(ws);
    zero_u32_kernel<<<dim3(static_cast<unsigned>(ceil_div(words, 256))), 256, 0, ctx->stream>>>(
        bitmap, words);
    srht_sample_kernel<<<1, 1, 0, ctx->stream>>>(seed, epoch, static_cast<uint32_t>(n_pad),
                                                 static_cast<uint32_t>(s), op, bitmap);
    RLA_LAUNCH_CHECK("srht_sample");
  }
  return RLA_OK;
}

int rla_host_srht_gen(uint64_t seed, uint64_t epoch, int64_t n_pad, int64_t s, int32_t* op,
                      uint64_t* sign_chunks) {
  int rc = check_srht_args(n_pad, op ? s : 1);
  if (rc) return rc;
  int64_t nchunks = ceil_div(n_pad, 64);
  if (sign_chunks)
    for (int64_t c = 0; c < nchunks; ++c) sign_chunks[c] = srht_sign_chunk(seed, epoch, c, n_pad);
  if (op) {
    std::vector<uint32_t> bitmap(static_cast<size_t>(ceil_div(n_pad, 32)), 0u);
    srht_sample(seed, epoch, static_cast<uint32_t>(n_pad), static_cast<uint32_t>(s), op,
                bitmap.data());
  }
  return RLA_OK;
}

int rla_gaussian_gen_f32(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t rows, int64_t cols,
                         double scale, float* out, int64_t ld) {
  return gaussian_gen<float>(ctx, seed, epoch, rows, cols, scale, out, ld);
}

int rla_gaussian_gen_f64(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t rows, int64_t cols,
                         double scale, double* out, int64_t ld) {
  return gaussian_gen<double>(ctx, seed, epoch, rows, cols, scale, out, ld);
}

int rla_host_gaussian_gen_f64(uint64_t seed, uint64_t epoch, int64_t rows, int64_t cols,
                              double scale, double* out, int64_t ld) {
  if (rows < 0 || cols < 0 || ld < rows) return fail(RLA_EARG, "bad Gaussian operator shape");
  int64_t total = rows * cols;
  for (int64_t q = 0; 2 * q < total; ++q) {
    double z0, z1;
    gauss_pair(seed, epoch, static_cast<uint64_t>(q), z0, z1);
    gauss_store<double>(out, rows, ld, 2 * q, z0, scale);
    if (2 * q + 1 < total) gauss_store<double>(out, rows, ld, 2 * q + 1, z1, scale);
  }
  return RLA_OK;
}

}  // extern "C"
