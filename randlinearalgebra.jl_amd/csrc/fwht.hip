// Walsh–Hadamard transforms and the SRHT operator (generic path).
//
// The transform of a length n = 2^L column is done as the reference's
// radix-2 network (src/Compressors/helpers/fwht.jl:21-118): stage b pairs
// entries whose index differs in bit b, for b = 0, 1, ..., L-1, and writes
// (z + y, z - y).  Stage 0 first multiplies each entry by +-scaling; the
// reference evaluates that product and the stage-0 sum in Float64 (scaling is
// a Float64), so we do too, then round to the element type.  Every later
// stage is a plain add/subtract in the element type.  Doing the same stages in
// the same order makes the result bit-identical to the reference loop,
// whatever grouping of stages into passes is used.
//
// Generic path: a pass handles a run of nb consecutive bits [lo, lo + nb) for
// every column.  A workgroup loads gpw butterfly groups (2^nb entries each,
// stride 2^lo) into LDS with coalesced reads, runs the nb stages with a
// barrier between stages, and writes them back.  The first pass may read
// straight from the (unpadded, strided) input and write to the padded
// workspace.
#pragma clang fp contract(off)

#include <algorithm>

#include "rla_internal.hpp"

namespace rla {

constexpr int kPassThreads = 512;
constexpr size_t kPassLds = 64 * 1024;

__device__ inline bool sign_bit(const uint64_t* __restrict__ s, int64_t i) {
  return (s[i >> 6] >> (i & 63)) & 1ull;
}

template <class T>
__global__ __launch_bounds__(kPassThreads) void fwht_pass_kernel(
    const T* __restrict__ src, int64_t src_rs, int64_t src_cs, int64_t m_valid, T* P,
    int64_t ldp, int64_t ncols, int L, int lo, int nb, int lg_gpw,
    const uint64_t* __restrict__ signs, double scaling, int scale_stage0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);
  const int gpw = 1 << lg_gpw;
  const int64_t groups_per_col = int64_t(1) << (L - nb);
  const int64_t total_groups = groups_per_col * ncols;
  const int64_t G0 = int64_t(blockIdx.x) * gpw;
  if (G0 >= total_groups) return;
  const int ne = gpw << nb;
  const int tmask = (1 << nb) - 1;
  // tile index of (t, gl): contiguous groups when lo == 0, interleaved
  // groups (consecutive lanes -> consecutive groups) when lo > 0.
  const int tstride = lo == 0 ? 1 : gpw;
  const int gstride = lo == 0 ? (1 << nb) : 1;

  auto addr_of = [&](int t, int gl, int64_t& col, int64_t& pos) {
    const int64_t G = G0 + gl;
    col = G >> (L - nb);
    const int64_t g = G & (groups_per_col - 1);
    if (lo == 0) {
      pos = (g << nb) + t;
    } else {
      const int64_t glo = g & ((int64_t(1) << lo) - 1), ghi = g >> lo;
      pos = (ghi << (lo + nb)) + (int64_t(t) << lo) + glo;
    }
  };

  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    const int t = lo == 0 ? (e & tmask) : (e >> lg_gpw);
    const int gl = lo == 0 ? (e >> nb) : (e & (gpw - 1));
    int64_t col, pos;
    addr_of(t, gl, col, pos);
    T v;
    if (col >= ncols) {
      v = T(0);
    } else if (src) {
      v = pos < m_valid ? src[pos * src_rs + col * src_cs] : T(0);
    } else {
      v = P[col * ldp + pos];
    }
    tile[t * tstride + gl * gstride] = v;
  }

  const int npairs = ne >> 1;
  for (int b = 0; b < nb; ++b) {
    __syncthreads();
    const int hmask = (1 << b) - 1;
    for (int q = threadIdx.x; q < npairs; q += blockDim.x) {
      const int qt = lo == 0 ? (q & (tmask >> 1)) : (q >> lg_gpw);
      const int gl = lo == 0 ? (q >> (nb - 1)) : (q & (gpw - 1));
      const int t0 = ((qt >> b) << (b + 1)) | (qt & hmask);
      const int t1 = t0 | (1 << b);
      T* a0 = &tile[t0 * tstride + gl * gstride];
      T* a1 = &tile[t1 * tstride + gl * gstride];
      if (b == 0 && scale_stage0) {
        // global positions of the pair (lo == 0 here)
        const int64_t G = G0 + gl;
        const int64_t g = G & (groups_per_col - 1);
        const int64_t p0 = (g << nb) + t0;
        const double s0 = (!signs || sign_bit(signs, p0)) ? scaling : -scaling;
        const double s1 = (!signs || sign_bit(signs, p0 + 1)) ? scaling : -scaling;
        const double z = double(*a0) * s0;
        const double y = double(*a1) * s1;
        *a0 = static_cast<T>(z + y);
        *a1 = static_cast<T>(z - y);
      } else {
        const T z = *a0, y = *a1;
        *a0 = z + y;
        *a1 = z - y;
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    const int t = lo == 0 ? (e & tmask) : (e >> lg_gpw);
    const int gl = lo == 0 ? (e >> nb) : (e & (gpw - 1));
    int64_t col, pos;
    addr_of(t, gl, col, pos);
    if (col < ncols) P[col * ldp + pos] = tile[t * tstride + gl * gstride];
  }
}

// Stages 14 .. L-1 of Float32 columns of length 2^L = NCH * 2^14, after the
// chunk kernel: thread (e, column) holds entries e + c * 2^14 (c < NCH; four
// consecutive e per thread when VEC) and runs the remaining butterflies in
// registers, in the reference's order (chunk bit 0 first).
template <int NCH, bool VEC>
__global__ __launch_bounds__(256) void fwht_cross_kernel(float* P, int64_t ldp) {
  constexpr int W = VEC ? 4 : 1;
  const int64_t e = (int64_t(blockIdx.x) * 256 + threadIdx.x) * W;
  float* pc = P + int64_t(blockIdx.y) * ldp + e;
  float a[NCH][W];
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    if constexpr (VEC) {
      const float4 q = *reinterpret_cast<const float4*>(pc + (int64_t(c) << 14));
      a[c][0] = q.x;
      a[c][1] = q.y;
      a[c][2] = q.z;
      a[c][3] = q.w;
    } else {
      a[c][0] = pc[int64_t(c) << 14];
    }
  }
#pragma unroll
  for (int h = 1; h < NCH; h <<= 1)
#pragma unroll
    for (int c = 0; c < NCH; ++c)
      if (!(c & h))
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const float z = a[c][w], y = a[c | h][w];
          a[c][w] = z + y;
          a[c | h][w] = z - y;
        }
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    if constexpr (VEC) {
      *reinterpret_cast<float4*>(pc + (int64_t(c) << 14)) =
          make_float4(a[c][0], a[c][1], a[c][2], a[c][3]);
    } else {
      pc[int64_t(c) << 14] = a[c][0];
    }
  }
}

template <int NCH>
static int launch_fwht_cross(rla_ctx* ctx, float* P, int64_t ldp, int64_t ncols) {
  const bool vec = (reinterpret_cast<uintptr_t>(P) % 16 == 0) && (ldp % 4 == 0);
  const unsigned bx = vec ? (1u << 14) / (256 * 4) : (1u << 14) / 256;
  for (int64_t c0 = 0; c0 < ncols; c0 += 65535) {
    const unsigned by = static_cast<unsigned>(std::min<int64_t>(65535, ncols - c0));
    if (vec)
      fwht_cross_kernel<NCH, true><<<dim3(bx, by), 256, 0, ctx->stream>>>(P + c0 * ldp, ldp);
    else
      fwht_cross_kernel<NCH, false><<<dim3(bx, by), 256, 0, ctx->stream>>>(P + c0 * ldp, ldp);
    RLA_LAUNCH_CHECK("fwht_cross");
  }
  return RLA_OK;
}

// Y[r, k] = beta*Y[r, k] + alpha*P[op[r], k] in double (the reference's
// broadcast `beta .* Cv .+ alpha .* padding[op, :]`, srht.jl:259); beta == 0
// does not read Y.
template <class T>
__global__ void srht_gather_kernel(const T* __restrict__ P, int64_t ldp, int64_t s,
                                   int64_t ncols, const int32_t* __restrict__ op,
                                   T* __restrict__ Y, int64_t y_rs, int64_t y_cs, double alpha,
                                   double beta) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= s * ncols) return;
  const int64_t r = e % s, k = e / s;
  T* yp = Y + r * y_rs + k * y_cs;
  const double p = double(P[k * ldp + op[r]]);
  *yp = static_cast<T>(beta == 0.0 ? alpha * p : beta * double(*yp) + alpha * p);
}

// P[op[r], k] = alpha * X[r, k] (srht.jl:311 `pv' .= alpha .* Av`, rounded to
// the element type); P is zero elsewhere.
template <class T>
__global__ void srht_scatter_kernel(const T* __restrict__ X, int64_t x_rs, int64_t x_cs,
                                    int64_t s, int64_t ncols, const int32_t* __restrict__ op,
                                    T* __restrict__ P, int64_t ldp, double alpha) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= s * ncols) return;
  const int64_t r = e % s, k = e / s;
  P[k * ldp + op[r]] = static_cast<T>(alpha * double(X[r * x_rs + k * x_cs]));
}

// Y[i, k] = (+-1)*P[i, k] + beta*Y[i, k] for i < m (srht.jl:321-327: the
// sign flip after the transform, then axpby!(1, pv', beta, Cv)).
template <class T>
__global__ void srht_adjoint_epilogue_kernel(const T* __restrict__ P, int64_t ldp, int64_t m,
                                             int64_t ncols, const uint64_t* __restrict__ signs,
                                             T* __restrict__ Y, int64_t y_rs, int64_t y_cs,
                                             double beta) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= m * ncols) return;
  const int64_t i = e % m, k = e / m;
  T p = P[k * ldp + i];
  if (!sign_bit(signs, i)) p = -p;
  T* yp = Y + i * y_rs + k * y_cs;
  *yp = static_cast<T>(beta == 0.0 ? double(p) : double(p) + beta * double(*yp));
}

template <class T>
__global__ void zero_kernel(T* p, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) p[i] = T(0);
}

// Runs every pass of the transform of ncols columns of length 2^L held in P
// (or, for the first pass, read from src with m_valid rows).
template <class T>
static int fwht_passes(rla_ctx* ctx, const T* src, int64_t src_rs, int64_t src_cs,
                       int64_t m_valid, T* P, int64_t ldp, int64_t ncols, int L,
                       const uint64_t* signs, double scaling) {
  const int cap_lg = ilog2(static_cast<int64_t>(kPassLds / sizeof(T)));  // 14 f32, 13 f64
  int lo = 0;
  if constexpr (sizeof(T) == 4) {
    // Float32 columns of 2^14 or more with unit row stride: stages 0..13 in
    // the wave-local chunk kernel, the rest in the passes below
    if (L >= 14 && (!src || src_rs == 1)) {
      int rc = fwht_chunks_f32(ctx, src, src_cs, src ? m_valid : (int64_t(1) << L), P, ldp,
                               ncols, L, signs, scaling);
      if (rc) return rc;
      lo = 14;
      // up to 16 chunks: the remaining stages in one register pass
      switch (L - 14) {
        case 0: return RLA_OK;
        case 1: return launch_fwht_cross<2>(ctx, reinterpret_cast<float*>(P), ldp, ncols);
        case 2: return launch_fwht_cross<4>(ctx, reinterpret_cast<float*>(P), ldp, ncols);
        case 3: return launch_fwht_cross<8>(ctx, reinterpret_cast<float*>(P), ldp, ncols);
        case 4: return launch_fwht_cross<16>(ctx, reinterpret_cast<float*>(P), ldp, ncols);
        default: break;
      }
    }
  }
  while (lo < L) {
    int nb, lg_gpw;
    if (lo == 0) {
      nb = std::min(L, cap_lg);
      lg_gpw = cap_lg - nb;  // several short columns per workgroup
      const int64_t total_groups = (int64_t(1) << (L - nb)) * ncols;
      while (lg_gpw > 0 && (int64_t(1) << lg_gpw) > total_groups) --lg_gpw;
    } else {
      nb = std::min(L - lo, cap_lg - 6);
      lg_gpw = std::min(lo, cap_lg - nb);
    }
    const int64_t total_groups = (int64_t(1) << (L - nb)) * ncols;
    const int64_t blocks = ceil_div(total_groups, int64_t(1) << lg_gpw);
    const size_t lds = (size_t(1) << (nb + lg_gpw)) * sizeof(T);
    const bool first = lo == 0;
    fwht_pass_kernel<T><<<dim3(static_cast<unsigned>(blocks)), kPassThreads, lds,
                          ctx->stream>>>(first ? src : nullptr, src_rs, src_cs, m_valid, P, ldp,
                                         ncols, L, lo, nb, lg_gpw, signs, scaling,
                                         first ? 1 : 0);
    RLA_LAUNCH_CHECK("fwht_pass");
    lo += nb;
  }
  return RLA_OK;
}

template <class T>
static int fwht(rla_ctx* ctx, int64_t n, int64_t ncols, const uint64_t* signs, double scaling,
                T* X, int64_t ldx) {
  RLA_CHECK_CTX(ctx);
  if (!is_pow2(n) || n < 2) return fail(RLA_EDIM, "Size of vector must be power of 2.");
  if (ncols < 0 || ldx < n) return fail(RLA_EARG, "bad column count or leading dimension");
  if (ncols == 0) return RLA_OK;
  int rc = use_device(ctx);
  if (rc) return rc;
  return fwht_passes<T>(ctx, nullptr, 0, 0, n, X, ldx, ncols, ilog2(n), signs, scaling);
}

// Columns per chunk so that the padded workspace stays below ~512 MiB.
template <class T>
static int64_t srht_chunk_cols(int64_t n_pad, int64_t ncols) {
  const int64_t budget = int64_t(512) << 20;
  int64_t c = budget / (n_pad * int64_t(sizeof(T)));
  if (c < 1) c = 1;
  return std::min(c, ncols);
}

template <class T>
int srht_fused_forward(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                       const int32_t* op, const uint64_t* signs, double scale, const void* plan,
                       const T* X, int64_t x_rs, int64_t x_cs, T* Y, int64_t y_rs, int64_t y_cs,
                       double alpha, double beta, bool exact, bool* handled);

template <class T>
static int srht_apply(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                      int64_t ncols, const int32_t* op, const uint64_t* signs, double scale,
                      const void* plan, const T* X, int64_t x_rs, int64_t x_cs, T* Y,
                      int64_t y_rs, int64_t y_cs, double alpha, double beta, bool exact) {
  RLA_CHECK_CTX(ctx);
  if (!is_pow2(n_pad) || n_pad < 2) return fail(RLA_EDIM, "Size of vector must be power of 2.");
  if (m < 1 || m > n_pad || s < 1 || s > n_pad || ncols < 0)
    return fail(RLA_EARG, "SRHT sizes must satisfy 1 <= m <= n_pad and 1 <= s <= n_pad");
  if (ncols == 0) return RLA_OK;
  if (!op || !signs || !X || !Y) return fail(RLA_EARG, "SRHT apply: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  const int L = ilog2(n_pad);
  if (!adjoint) {
    bool handled = false;
    rc = srht_fused_forward<T>(ctx, m, n_pad, s, ncols, op, signs, scale, plan, X, x_rs, x_cs, Y,
                               y_rs, y_cs, alpha, beta, exact, &handled);
    if (rc || handled) return rc;
  }
  const int64_t cc = srht_chunk_cols<T>(n_pad, ncols);
  void* ws = nullptr;
  rc = workspace(ctx, static_cast<size_t>(n_pad * cc) * sizeof(T), &ws);
  if (rc) return rc;
  T* P = static_cast<T*>(ws);
  for (int64_t k0 = 0; k0 < ncols; k0 += cc) {
    const int64_t kc = std::min(cc, ncols - k0);
    const T* Xk = X + k0 * x_cs;
    T* Yk = Y + k0 * y_cs;
    if (!adjoint) {
      rc = fwht_passes<T>(ctx, Xk, x_rs, x_cs, m, P, n_pad, kc, L, signs, scale);
      if (rc) return rc;
      srht_gather_kernel<T><<<dim3(static_cast<unsigned>(ceil_div(s * kc, 256))), 256, 0,
                              ctx->stream>>>(P, n_pad, s, kc, op, Yk, y_rs, y_cs, alpha, beta);
      RLA_LAUNCH_CHECK("srht_gather");
    } else {
      zero_kernel<T><<<dim3(static_cast<unsigned>(ceil_div(n_pad * kc, 256))), 256, 0,
                       ctx->stream>>>(P, n_pad * kc);
      srht_scatter_kernel<T><<<dim3(static_cast<unsigned>(ceil_div(s * kc, 256))), 256, 0,
                               ctx->stream>>>(Xk, x_rs, x_cs, s, kc, op, P, n_pad, alpha);
      rc = fwht_passes<T>(ctx, nullptr, 0, 0, n_pad, P, n_pad, kc, L, nullptr, scale);
      if (rc) return rc;
      srht_adjoint_epilogue_kernel<T><<<dim3(static_cast<unsigned>(ceil_div(m * kc, 256))), 256,
                                        0, ctx->stream>>>(P, n_pad, m, kc, signs, Yk, y_rs, y_cs,
                                                          beta);
      RLA_LAUNCH_CHECK("srht_adjoint");
    }
  }
  return RLA_OK;
}

template <class T>
static int sparse_apply_t(rla_ctx* ctx, int trans_m, int64_t nrows, int64_t ncols,
                          const int32_t* colptr, const int32_t* rowidx, const T* val,
                          const int32_t* rowptr, const int32_t* colidx, const T* val_t,
                          int64_t n_rhs, const T* X, int64_t x_rs, int64_t x_cs, T* C,
                          int64_t c_rs, int64_t c_cs, double alpha, double beta) {
  if constexpr (sizeof(T) == 4)
    return rla_sparse_apply_f32(ctx, trans_m, nrows, ncols, colptr, rowidx, val, rowptr, colidx,
                                val_t, n_rhs, X, x_rs, x_cs, C, c_rs, c_cs, alpha, beta);
  else
    return rla_sparse_apply_f64(ctx, trans_m, nrows, ncols, colptr, rowidx, val, rowptr, colidx,
                                val_t, n_rhs, X, x_rs, x_cs, C, c_rs, c_cs, alpha, beta);
}

// FJLT (fjlt.jl:257-535) with T = scale * K * H * D * pad, an s x m operator
// (K: s x n_pad, CSC (colptr, rowidx, val) and CSR (rowptr, colidx, val_t)):
//   forward (adjoint = 0), Y = alpha * T * X + beta * Y: the FWHT passes (signs
//     and scale on the first stage, as fwht!(x, signs; scaling)) into the
//     padded workspace, then the CSR row-gather product with K
//     (mul!(Cv, S.op, S.padding, alpha, beta), fjlt.jl:284-296);
//   adjoint (adjoint = 1), Y = alpha * T' * X + beta * Y: W = alpha * K' * X
//     (mul!(pv', Av, S.op, alpha, 0), fjlt.jl:341), the FWHT passes with the
//     scale only, then the sign flip and Y = (D W)[0:m] + beta * Y
//     (fjlt.jl:348-357).
// Columns are processed in chunks that bound the padded workspace, in place
// of the reference's block_size column blocks.
template <class T>
static int fjlt_apply(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                      int64_t ncols, const int32_t* colptr, const int32_t* rowidx, const T* val,
                      const int32_t* rowptr, const int32_t* colidx, const T* val_t,
                      const uint64_t* signs, double scale, const T* X, int64_t x_rs,
                      int64_t x_cs, T* Y, int64_t y_rs, int64_t y_cs, double alpha,
                      double beta) {
  RLA_CHECK_CTX(ctx);
  if (!is_pow2(n_pad) || n_pad < 2) return fail(RLA_EDIM, "Size of vector must be power of 2.");
  if (m < 1 || m > n_pad || s < 1 || ncols < 0)
    return fail(RLA_EARG, "FJLT sizes must satisfy 1 <= m <= n_pad and s >= 1");
  if (ncols == 0) return RLA_OK;
  if (!colptr || !rowidx || !val || !rowptr || !colidx || !val_t || !signs || !X || !Y)
    return fail(RLA_EARG, "FJLT apply: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  const int L = ilog2(n_pad);
  const int64_t cc = srht_chunk_cols<T>(n_pad, ncols);
  void* ws = nullptr;
  rc = workspace(ctx, static_cast<size_t>(n_pad * cc) * sizeof(T), &ws);
  if (rc) return rc;
  T* P = static_cast<T*>(ws);
  for (int64_t k0 = 0; k0 < ncols; k0 += cc) {
    const int64_t kc = std::min(cc, ncols - k0);
    const T* Xk = X + k0 * x_cs;
    T* Yk = Y + k0 * y_cs;
    if (!adjoint) {
      rc = fwht_passes<T>(ctx, Xk, x_rs, x_cs, m, P, n_pad, kc, L, signs, scale);
      if (rc) return rc;
      rc = sparse_apply_t<T>(ctx, 0, s, n_pad, colptr, rowidx, val, rowptr, colidx, val_t, kc,
                             P, 1, n_pad, Yk, y_rs, y_cs, alpha, beta);
      if (rc) return rc;
    } else {
      rc = sparse_apply_t<T>(ctx, 1, s, n_pad, colptr, rowidx, val, rowptr, colidx, val_t, kc,
                             Xk, x_rs, x_cs, P, 1, n_pad, alpha, 0.0);
      if (rc) return rc;
      rc = fwht_passes<T>(ctx, nullptr, 0, 0, n_pad, P, n_pad, kc, L, nullptr, scale);
      if (rc) return rc;
      srht_adjoint_epilogue_kernel<T><<<dim3(static_cast<unsigned>(ceil_div(m * kc, 256))), 256,
                                        0, ctx->stream>>>(P, n_pad, m, kc, signs, Yk, y_rs, y_cs,
                                                          beta);
      RLA_LAUNCH_CHECK("fjlt_adjoint");
    }
  }
  return RLA_OK;
}

// The forward FJLT with K's rows given as the row-interleaved copy (ptr of
// K's CSR, ell_idx/ell_val from rla_sparse_ell_build): the same FWHT passes,
// then the interleaved gather, whose lanes read consecutive words of K.
// Bit-identical to fjlt_apply with adjoint = 0.
template <class T>
static int fjlt_apply_ell(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                          const int32_t* ptr, const int32_t* ell_idx, const T* ell_val,
                          const uint64_t* signs, double scale, const T* X, int64_t x_rs,
                          int64_t x_cs, T* Y, int64_t y_rs, int64_t y_cs, double alpha,
                          double beta) {
  RLA_CHECK_CTX(ctx);
  if (!is_pow2(n_pad) || n_pad < 2) return fail(RLA_EDIM, "Size of vector must be power of 2.");
  if (m < 1 || m > n_pad || s < 1 || ncols < 0)
    return fail(RLA_EARG, "FJLT sizes must satisfy 1 <= m <= n_pad and s >= 1");
  if (ncols == 0) return RLA_OK;
  if (!ptr || !ell_idx || !ell_val || !signs || !X || !Y)
    return fail(RLA_EARG, "FJLT apply: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  const int L = ilog2(n_pad);
  const int64_t cc = srht_chunk_cols<T>(n_pad, ncols);
  void* ws = nullptr;
  rc = workspace(ctx, static_cast<size_t>(n_pad * cc) * sizeof(T), &ws);
  if (rc) return rc;
  T* P = static_cast<T*>(ws);
  for (int64_t k0 = 0; k0 < ncols; k0 += cc) {
    const int64_t kc = std::min(cc, ncols - k0);
    rc = fwht_passes<T>(ctx, X + k0 * x_cs, x_rs, x_cs, m, P, n_pad, kc, L, signs, scale);
    if (rc) return rc;
    if constexpr (sizeof(T) == 4)
      rc = rla_sparse_apply_ell_f32(ctx, s, n_pad, ptr, ell_idx, ell_val, kc, P, 1, n_pad,
                                    Y + k0 * y_cs, y_rs, y_cs, alpha, beta);
    else
      rc = rla_sparse_apply_ell_f64(ctx, s, n_pad, ptr, ell_idx, ell_val, kc, P, 1, n_pad,
                                    Y + k0 * y_cs, y_rs, y_cs, alpha, beta);
    if (rc) return rc;
  }
  return RLA_OK;
}

// The forward product through the block-interleaved copy of K (rows s, X
// blocks of rb rows): the same transform, then rla_sparse_apply_bell, whose
// sums run in the same order as the CSR gather's.
template <class T>
static int fjlt_apply_bell(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                           int64_t rb, const int32_t* widths, const int32_t* bidx, const T* bval,
                           const uint64_t* signs, double scale, const T* X, int64_t x_rs,
                           int64_t x_cs, T* Y, int64_t y_rs, int64_t y_cs, double alpha,
                           double beta) {
  RLA_CHECK_CTX(ctx);
  if (!is_pow2(n_pad) || n_pad < 2) return fail(RLA_EDIM, "Size of vector must be power of 2.");
  if (m < 1 || m > n_pad || s < 1 || ncols < 0)
    return fail(RLA_EARG, "FJLT sizes must satisfy 1 <= m <= n_pad and s >= 1");
  if (ncols == 0) return RLA_OK;
  if (!widths || !bidx || !bval || !signs || !X || !Y)
    return fail(RLA_EARG, "FJLT apply: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  const int L = ilog2(n_pad);
  const int64_t cc = srht_chunk_cols<T>(n_pad, ncols);
  void* ws = nullptr;
  rc = workspace(ctx, static_cast<size_t>(n_pad * cc) * sizeof(T), &ws);
  if (rc) return rc;
  T* P = static_cast<T*>(ws);
  for (int64_t k0 = 0; k0 < ncols; k0 += cc) {
    const int64_t kc = std::min(cc, ncols - k0);
    rc = fwht_passes<T>(ctx, X + k0 * x_cs, x_rs, x_cs, m, P, n_pad, kc, L, signs, scale);
    if (rc) return rc;
    if constexpr (sizeof(T) == 4)
      rc = rla_sparse_apply_bell_f32(ctx, s, n_pad, rb, widths, bidx, bval, kc, P, 1, n_pad,
                                     Y + k0 * y_cs, y_rs, y_cs, alpha, beta);
    else
      rc = rla_sparse_apply_bell_f64(ctx, s, n_pad, rb, widths, bidx, bval, kc, P, 1, n_pad,
                                     Y + k0 * y_cs, y_rs, y_cs, alpha, beta);
    if (rc) return rc;
  }
  return RLA_OK;
}

// Forward FJLT with K's row-group tile copy (values carried): each column
// chunk's FWHT goes to the first part of the workspace, the ranges' sums of
// the tile gather to the second; within rounding of the CSR gather.
static int fjlt_apply_tile(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                           int64_t rec_words, const uint32_t* tiles, const uint64_t* signs,
                           double scale, const float* X, int64_t x_rs, int64_t x_cs, float* Y,
                           int64_t y_rs, int64_t y_cs, double alpha, double beta) {
  RLA_CHECK_CTX(ctx);
  if (!is_pow2(n_pad) || n_pad < 2) return fail(RLA_EDIM, "Size of vector must be power of 2.");
  if (m < 1 || m > n_pad || s < 1 || ncols < 0)
    return fail(RLA_EARG, "FJLT sizes must satisfy 1 <= m <= n_pad and s >= 1");
  if (ncols == 0) return RLA_OK;
  if (!tiles || !signs || !X || !Y) return fail(RLA_EARG, "FJLT apply: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  const int L = ilog2(n_pad);
  const int64_t cc = srht_chunk_cols<float>(n_pad, ncols);
  const size_t p_bytes = (static_cast<size_t>(n_pad * cc) * sizeof(float) + 255) / 256 * 256;
  // sums for a full chunk of columns or the narrower last one (more ranges)
  size_t part_bytes = sparse_tile_part_bytes(s, n_pad, cc);
  if (ncols % cc) part_bytes = std::max(part_bytes, sparse_tile_part_bytes(s, n_pad, ncols % cc));
  void* ws = nullptr;
  rc = workspace(ctx, p_bytes + part_bytes, &ws);
  if (rc) return rc;
  float* P = static_cast<float*>(ws);
  float* part = reinterpret_cast<float*>(static_cast<char*>(ws) + p_bytes);
  for (int64_t k0 = 0; k0 < ncols; k0 += cc) {
    const int64_t kc = std::min(cc, ncols - k0);
    rc = fwht_passes<float>(ctx, X + k0 * x_cs, x_rs, x_cs, m, P, n_pad, kc, L, signs, scale);
    if (rc) return rc;
    rc = sparse_apply_tile_part(ctx, s, n_pad, 1, rec_words, tiles, 1.0, kc, P, 1, n_pad,
                                Y + k0 * y_cs, y_rs, y_cs, alpha, beta, part);
    if (rc) return rc;
  }
  return RLA_OK;
}

}  // namespace rla

using namespace rla;

extern "C" {

int rla_fjlt_apply_tile_f32(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                            int64_t rec_words, const uint32_t* tiles,
                            const uint64_t* sign_chunks, double scale, const float* X,
                            int64_t x_rs, int64_t x_cs, float* Y, int64_t y_rs, int64_t y_cs,
                            double alpha, double beta) {
  return fjlt_apply_tile(ctx, m, n_pad, s, ncols, rec_words, tiles, sign_chunks, scale, X, x_rs,
                         x_cs, Y, y_rs, y_cs, alpha, beta);
}

int rla_fjlt_apply_f32(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                       int64_t ncols, const int32_t* colptr, const int32_t* rowidx,
                       const float* val, const int32_t* rowptr, const int32_t* colidx,
                       const float* val_t, const uint64_t* sign_chunks, double scale,
                       const float* X, int64_t x_rs, int64_t x_cs, float* Y, int64_t y_rs,
                       int64_t y_cs, double alpha, double beta) {
  return fjlt_apply<float>(ctx, adjoint, m, n_pad, s, ncols, colptr, rowidx, val, rowptr, colidx,
                           val_t, sign_chunks, scale, X, x_rs, x_cs, Y, y_rs, y_cs, alpha, beta);
}

int rla_fjlt_apply_f64(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                       int64_t ncols, const int32_t* colptr, const int32_t* rowidx,
                       const double* val, const int32_t* rowptr, const int32_t* colidx,
                       const double* val_t, const uint64_t* sign_chunks, double scale,
                       const double* X, int64_t x_rs, int64_t x_cs, double* Y, int64_t y_rs,
                       int64_t y_cs, double alpha, double beta) {
  return fjlt_apply<double>(ctx, adjoint, m, n_pad, s, ncols, colptr, rowidx, val, rowptr,
                            colidx, val_t, sign_chunks, scale, X, x_rs, x_cs, Y, y_rs, y_cs,
                            alpha, beta);
}

int rla_fwht_f32(rla_ctx* ctx, int64_t n, int64_t ncols, const uint64_t* sign_chunks,
                 double scaling, float* X, int64_t ldx) {
  return fwht<float>(ctx, n, ncols, sign_chunks, scaling, X, ldx);
}

int rla_fwht_f64(rla_ctx* ctx, int64_t n, int64_t ncols, const uint64_t* sign_chunks,
                 double scaling, double* X, int64_t ldx) {
  return fwht<double>(ctx, n, ncols, sign_chunks, scaling, X, ldx);
}

int rla_srht_apply_f32(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                       int64_t ncols, const int32_t* op, const uint64_t* sign_chunks,
                       double scale, const void* plan, const float* X, int64_t x_rs,
                       int64_t x_cs, float* Y, int64_t y_rs, int64_t y_cs, double alpha,
                       double beta) {
  return srht_apply<float>(ctx, adjoint, m, n_pad, s, ncols, op, sign_chunks, scale, plan, X,
                           x_rs, x_cs, Y, y_rs, y_cs, alpha, beta, false);
}

int rla_srht_apply_exact_f32(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                             int64_t ncols, const int32_t* op, const uint64_t* sign_chunks,
                             double scale, const void* plan, const float* X, int64_t x_rs,
                             int64_t x_cs, float* Y, int64_t y_rs, int64_t y_cs, double alpha,
                             double beta) {
  return srht_apply<float>(ctx, adjoint, m, n_pad, s, ncols, op, sign_chunks, scale, plan, X,
                           x_rs, x_cs, Y, y_rs, y_cs, alpha, beta, true);
}

int rla_srht_apply_f64(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                       int64_t ncols, const int32_t* op, const uint64_t* sign_chunks,
                       double scale, const void* plan, const double* X, int64_t x_rs,
                       int64_t x_cs, double* Y, int64_t y_rs, int64_t y_cs, double alpha,
                       double beta) {
  return srht_apply<double>(ctx, adjoint, m, n_pad, s, ncols, op, sign_chunks, scale, plan, X,
                            x_rs, x_cs, Y, y_rs, y_cs, alpha, beta, true);
}

int rla_fjlt_apply_ell_f32(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                           const int32_t* ptr, const int32_t* ell_idx, const float* ell_val,
                           const uint64_t* sign_chunks, double scale, const float* X,
                           int64_t x_rs, int64_t x_cs, float* Y, int64_t y_rs, int64_t y_cs,
                           double alpha, double beta) {
  return fjlt_apply_ell<float>(ctx, m, n_pad, s, ncols, ptr, ell_idx, ell_val, sign_chunks, scale,
                               X, x_rs, x_cs, Y, y_rs, y_cs, alpha, beta);
}

int rla_fjlt_apply_ell_f64(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                           const int32_t* ptr, const int32_t* ell_idx, const double* ell_val,
                           const uint64_t* sign_chunks, double scale, const double* X,
                           int64_t x_rs, int64_t x_cs, double* Y, int64_t y_rs, int64_t y_cs,
                           double alpha, double beta) {
  return fjlt_apply_ell<double>(ctx, m, n_pad, s, ncols, ptr, ell_idx, ell_val, sign_chunks,
                                scale, X, x_rs, x_cs, Y, y_rs, y_cs, alpha, beta);
}

int rla_fjlt_apply_bell_f32(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                            int64_t rb, const int32_t* widths, const int32_t* bidx,
                            const float* bval, const uint64_t* sign_chunks, double scale,
                            const float* X, int64_t x_rs, int64_t x_cs, float* Y, int64_t y_rs,
                            int64_t y_cs, double alpha, double beta) {
  return fjlt_apply_bell<float>(ctx, m, n_pad, s, ncols, rb, widths, bidx, bval, sign_chunks,
                                scale, X, x_rs, x_cs, Y, y_rs, y_cs, alpha, beta);
}

int rla_fjlt_apply_bell_f64(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                            int64_t rb, const int32_t* widths, const int32_t* bidx,
                            const double* bval, const uint64_t* sign_chunks, double scale,
                            const double* X, int64_t x_rs, int64_t x_cs, double* Y, int64_t y_rs,
                            int64_t y_cs, double alpha, double beta) {
  return fjlt_apply_bell<double>(ctx, m, n_pad, s, ncols, rb, widths, bidx, bval, sign_chunks,
                                 scale, X, x_rs, x_cs, Y, y_rs, y_cs, alpha, beta);
}

}  // extern "C"
