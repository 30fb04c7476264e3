// BF16 MFMA GEMM with FP32 accumulation, and the BF16 conversion kernels.
//
// C[i, j] = alpha * sum_l A[l + i*lda] * B[l + j*ldb] + beta * C[i, j]
//
// Both operands are K-contiguous ("TN"): row i of op(A) and column j of B are
// contiguous runs of K bf16 values.  The RangeFinder keeps A resident as two
// BF16 copies, row-major (rows of A contiguous: for A*X) and column-major
// (columns of A contiguous: for A'*X), so every product streams A once with
// coalesced 16-byte loads.
//
// Workgroup tile: 128 rows (4 wavefronts x 32 rows) x NB*32 columns.  Per
// 64-wide K step the B tile (NB*32 columns x 64) is staged in LDS (padded
// row stride, conflict-free 16-byte fragment reads) and every wavefront runs
// 4 x NB v_mfma_f32_32x32x16_bf16 on its 32 rows.  The A fragments and the
// next B tile are loaded while the current tile is multiplied.  Split-K
// writes partial tiles that a second kernel sums in split order, so results
// are deterministic.
#include <algorithm>

#include "rla_internal.hpp"

namespace rla {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128;        // rows per workgroup
constexpr int kBK = 64;         // K per step
constexpr int kBstride = 72;    // bf16 per staged B column (64 + 8 pad)

// Round-to-nearest-even float -> bf16 bits (NaN kept quiet).
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return uint16_t((u >> 16) | 0x40);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

// dst[l + i*ldd] = bf16(src[i, l]) for i < rows, l < cols of a strided
// source (the source's (rows x cols) view); i.e. rows of the view become
// contiguous runs.  Tiled through LDS so both sides stay coalesced.
template <class T>
__global__ __launch_bounds__(256) void to_bf16_rows_kernel(int64_t rows, int64_t cols,
                                                           const T* __restrict__ src,
                                                           int64_t s_rs, int64_t s_cs,
                                                           uint16_t* __restrict__ dst,
                                                           int64_t ldd) {
  __shared__ float tile[64][65];
  const int64_t i0 = int64_t(blockIdx.y) * 64, l0 = int64_t(blockIdx.x) * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // read: consecutive threads along the source's unit-stride direction
  const bool rows_fast = s_rs == 1;
  for (int r = ty; r < 64; r += 4) {
    const int64_t i = rows_fast ? i0 + tx : i0 + r;
    const int64_t l = rows_fast ? l0 + r : l0 + tx;
    float v = 0.f;
    if (i < rows && l < cols) v = static_cast<float>(src[i * s_rs + l * s_cs]);
    if (rows_fast) tile[tx][r] = v; else tile[r][tx] = v;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int64_t i = i0 + r, l = l0 + tx;
    if (i < rows && l < cols) dst[l + i * ldd] = f32_to_bf16_bits(tile[r][tx]);
  }
}

// 8 bf16 at p with entries at or beyond n_valid replaced by zero
__device__ __forceinline__ u32x4 load8_tail(const uint16_t* p, int64_t n_valid) {
  uint32_t w[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t lo = (2 * q < n_valid) ? p[2 * q] : 0u;
    const uint32_t hi = (2 * q + 1 < n_valid) ? p[2 * q + 1] : 0u;
    w[q] = lo | (hi << 16);
  }
  return u32x4{w[0], w[1], w[2], w[3]};
}

template <int NB>
__global__ __launch_bounds__(256) void gemm_bf16_tn_kernel(
    int64_t M, int64_t N, int64_t K, const uint16_t* __restrict__ A, int64_t lda,
    const uint16_t* __restrict__ B, int64_t ldb, float* __restrict__ C, int64_t c_rs,
    int64_t c_cs, float alpha, float beta, int ksplit, int64_t k_per_split,
    float* __restrict__ part) {
  constexpr int BN = NB * 32;
  __shared__ __align__(16) uint16_t Bs[BN * kBstride];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tiles_n = static_cast<int>((N + BN - 1) / BN);
  const int bid = blockIdx.x;
  const int split = bid % ksplit;
  const int tn = (bid / ksplit) % tiles_n;
  const int64_t tm = (bid / ksplit) / tiles_n;
  const int64_t row0 = tm * kBM + wave * 32;
  const int64_t col0 = int64_t(tn) * BN;
  const int64_t kb = split * k_per_split;
  const int64_t ke = std::min<int64_t>(K, kb + k_per_split);

  f32x16 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[nb][r] = 0.f;

  // A fragment source: row row0 + (lane & 31), K offset (lane >> 5) * 8
  const int64_t arow = row0 + (lane & 31);
  const bool arow_ok = arow < M;
  const uint16_t* ap = A + (arow_ok ? arow : 0) * lda;
  const int akq = (lane >> 5) * 8;

  // B staging: 256 threads move BN columns x 64 bf16 (8 chunks of 16 B per column)
  constexpr int CHUNKS = BN * 8;
  constexpr int PER_T = (CHUNKS + 255) / 256;
  u32x4 breg[PER_T];
  auto load_b = [&](int64_t k0) {
#pragma unroll
    for (int u = 0; u < PER_T; ++u) {
      const int c = threadIdx.x + u * 256;
      const int col = c >> 3, part8 = c & 7;
      const int64_t gcol = col0 + col;
      const int64_t gk = k0 + part8 * 8;
      u32x4 v = {0u, 0u, 0u, 0u};
      if (c < CHUNKS && gcol < N) {
        const uint16_t* bp = B + gcol * ldb + gk;
        if (gk + 8 <= ke) v = *reinterpret_cast<const u32x4*>(bp);
        else if (gk < ke) v = load8_tail(bp, ke - gk);
      }
      breg[u] = v;
    }
  };
  auto store_b = [&]() {
#pragma unroll
    for (int u = 0; u < PER_T; ++u) {
      const int c = threadIdx.x + u * 256;
      if (c < CHUNKS) {
        const int col = c >> 3, part8 = c & 7;
        *reinterpret_cast<u32x4*>(&Bs[col * kBstride + part8 * 8]) = breg[u];
      }
    }
  };
  auto load_a = [&](int64_t k0, bf16x8 (&af)[4]) {
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int64_t gk = k0 + kk * 16 + akq;
      u32x4 v = {0u, 0u, 0u, 0u};
      if (arow_ok) {
        if (gk + 8 <= ke) v = *reinterpret_cast<const u32x4*>(ap + gk);
        else if (gk < ke) v = load8_tail(ap + gk, ke - gk);
      }
      af[kk] = __builtin_bit_cast(bf16x8, v);
    }
  };

  bf16x8 af[4];
  if (kb < ke) {
    load_b(kb);
    load_a(kb, af);
  }
  for (int64_t k0 = kb; k0 < ke; k0 += kBK) {
    __syncthreads();  // previous tile's readers are done
    store_b();
    __syncthreads();
    bf16x8 af_cur[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) af_cur[kk] = af[kk];
    if (k0 + kBK < ke) {
      load_b(k0 + kBK);
      load_a(k0 + kBK, af);
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const u32x4 bv = *reinterpret_cast<const u32x4*>(
            &Bs[(nb * 32 + (lane & 31)) * kBstride + kk * 16 + akq]);
        acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af_cur[kk],
                                                          __builtin_bit_cast(bf16x8, bv),
                                                          acc[nb], 0, 0, 0);
      }
    }
  }

  // acc[nb][r]: row row0 + (r/4)*8 + (lane/32)*4 + (r%4), column col0 + nb*32 + lane%32
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int64_t j = col0 + nb * 32 + (lane & 31);
    if (j >= N) continue;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int64_t i = row0 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
      if (i >= M) continue;
      if (ksplit == 1) {
        float* cp = C + i * c_rs + j * c_cs;
        *cp = beta == 0.f ? alpha * acc[nb][r] : alpha * acc[nb][r] + beta * (*cp);
      } else {
        part[(int64_t(split) * N + j) * M + i] = acc[nb][r];
      }
    }
  }
}

__global__ void splitk_reduce_kernel(int64_t M, int64_t N, int ksplit,
                                     const float* __restrict__ part, float* __restrict__ C,
                                     int64_t c_rs, int64_t c_cs, float alpha, float beta) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= M * N) return;
  const int64_t i = e % M, j = e / M;
  float s = 0.f;
  for (int q = 0; q < ksplit; ++q) s += part[(int64_t(q) * N + j) * M + i];
  float* cp = C + i * c_rs + j * c_cs;
  *cp = beta == 0.f ? alpha * s : alpha * s + beta * (*cp);
}

static int num_cus_dev(int device) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 256;
  return p.multiProcessorCount;
}

template <int NB>
static int launch_gemm(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, const uint16_t* A,
                       int64_t lda, const uint16_t* B, int64_t ldb, float* C, int64_t c_rs,
                       int64_t c_cs, float alpha, float beta) {
  constexpr int BN = NB * 32;
  const int64_t tiles = ceil_div(M, kBM) * ceil_div(N, BN);
  static int cus = 0;
  if (!cus) cus = num_cus_dev(ctx->device);
  // split K until about one workgroup per CU, keeping >= 1024 K per split
  int ksplit = 1;
  while (tiles * ksplit * 2 <= cus && K / (ksplit * 2) >= 1024) ksplit *= 2;
  const int64_t kps = ceil_div(ceil_div(K, ksplit), kBK) * kBK;
  ksplit = static_cast<int>(ceil_div(K, kps));
  float* part = nullptr;
  if (ksplit > 1) {
    void* ws = nullptr;
    int rc = workspace(ctx, static_cast<size_t>(ksplit) * M * N * sizeof(float), &ws);
    if (rc) return rc;
    part = static_cast<float*>(ws);
  }
  gemm_bf16_tn_kernel<NB><<<dim3(static_cast<unsigned>(tiles * ksplit)), 256, 0, ctx->stream>>>(
      M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta, ksplit, kps, part);
  RLA_LAUNCH_CHECK("gemm_bf16");
  if (ksplit > 1) {
    splitk_reduce_kernel<<<dim3(static_cast<unsigned>(ceil_div(M * N, 256))), 256, 0,
                           ctx->stream>>>(M, N, ksplit, part, C, c_rs, c_cs, alpha, beta);
    RLA_LAUNCH_CHECK("gemm_bf16/splitk");
  }
  return RLA_OK;
}

int gemm_bf16_tn(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, float alpha, const uint16_t* A,
                 int64_t lda, const uint16_t* B, int64_t ldb, float beta, float* C, int64_t c_rs,
                 int64_t c_cs) {
  if (M <= 0 || N <= 0) return RLA_OK;
  if (K <= 0) return fail(RLA_EARG, "gemm_bf16: K must be positive");
  if ((lda % 8) || (ldb % 8) || (reinterpret_cast<uintptr_t>(A) % 16) ||
      (reinterpret_cast<uintptr_t>(B) % 16))
    return fail(RLA_EARG, "gemm_bf16: operands must be 16-byte aligned with ld a multiple of 8");
  // widest N tile that covers N in one pass (up to 288 columns), else 256
  const int64_t nbk = ceil_div(N, 32);
  if (nbk <= 4) return launch_gemm<4>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
  if (nbk <= 6) return launch_gemm<6>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
  if (nbk <= 9) return launch_gemm<9>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
  return launch_gemm<8>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
}

template <class T>
int to_bf16_rows(rla_ctx* ctx, int64_t rows, int64_t cols, const T* src, int64_t s_rs,
                 int64_t s_cs, uint16_t* dst, int64_t ldd) {
  if (rows <= 0 || cols <= 0) return RLA_OK;
  dim3 grid(static_cast<unsigned>(ceil_div(cols, 64)), static_cast<unsigned>(ceil_div(rows, 64)));
  to_bf16_rows_kernel<T><<<grid, 256, 0, ctx->stream>>>(rows, cols, src, s_rs, s_cs, dst, ldd);
  RLA_LAUNCH_CHECK("to_bf16");
  return RLA_OK;
}

template int to_bf16_rows<float>(rla_ctx*, int64_t, int64_t, const float*, int64_t, int64_t,
                                 uint16_t*, int64_t);
template int to_bf16_rows<double>(rla_ctx*, int64_t, int64_t, const double*, int64_t, int64_t,
                                  uint16_t*, int64_t);

}  // namespace rla

using namespace rla;

extern "C" {

int rla_to_bf16_f32(rla_ctx* ctx, int64_t rows, int64_t cols, const float* src, int64_t s_rs,
                    int64_t s_cs, uint16_t* dst, int64_t ldd) {
  RLA_CHECK_CTX(ctx);
  if (ldd < cols) return fail(RLA_EARG, "rla_to_bf16: ldd smaller than the run length");
  int rc = use_device(ctx);
  if (rc) return rc;
  return to_bf16_rows<float>(ctx, rows, cols, src, s_rs, s_cs, dst, ldd);
}

int rla_to_bf16_f64(rla_ctx* ctx, int64_t rows, int64_t cols, const double* src, int64_t s_rs,
                    int64_t s_cs, uint16_t* dst, int64_t ldd) {
  RLA_CHECK_CTX(ctx);
  if (ldd < cols) return fail(RLA_EARG, "rla_to_bf16: ldd smaller than the run length");
  int rc = use_device(ctx);
  if (rc) return rc;
  return to_bf16_rows<double>(ctx, rows, cols, src, s_rs, s_cs, dst, ldd);
}

int rla_gemm_bf16(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, float alpha, const uint16_t* A,
                  int64_t lda, const uint16_t* B, int64_t ldb, float beta, float* C, int64_t c_rs,
                  int64_t c_cs) {
  RLA_CHECK_CTX(ctx);
  int rc = use_device(ctx);
  if (rc) return rc;
  return gemm_bf16_tn(ctx, M, N, K, alpha, A, lda, B, ldb, beta, C, c_rs, c_cs);
}

}  // extern "C"
