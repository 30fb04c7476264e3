// Dense products of strided views on the FP32 / FP64 matrix cores:
//   C = alpha * A * B + beta * C,
// A (M x K) at A[i*a_rs + l*a_cs], B (K x N) at B[l*b_rs + j*b_cs], C at
// C[i*c_rs + j*c_cs].  Any strides, so Julia's column-major arrays, their
// adjoints and views map to stride swaps without copies.  This is the
// Gaussian compressor's mul! (gaussian.jl:145-170: mul!(C, S.op, A, alpha,
// beta) and mul!(C, A, S.op, alpha, beta)).
//
// Tile 64 x 64 per workgroup: 4 wavefronts, each 32 x 32 = 2 x 2 blocks of
// v_mfma_f32_16x16x4_f32 (FP32) or v_mfma_f64_16x16x4_f64 (FP64), so products
// are formed and accumulated in the data's own precision.  K runs in stages
// of 16 through two LDS buffers: the next stage's global loads are issued
// into registers before the current stage's MFMAs and stored after them, one
// barrier per stage.  Each operand is staged with consecutive threads along
// its unit-stride index (rows or columns), so global loads coalesce for
// column-major and row-major views alike.  Tall-skinny products (few tiles,
// long K) split K over workgroups; the partial tiles are summed in split
// order by a second kernel, so results do not depend on timing.
#include <algorithm>
#include <mutex>

#include "rla_internal.hpp"

namespace rla {

namespace {

constexpr int kGT = 64;   // tile rows and columns
constexpr int kGK = 16;   // K per stage
constexpr int kGThreads = 256;
constexpr int kGPerThread = kGT * kGK / kGThreads;  // elements of each operand per thread

template <class T>
struct Mma;

template <>
struct Mma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // row (within the 16 x 16 block) of accumulator element r of a lane
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <>
struct Mma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

struct GemmArgs {
  int64_t M, N, K;
  const void* A;
  int64_t a_rs, a_cs;
  const void* B;
  int64_t b_rs, b_cs;
  void* C;
  int64_t c_rs, c_cs;
  double alpha, beta;
  int ksplit;
  int64_t kps;      // K per split (a multiple of kGK)
  int64_t tiles_m;
  void* part;       // ksplit x N x M partial sums when ksplit > 1
};

// staging map of element e (0 .. kGT*kGK-1) of a tile: (index along the
// tile's 64-wide side, index along K), consecutive e along the unit-stride one
__device__ __forceinline__ void stage_map(int e, bool wide_unit, int& w, int& l) {
  if (wide_unit) {
    w = e % kGT;
    l = e / kGT;
  } else {
    l = e % kGK;
    w = e / kGK;
  }
}

template <class T>
__global__ __launch_bounds__(kGThreads) void gemm_mfma_kernel(GemmArgs g) {
  using acc_t = typename Mma<T>::acc_t;
  __shared__ T As[2][kGK][kGT + 1];  // As[l][i]
  __shared__ T Bs[2][kGK][kGT + 1];  // Bs[l][j]
  const T* __restrict__ A = static_cast<const T*>(g.A);
  const T* __restrict__ B = static_cast<const T*>(g.B);
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int split = blockIdx.x % g.ksplit;
  const int64_t tile = blockIdx.x / g.ksplit;
  const int64_t i0 = (tile % g.tiles_m) * kGT, j0 = (tile / g.tiles_m) * kGT;
  const int64_t kb = split * g.kps, ke = std::min<int64_t>(g.K, kb + g.kps);
  const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
  // A tile: rows i0.., unit stride along rows when a_rs == 1; B tile:
  // columns j0.., unit stride along columns when b_cs == 1
  const bool a_rows_unit = g.a_rs == 1;
  const bool b_cols_unit = g.b_cs == 1 && g.b_rs != 1;
  int aw[kGPerThread], al[kGPerThread], bw[kGPerThread], bl[kGPerThread];
#pragma unroll
  for (int q = 0; q < kGPerThread; ++q) {
    stage_map(t + q * kGThreads, a_rows_unit, aw[q], al[q]);
    stage_map(t + q * kGThreads, b_cols_unit, bw[q], bl[q]);
  }
  T ra[kGPerThread], rb[kGPerThread];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < kGPerThread; ++q) {
      const int64_t gi = i0 + aw[q], gl = k0 + al[q];
      ra[q] = (gi < g.M && gl < ke) ? A[gi * g.a_rs + gl * g.a_cs] : T(0);
      const int64_t gj = j0 + bw[q], gl2 = k0 + bl[q];
      rb[q] = (gj < g.N && gl2 < ke) ? B[gl2 * g.b_rs + gj * g.b_cs] : T(0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < kGPerThread; ++q) {
      As[buf][al[q]][aw[q]] = ra[q];
      Bs[buf][bl[q]][bw[q]] = rb[q];
    }
  };

  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  int buf = 0;
  if (kb < ke) {
    load(kb);
    store(0);
  }
  __syncthreads();
  for (int64_t k0 = kb; k0 < ke; k0 += kGK) {
    const bool more = k0 + kGK < ke;
    if (more) load(k0 + kGK);  // in flight during this stage's MFMAs
#pragma unroll
    for (int kk = 0; kk < kGK; kk += 4) {
      const int l = kk + (lane >> 4);
      T av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = As[buf][l][wi + a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = Bs[buf][l][wj + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Mma<T>::run(av[a], bv[b], acc[a][b]);
    }
    if (more) store(buf ^ 1);  // the other buffer: last read a stage ago
    __syncthreads();
    buf ^= 1;
  }

  T* __restrict__ C = static_cast<T*>(g.C);
  T* __restrict__ part = static_cast<T*>(g.part);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wi + a * 16 + Mma<T>::row(lane, r);
        const int64_t j = j0 + wj + b * 16 + (lane & 15);
        if (i >= g.M || j >= g.N) continue;
        const T v = acc[a][b][r];
        if (g.ksplit > 1) {
          part[(int64_t(split) * g.N + j) * g.M + i] = v;
        } else {
          T* cp = C + i * g.c_rs + j * g.c_cs;
          *cp = g.beta == 0.0 ? T(g.alpha) * v : T(g.alpha) * v + T(g.beta) * (*cp);
        }
      }
}

template <class T>
__global__ void gemm_mfma_reduce_kernel(GemmArgs g) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= g.M * g.N) return;
  const int64_t i = e % g.M, j = e / g.M;
  const T* __restrict__ part = static_cast<const T*>(g.part);
  T s = T(0);
  for (int q = 0; q < g.ksplit; ++q) s += part[(int64_t(q) * g.N + j) * g.M + i];
  T* cp = static_cast<T*>(g.C) + i * g.c_rs + j * g.c_cs;
  *cp = g.beta == 0.0 ? T(g.alpha) * s : T(g.alpha) * s + T(g.beta) * (*cp);
}

// split count: enough workgroups to fill the chip, at least 256 of K each
int choose_ksplit(int64_t tiles, int64_t K) {
  int ks = 1;
  while (ks < 64 && tiles * ks < 1024 && K / (2 * ks) >= 256) ks *= 2;
  return ks;
}

__global__ void scale_kernel_f32(int64_t M, int64_t N, float* C, int64_t c_rs, int64_t c_cs,
                                 double beta) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= M * N) return;
  float* cp = C + (e % M) * c_rs + (e / M) * c_cs;
  *cp = beta == 0.0 ? 0.0f : float(beta) * (*cp);
}
__global__ void scale_kernel_f64(int64_t M, int64_t N, double* C, int64_t c_rs, int64_t c_cs,
                                 double beta) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= M * N) return;
  double* cp = C + (e % M) * c_rs + (e / M) * c_cs;
  *cp = beta == 0.0 ? 0.0 : beta * (*cp);
}

}  // namespace

// Partial sums of split products go to `scratch` (scratch_elems elements)
// when it is given, else to the context workspace.
template <class T>
static int gemm_mfma_impl(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, double alpha,
                          const T* A, int64_t a_rs, int64_t a_cs, const T* B, int64_t b_rs,
                          int64_t b_cs, double beta, T* C, int64_t c_rs, int64_t c_cs,
                          T* scratch, size_t scratch_elems) {
  if (M < 0 || N < 0 || K < 0) return fail(RLA_EARG, "gemm: negative size");
  if (M == 0 || N == 0) return RLA_OK;
  if (!C || (K > 0 && (!A || !B))) return fail(RLA_EARG, "gemm: NULL operand");
  int rc = use_device(ctx);
  if (rc) return rc;
  if (K == 0) {  // C = beta * C
    const int64_t n = M * N;
    if constexpr (sizeof(T) == 4)
      scale_kernel_f32<<<dim3(unsigned(ceil_div(n, 256))), 256, 0, ctx->stream>>>(M, N, C, c_rs,
                                                                                 c_cs, beta);
    else
      scale_kernel_f64<<<dim3(unsigned(ceil_div(n, 256))), 256, 0, ctx->stream>>>(M, N, C, c_rs,
                                                                                 c_cs, beta);
    RLA_LAUNCH_CHECK("gemm scale");
    return RLA_OK;
  }
  GemmArgs g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.a_rs = a_rs;
  g.a_cs = a_cs;
  g.B = B;
  g.b_rs = b_rs;
  g.b_cs = b_cs;
  g.C = C;
  g.c_rs = c_rs;
  g.c_cs = c_cs;
  g.alpha = alpha;
  g.beta = beta;
  g.tiles_m = ceil_div(M, kGT);
  const int64_t tiles = g.tiles_m * ceil_div(N, kGT);
  g.ksplit = choose_ksplit(tiles, K);
  g.kps = ceil_div(ceil_div(K, g.ksplit), kGK) * kGK;
  g.ksplit = static_cast<int>(ceil_div(K, g.kps));  // no empty splits
  if (g.ksplit > 1) {
    const size_t need = size_t(g.ksplit) * M * N;
    if (scratch) {
      if (need > scratch_elems)
        return fail(RLA_ENOTSUP, "gemm: the split partial sums (" + std::to_string(need * sizeof(T)) +
                                     " bytes) exceed the allocation-free scratch (" +
                                     std::to_string(scratch_elems * sizeof(T)) + " bytes)");
      g.part = scratch;
    } else {
      void* ws = nullptr;
      rc = workspace(ctx, need * sizeof(T), &ws);
      if (rc) return rc;
      g.part = ws;
    }
  }
  const int64_t blocks = tiles * g.ksplit;
  if (blocks > INT32_MAX) return fail(RLA_ENOTSUP, "gemm: too many tiles");
  gemm_mfma_kernel<T><<<dim3(static_cast<unsigned>(blocks)), kGThreads, 0, ctx->stream>>>(g);
  RLA_LAUNCH_CHECK("gemm_mfma");
  if (g.ksplit > 1) {
    gemm_mfma_reduce_kernel<T><<<dim3(unsigned(ceil_div(M * N, 256))), 256, 0, ctx->stream>>>(g);
    RLA_LAUNCH_CHECK("gemm_mfma_reduce");
  }
  return RLA_OK;
}

template <class T>
int gemm_mfma(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, double alpha, const T* A,
              int64_t a_rs, int64_t a_cs, const T* B, int64_t b_rs, int64_t b_cs, double beta,
              T* C, int64_t c_rs, int64_t c_cs) {
  return gemm_mfma_impl<T>(ctx, M, N, K, alpha, A, a_rs, a_cs, B, b_rs, b_cs, beta, C, c_rs, c_cs,
                           nullptr, 0);
}

template int gemm_mfma<float>(rla_ctx*, int64_t, int64_t, int64_t, double, const float*, int64_t,
                              int64_t, const float*, int64_t, int64_t, double, float*, int64_t,
                              int64_t);
template int gemm_mfma<double>(rla_ctx*, int64_t, int64_t, int64_t, double, const double*,
                               int64_t, int64_t, const double*, int64_t, int64_t, double,
                               double*, int64_t, int64_t);

// Operands of a Gaussian apply as a product: side 0 (Left) C = S A, side 1
// (Right) C = A S, S given as a strided view (an adjoint is a stride swap).
template <class T>
static int gaussian_apply(rla_ctx* ctx, int side, int64_t s_rows, int64_t s_cols, const T* S,
                          int64_t s_rs, int64_t s_cs, int64_t a_rows, int64_t a_cols, const T* A,
                          int64_t a_rs, int64_t a_cs, double alpha, double beta, T* C,
                          int64_t c_rs, int64_t c_cs, T* scratch = nullptr,
                          size_t scratch_elems = 0) {
  RLA_CHECK_CTX(ctx);
  if (side != 0 && side != 1) return fail(RLA_EARG, "rla_gaussian_apply: side is 0 or 1");
  if (side == 0) {
    if (s_cols != a_rows)
      return fail(RLA_EDIM, "Number of columns of S, " + std::to_string(s_cols) +
                                ", does not match number of rows of A, " +
                                std::to_string(a_rows) + ".");
    return gemm_mfma_impl<T>(ctx, s_rows, a_cols, s_cols, alpha, S, s_rs, s_cs, A, a_rs, a_cs,
                             beta, C, c_rs, c_cs, scratch, scratch_elems);
  }
  if (a_cols != s_rows)
    return fail(RLA_EDIM, "Number of columns of A, " + std::to_string(a_cols) +
                              ", does not match number of rows of S, " +
                              std::to_string(s_rows) + ".");
  return gemm_mfma_impl<T>(ctx, a_rows, s_cols, a_cols, alpha, A, a_rs, a_cs, S, s_rs, s_cs, beta,
                           C, c_rs, c_cs, scratch, scratch_elems);
}

// ---- allocation-free Gaussian apply ------------------------------------------
// Split partial sums in a file-scope device array instead of the context
// workspace.  choose_ksplit stops at fewer than 2048 (tile, split) pairs of
// 64 x 64, so partials stay below 2048 x 4096 = 8 Mi elements and every shape
// fits; a larger need would be rejected.  Host code uses the array through
// its device address, looked up once per device.  Calls that overlap on one
// device share the array and race.
constexpr size_t kNoallocScratch = size_t(8) << 20;  // elements (32 MiB of Float32)
__device__ __align__(256) float scratch[kNoallocScratch];

namespace {
std::mutex g_scratch_mu;
float* g_scratch_ptr[64];
}  // namespace

// the calling thread's current device must be `device`
static int scratch_address(int device, float** out) {
  if (device < 0 || device >= 64) return fail(RLA_EARG, "scratch: device out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    RLA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[device];
  return RLA_OK;
}

}  // namespace rla

using namespace rla;

extern "C" {

int rla_gemm_strided_f32(rla_ctx* ctx, int64_t m, int64_t n, int64_t k, double alpha,
                         const float* A, int64_t a_rs, int64_t a_cs, const float* B, int64_t b_rs,
                         int64_t b_cs, double beta, float* C, int64_t c_rs, int64_t c_cs) {
  RLA_CHECK_CTX(ctx);
  return gemm_mfma<float>(ctx, m, n, k, alpha, A, a_rs, a_cs, B, b_rs, b_cs, beta, C, c_rs, c_cs);
}

int rla_gemm_strided_f64(rla_ctx* ctx, int64_t m, int64_t n, int64_t k, double alpha,
                         const double* A, int64_t a_rs, int64_t a_cs, const double* B,
                         int64_t b_rs, int64_t b_cs, double beta, double* C, int64_t c_rs,
                         int64_t c_cs) {
  RLA_CHECK_CTX(ctx);
  return gemm_mfma<double>(ctx, m, n, k, alpha, A, a_rs, a_cs, B, b_rs, b_cs, beta, C, c_rs,
                           c_cs);
}

int rla_gaussian_apply_f32(rla_ctx* ctx, int side, int64_t s_rows, int64_t s_cols,
                           const float* S, int64_t s_rs, int64_t s_cs, int64_t a_rows,
                           int64_t a_cols, const float* A, int64_t a_rs, int64_t a_cs,
                           double alpha, double beta, float* C, int64_t c_rs, int64_t c_cs) {
  return gaussian_apply<float>(ctx, side, s_rows, s_cols, S, s_rs, s_cs, a_rows, a_cols, A, a_rs,
                               a_cs, alpha, beta, C, c_rs, c_cs);
}

int rla_gaussian_apply_noalloc_f32(rla_ctx* ctx, int side, int64_t s_rows, int64_t s_cols,
                                   const float* S, int64_t s_rs, int64_t s_cs, int64_t a_rows,
                                   int64_t a_cols, const float* A, int64_t a_rs, int64_t a_cs,
                                   double alpha, double beta, float* C, int64_t c_rs,
                                   int64_t c_cs) {
  RLA_CHECK_CTX(ctx);
  int rc = use_device(ctx);
  if (rc) return rc;
  float* scr = nullptr;
  rc = scratch_address(ctx->device, &scr);
  if (rc) return rc;
  return gaussian_apply<float>(ctx, side, s_rows, s_cols, S, s_rs, s_cs, a_rows, a_cols, A, a_rs,
                               a_cs, alpha, beta, C, c_rs, c_cs, scr, kNoallocScratch);
}

int rla_gaussian_apply_f64(rla_ctx* ctx, int side, int64_t s_rows, int64_t s_cols,
                           const double* S, int64_t s_rs, int64_t s_cs, int64_t a_rows,
                           int64_t a_cols, const double* A, int64_t a_rs, int64_t a_cs,
                           double alpha, double beta, double* C, int64_t c_rs, int64_t c_cs) {
  return gaussian_apply<double>(ctx, side, s_rows, s_cols, S, s_rs, s_cs, a_rows, a_cols, A,
                                a_rs, a_cs, alpha, beta, C, c_rs, c_cs);
}

}  // extern "C"
