// Cholesky factor of the small k x k Gram matrix of CholeskyQR (k <= 288),
// G = R'R with R upper and a positive diagonal, and the in-place triangular
// solve Y <- Y R^-1 that completes a CholeskyQR pass.
//
// Factorisation: one workgroup of 8 wavefronts.  The upper-triangular
// 16 x 16 blocks of G are spread round-robin over the wavefronts and stay in
// their MFMA accumulator registers for the whole factorisation (at most
// 18 x 19 / 2 blocks, 22 per wavefront), so the k^3/3 flops of the trailing
// updates touch no memory.  Block step s:
//   1. the owners of block row s copy it to LDS,
//   2. wavefront 0 factors the 16 x 16 diagonal block in LDS,
//   3. one thread per column solves R_ss' p = g for the panel row,
//   4. every wavefront subtracts P_i' P_j from its trailing blocks (MFMA).
// R is written column-major and row-major (the solve's operand layout), and
// the inverses of the diagonal blocks are formed at the end, one block per
// wavefront.
//
// A pivot that is not finite or not above k * eps * G_jj stops the
// factorisation and reports the 1-based column in *info (0 on success).  The
// relative floor (LAPACK's potrf only tests for <= 0) makes a numerically
// rank-deficient Y report a breakdown instead of returning a factor whose
// inverse is dominated by rounding error.
//
// Solve: each wavefront owns 16 rows of Y and walks the column blocks in
// order, Ynew_b = (Y_b - sum_{l<b} Ynew_l R_lb) inv(R_bb), keeping the solved
// blocks in registers in the MFMA operand layout; it reads only its own rows
// and writes each block after reading it, so Y is overwritten in place.  The
// solve is skipped on the device when *info != 0.
#include <cmath>

#include "rla_internal.hpp"

namespace rla {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kCholThreads = 512;
constexpr int kCholWaves = kCholThreads / 64;
constexpr int kMaxCholK = 288;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

__device__ __forceinline__ void upper_block(int q, int& bi, int& bj) {
  bj = 0;
  while (q > bj) {
    q -= bj + 1;
    ++bj;
  }
  bi = q;
}

template <int NBLK>
__global__ __launch_bounds__(kCholThreads) void chol_blocks_kernel(
    int k, const double* G, int ldg, double* R, int ldr, double* __restrict__ Rt, int ldt,
    double* __restrict__ Dinv, int* __restrict__ info) {
  constexpr int T = NBLK * (NBLK + 1) / 2;
  constexpr int MAXB = (T + kCholWaves - 1) / kCholWaves;
  constexpr int CP = NBLK * 16;
  __shared__ double P[16][CP + 1];  // block row s: G after updates, then R
  __shared__ double fl[CP];         // pivot floors k * eps * G_jj
  __shared__ int fail;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int lr = lane >> 4, lc = lane & 15;

  int bi[MAXB], bj[MAXB];
  f64x4 acc[MAXB];
#pragma unroll
  for (int u = 0; u < MAXB; ++u) {
    const int q = wave + kCholWaves * u;
    upper_block(q < T ? q : 0, bi[u], bj[u]);
    if (q >= T) bi[u] = bj[u] = NBLK;  // never active
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = bi[u] * 16 + lr + 4 * r, j = bj[u] * 16 + lc;
      acc[u][r] = (q < T && i < k && j < k) ? G[i + j * ldg] : 0.0;
    }
  }
  for (int j = t; j < CP; j += kCholThreads)
    fl[j] = j < k ? k * 2.220446049250313e-16 * G[j + j * ldg] : 0.0;
  if (t == 0) fail = 0;
  // zero R below the block diagonal (the diagonal blocks are written whole);
  // G's lower blocks are never read, so R may alias G
  for (int e = t; e < k * k; e += kCholThreads) {
    const int i = e % k, j = e / k;
    if ((i >> 4) > (j >> 4)) R[i + j * ldr] = 0.0;
  }
  __syncthreads();

  double* pl = &P[lr][lc];  // this lane's element of a 4 x 16 slab; slab rows are CP+1 apart
  for (int s = 0; s < NBLK; ++s) {
    const int js = s * 16;
    if (js >= k) break;
    const int nb = min(16, k - js);
    // re-derive the lane pointer each step so the compiler does not keep one
    // address register per block alive across the loop
    asm volatile("" : "+v"(pl));
    // 1. block row s to LDS
#pragma unroll
    for (int u = 0; u < MAXB; ++u)
      if (bi[u] == s) {
#pragma unroll
        for (int r = 0; r < 4; ++r) pl[4 * r * (CP + 1) + bj[u] * 16] = acc[u][r];
      }
    __syncthreads();
    // 2. diagonal block, wavefront 0; lanes own elements (i, c) = (e / 16, e % 16)
    if (wave == 0) {
      for (int j = 0; j < nb; ++j) {
        const double dj = P[j][js + j];
        if (!(dj > fl[js + j]) || !(dj > 0.0) || !isfinite(dj)) {
          if (lane == 0) fail = js + j + 1;
          break;
        }
        const double d = sqrt(dj), rd = 1.0 / d;
        wave_sync();
        if (lane >= j && lane < nb) P[j][js + lane] = lane == j ? d : P[j][js + lane] * rd;
        wave_sync();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = lane + 64 * u, i = e >> 4, c = e & 15;
          if (i > j && i <= c && c < nb) P[i][js + c] -= P[j][js + i] * P[j][js + c];
        }
        wave_sync();
      }
    }
    __syncthreads();
    if (fail) break;
    // 3. panel row (columns past the diagonal block) and R / Rt for block row s
    const int n2 = k - js - 16;
    if (t < n2) {
      // forward substitution in place in LDS (the trailing blocks hold the
      // registers); column c belongs to this thread alone
      const int c = js + 16 + t;
      for (int i = 0; i < 16; ++i) {
        double v = P[i][c];
        for (int l = 0; l < i; ++l) v -= P[l][js + i] * P[l][c];
        v /= P[i][js + i];
        P[i][c] = v;
        R[(js + i) + int64_t(c) * ldr] = v;
        Rt[int64_t(js + i) * ldt + c] = v;
      }
    }
    if (t < 256) {
      const int i = t >> 4, c = t & 15;
      if (i < nb && c < nb) {
        const double v = i <= c ? P[i][js + c] : 0.0;
        R[(js + i) + int64_t(js + c) * ldr] = v;
        Rt[int64_t(js + i) * ldt + js + c] = v;
      }
    }
    __syncthreads();
    // 4. trailing update of the register-resident blocks
#pragma unroll
    for (int u = 0; u < MAXB; ++u)
      if (bi[u] > s && bi[u] < NBLK) {
        double a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          a[q] = pl[4 * q * (CP + 1) + bi[u] * 16];
          b[q] = pl[4 * q * (CP + 1) + bj[u] * 16];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[q], b[q], acc[u], 0, 0, 0);
      }
    __syncthreads();
  }
  if (t == 0) *info = fail;
  if (fail) return;
  // inverses of the diagonal blocks: lane c < nb solves R_bb x = e_c
  for (int b = wave; b * 16 < k; b += kCholWaves) {
    const int js = b * 16, nb = min(16, k - js);
    if (lane < 16) {
      double x[16];
#pragma unroll
      for (int i = 15; i >= 0; --i) {
        double v = 0.0;
        if (i < nb && i <= lane && lane < nb) {
          v = i == lane ? 1.0 : 0.0;
#pragma unroll
          for (int l = 0; l < 16; ++l)
            if (l > i && l < nb) v -= Rt[int64_t(js + i) * ldt + js + l] * x[l];
          v /= Rt[int64_t(js + i) * ldt + js + i];
        }
        x[i] = v;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) Dinv[b * 256 + i * 16 + lane] = x[i];
    }
  }
}

template <int NBLK>
__global__ __launch_bounds__(256) void trsm_ru_f64_kernel(int64_t m, int k, double* Y, int64_t ldy,
                                                          const double* __restrict__ Rt, int ldt,
                                                          const double* __restrict__ Dinv,
                                                          const int* __restrict__ info) {
  if (info && *info) return;
  __shared__ double Ts[4][16][17];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane >> 4, lc = lane & 15;
  const int64_t rbase = int64_t(blockIdx.x) * 64 + wave * 16;
  const int64_t row = rbase + lc;  // operand layout: row lane%16, column 4q + lane/16
  const bool row_ok = row < m;
  double (*ts)[17] = Ts[wave];
  double ya[NBLK][4];
#pragma unroll
  for (int b = 0; b < NBLK; ++b) {
    if (b * 16 >= k) break;
    // Y_b in the accumulator layout (through LDS)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = b * 16 + 4 * q + lr;
      ts[lc][4 * q + lr] = (row_ok && c < k) ? Y[row + int64_t(c) * ldy] : 0.0;
    }
    wave_sync();
    f64x4 acc;
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = ts[lr + 4 * r][lc];
    wave_sync();
    // T = Y_b - sum_{l<b} Ynew_l R_lb
    const int cb = b * 16 + lc;
#pragma unroll
    for (int l = 0; l < b; ++l) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double rv = cb < k ? Rt[int64_t(l * 16 + 4 * q + lr) * ldt + cb] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ya[l][q], -rv, acc, 0, 0, 0);
      }
    }
    // Ynew_b = T inv(R_bb)
#pragma unroll
    for (int r = 0; r < 4; ++r) ts[lr + 4 * r][lc] = acc[r];
    wave_sync();
    double ta[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) ta[q] = ts[lc][4 * q + lr];
    wave_sync();
    acc = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ta[q], Dinv[b * 256 + (4 * q + lr) * 16 + lc],
                                                 acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) ts[lr + 4 * r][lc] = acc[r];
    wave_sync();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ya[b][q] = ts[lc][4 * q + lr];
      const int c = b * 16 + 4 * q + lr;
      if (row_ok && c < k) Y[row + int64_t(c) * ldy] = ya[b][q];
    }
    wave_sync();
  }
}

static int chol_nblk(int64_t k) {
  if (k < 1 || k > kMaxCholK) return 0;
  for (int nb : {2, 4, 8, 12, 16, 17, 18})
    if (k <= 16 * nb) return nb;
  return 0;
}

bool cholqr_small_k(int64_t k) { return chol_nblk(k) != 0; }

// R (column-major, ldr; may alias G), Rt (row-major, ldt) and Dinv (16 x 16
// row-major inverse of each diagonal block, 256 * ceil(k/16) doubles).
int chol_blocks(rla_ctx* ctx, int64_t k, const double* G, int64_t ldg, double* R, int64_t ldr,
                double* Rt, int64_t ldt, double* Dinv, int* info_dev) {
  const int nb = chol_nblk(k);
  if (!nb) return fail(RLA_ENOTSUP, "chol_blocks: k must be in 1..288");
  const int ki = static_cast<int>(k);
#define RLA_CHOL(NB)                                                                       \
  case NB:                                                                                 \
    chol_blocks_kernel<NB><<<1, kCholThreads, 0, ctx->stream>>>(                           \
        ki, G, static_cast<int>(ldg), R, static_cast<int>(ldr), Rt, static_cast<int>(ldt), \
        Dinv, info_dev);                                                                   \
    break;
  switch (nb) {
    RLA_CHOL(2) RLA_CHOL(4) RLA_CHOL(8) RLA_CHOL(12) RLA_CHOL(16) RLA_CHOL(17) RLA_CHOL(18)
  }
#undef RLA_CHOL
  RLA_LAUNCH_CHECK("chol_blocks");
  return RLA_OK;
}

int trsm_ru_f64(rla_ctx* ctx, int64_t m, int64_t k, double* Y, int64_t ldy, const double* Rt,
                int64_t ldt, const double* Dinv, const int* info) {
  const int nb = chol_nblk(k);
  if (!nb) return fail(RLA_ENOTSUP, "trsm_ru_f64: k must be in 1..288");
  if (m <= 0) return RLA_OK;
  const int ki = static_cast<int>(k);
  const unsigned grid = static_cast<unsigned>(ceil_div(m, 64));
#define RLA_TRSM(NB)                                                                         \
  case NB:                                                                                   \
    trsm_ru_f64_kernel<NB><<<dim3(grid), 256, 0, ctx->stream>>>(m, ki, Y, ldy, Rt,           \
                                                                static_cast<int>(ldt), Dinv, \
                                                                info);                       \
    break;
  switch (nb) {
    RLA_TRSM(2) RLA_TRSM(4) RLA_TRSM(8) RLA_TRSM(12) RLA_TRSM(16) RLA_TRSM(17) RLA_TRSM(18)
  }
#undef RLA_TRSM
  RLA_LAUNCH_CHECK("trsm_ru_f64");
  return RLA_OK;
}

__global__ void identity_kernel(int k, double* X, int64_t ldx) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= int64_t(k) * k) return;
  const int i = static_cast<int>(e % k), j = static_cast<int>(e / k);
  X[i + j * ldx] = i == j ? 1.0 : 0.0;
}

}  // namespace rla

using namespace rla;

extern "C" int rla_chol_inv_f64(rla_ctx* ctx, int64_t k, const double* G, int64_t ldg, double* R,
                                int64_t ldr, double* X, int64_t ldx, int* info_host) {
  RLA_CHECK_CTX(ctx);
  if (k < 1 || ldg < k || ldr < k || (X && ldx < k))
    return fail(RLA_EARG, "rla_chol_inv_f64: bad sizes");
  if (k > kMaxCholK) return fail(RLA_ENOTSUP, "rla_chol_inv_f64: k above 288");
  int rc = use_device(ctx);
  if (rc) return rc;
  const int64_t kp = ceil_div(k, 16) * 16;
  void* ws = nullptr;
  const size_t rt_bytes = sizeof(double) * k * k, dinv_bytes = sizeof(double) * kp * 16;
  rc = workspace(ctx, 256 + rt_bytes + dinv_bytes, &ws);
  if (rc) return rc;
  int* info = static_cast<int*>(ws);
  double* Rt = reinterpret_cast<double*>(static_cast<char*>(ws) + 256);
  double* Dinv = reinterpret_cast<double*>(static_cast<char*>(ws) + 256 + rt_bytes);
  rc = chol_blocks(ctx, k, G, ldg, R, ldr, Rt, k, Dinv, info);
  if (rc) return rc;
  if (X) {
    // X = I R^-1 (skipped on the device after a breakdown)
    identity_kernel<<<dim3(unsigned(ceil_div(k * k, 256))), 256, 0, ctx->stream>>>(
        static_cast<int>(k), X, ldx);
    RLA_LAUNCH_CHECK("identity");
    rc = trsm_ru_f64(ctx, k, k, X, ldx, Rt, k, Dinv, info);
    if (rc) return rc;
  }
  int h = 0;
  RLA_HIP(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (info_host) *info_host = h;
  return RLA_OK;
}
