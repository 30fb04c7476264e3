// FP64 MFMA GEMM for the tall-skinny products of CholeskyQR2:
//   C = alpha * op(A) * op(B) + beta * C,  op(X) = X or X' (column-major).
// Used for the Gram matrix G = Y'Y (m x k, m >> k: split over K) and for
// Q = Y * R^-1 (k x k right factor).
//
// Tile 64 x 64 per workgroup (4 wavefronts, 32 x 32 each = 2 x 2 blocks of
// v_mfma_f64_16x16x4f64), K staged through LDS 16 at a time.  Split-K writes
// partial tiles that a second kernel sums in split order (deterministic).
#include <algorithm>

#include "rla_internal.hpp"

namespace rla {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kT = 64;    // tile rows / cols
constexpr int kKS = 16;   // K per LDS stage

// element (i, l) of op(X) for a column-major X with leading dimension ld
__device__ __forceinline__ double opget(const double* __restrict__ X, int64_t ld, bool trans,
                                        int64_t i, int64_t l) {
  return trans ? X[l + i * ld] : X[i + l * ld];
}

__global__ __launch_bounds__(256) void gemm_f64_kernel(int64_t M, int64_t N, int64_t K,
                                                       double alpha, const double* __restrict__ A,
                                                       int64_t lda, int ta,
                                                       const double* __restrict__ B, int64_t ldb,
                                                       int tb, double beta, double* __restrict__ C,
                                                       int64_t ldc, int ksplit, int64_t kps,
                                                       double* __restrict__ part) {
  __shared__ double As[kKS][kT + 1];  // As[l][i]
  __shared__ double Bs[kKS][kT + 1];  // Bs[l][j]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t tiles_m = (M + kT - 1) / kT;
  const int split = blockIdx.x % ksplit;
  const int64_t tile = blockIdx.x / ksplit;
  const int64_t i0 = (tile % tiles_m) * kT, j0 = (tile / tiles_m) * kT;
  const int64_t kb = split * kps, ke = std::min<int64_t>(K, kb + kps);
  const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;

  f64x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};

  for (int64_t k0 = kb; k0 < ke; k0 += kKS) {
    __syncthreads();
    // stage 64 x 16 of op(A) and 16 x 64 of op(B); the unit-stride index of
    // each operand runs along consecutive threads
    for (int e = threadIdx.x; e < kT * kKS; e += 256) {
      int ii, ll;
      if (!ta) { ii = e % kT; ll = e / kT; } else { ll = e % kKS; ii = e / kKS; }
      const int64_t gi = i0 + ii, gl = k0 + ll;
      As[ll][ii] = (gi < M && gl < ke) ? opget(A, lda, ta, gi, gl) : 0.0;
      int jj, l2;
      if (tb) { jj = e % kT; l2 = e / kT; } else { l2 = e % kKS; jj = e / kKS; }
      const int64_t gj = j0 + jj, gl2 = k0 + l2;
      Bs[l2][jj] = (gj < N && gl2 < ke) ? opget(B, ldb, tb, gl2, gj) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kKS; kk += 4) {
      const int l = kk + (lane >> 4);
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const double av = As[l][wi + a * 16 + (lane & 15)];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const double bv = Bs[l][wj + b * 16 + (lane & 15)];
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[a][b], 0, 0, 0);
        }
      }
    }
  }
  // acc[a][b][r]: row wi + a*16 + lane/16 + 4*r, column wj + b*16 + lane%16
  // (the f64 16x16x4 accumulator layout, measured)
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wi + a * 16 + (lane >> 4) + 4 * r;
        const int64_t j = j0 + wj + b * 16 + (lane & 15);
        if (i >= M || j >= N) continue;
        if (ksplit == 1) {
          double* cp = C + i + j * ldc;
          *cp = beta == 0.0 ? alpha * acc[a][b][r] : alpha * acc[a][b][r] + beta * (*cp);
        } else {
          part[(int64_t(split) * N + j) * M + i] = acc[a][b][r];
        }
      }
}

__global__ void gemm_f64_reduce_kernel(int64_t M, int64_t N, int ksplit,
                                       const double* __restrict__ part, double alpha,
                                       double beta, double* __restrict__ C, int64_t ldc) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= M * N) return;
  const int64_t i = e % M, j = e / M;
  double s = 0.0;
  for (int q = 0; q < ksplit; ++q) s += part[(int64_t(q) * N + j) * M + i];
  double* cp = C + i + j * ldc;
  *cp = beta == 0.0 ? alpha * s : alpha * s + beta * (*cp);
}

// part: caller scratch of ksplit*M*N doubles when the call splits K
// (gemm_f64_split_bytes tells how much).
int gemm_f64(rla_ctx* ctx, bool ta, bool tb, int64_t M, int64_t N, int64_t K, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
             int64_t ldc, double* part, size_t part_bytes) {
  if (M <= 0 || N <= 0) return RLA_OK;
  const int64_t tiles = ceil_div(M, kT) * ceil_div(N, kT);
  int ksplit = 1;
  while (ksplit < 16 && tiles * ksplit * 2 <= 512 && K / (ksplit * 2) >= 512) ksplit *= 2;
  if (ksplit > 1 && (!part || part_bytes < size_t(ksplit) * M * N * sizeof(double))) {
    ksplit = static_cast<int>(std::min<int64_t>(
        ksplit, part ? int64_t(part_bytes / (sizeof(double) * M * N)) : 1));
    if (ksplit < 1) ksplit = 1;
  }
  const int64_t kps = ceil_div(ceil_div(K, ksplit), kKS) * kKS;
  ksplit = static_cast<int>(std::max<int64_t>(1, ceil_div(K, kps)));
  gemm_f64_kernel<<<dim3(static_cast<unsigned>(tiles * ksplit)), 256, 0, ctx->stream>>>(
      M, N, K, alpha, A, lda, ta ? 1 : 0, B, ldb, tb ? 1 : 0, beta, C, ldc, ksplit, kps, part);
  RLA_LAUNCH_CHECK("gemm_f64");
  if (ksplit > 1) {
    gemm_f64_reduce_kernel<<<dim3(static_cast<unsigned>(ceil_div(M * N, 256))), 256, 0,
                             ctx->stream>>>(M, N, ksplit, part, alpha, beta, C, ldc);
    RLA_LAUNCH_CHECK("gemm_f64/reduce");
  }
  return RLA_OK;
}

size_t gemm_f64_split_bytes(int64_t M, int64_t N) { return size_t(16) * M * N * sizeof(double); }

// ---- CholeskyQR kernels for k <= 288 columns ---------------------------------
//
// Gram G = Y'Y: each workgroup takes a contiguous range of rows of Y and
// computes every upper-triangular 16 x 16 block of G for it, so Y is read
// from HBM exactly once.  Y is FP64, or FP32 (widened exactly on load, so
// the result equals the Gram matrix of the FP64 copy).  Rows are staged 16 at a time in LDS (double
// buffered, next chunk prefetched into registers); the 8 wavefronts share the
// NBLK(NBLK+1)/2 blocks round-robin.  Per-workgroup partial blocks go to
// scratch and a second kernel sums them in workgroup order (deterministic).

constexpr int kGramRows = 16;  // rows per LDS chunk

// 8 wavefronts hold up to 16 x 17 / 2 blocks in 256 registers each; the
// 18 x 19 / 2 blocks of k <= 288 take 16 wavefronts (128 registers each)
constexpr int gram_threads(int nblk) { return nblk > 16 ? 1024 : 512; }

__device__ __forceinline__ void tri_block(int q, int& bi, int& bj) {
  bj = 0;
  while (q > bj) {
    q -= bj + 1;
    ++bj;
  }
  bi = q;
}

template <int NBLK, class TY>
__global__ __launch_bounds__(gram_threads(NBLK)) void gram_f64_kernel(int64_t m, int k,
                                                                const TY* __restrict__ Y,
                                                                int64_t ldy, int64_t rows_per_wg,
                                                                double* __restrict__ part) {
  constexpr int CP = NBLK * 16;
  constexpr int LDSS = (CP + 31) / 32 * 32 + 16;  // rows 2 apart map to opposite bank halves
  constexpr int T = NBLK * (NBLK + 1) / 2;
  constexpr int NT = gram_threads(NBLK);
  constexpr int NW = NT / 64;
  constexpr int MAXB = (T + NW - 1) / NW;
  constexpr int CSTEP = NT / kGramRows;               // chunk columns per load pass
  constexpr int PER = (CP + CSTEP - 1) / CSTEP;
  __shared__ double Ys[2][kGramRows * LDSS];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t r_begin = int64_t(blockIdx.x) * rows_per_wg;
  const int64_t r_end = std::min<int64_t>(m, r_begin + rows_per_wg);
  int offa[MAXB], offb[MAXB];
#pragma unroll
  for (int s = 0; s < MAXB; ++s) {
    const int q = wave + NW * s;
    int bi = 0, bj = 0;
    tri_block(q, bi, bj);
    offa[s] = q < T ? bi * 16 : -1;
    offb[s] = bj * 16;
  }
  f64x4 acc[MAXB];
#pragma unroll
  for (int s = 0; s < MAXB; ++s) acc[s] = f64x4{0.0, 0.0, 0.0, 0.0};

  const int lr = t & 15, lc = t >> 4;  // chunk element (row lr, column lc + CSTEP p)
  double v[PER];
  // loads are unconditional from a clamped (in-bounds) address; out-of-range
  // entries are zeroed afterwards, so the prefetch needs no branches
  auto load = [&](int64_t r0) {
    const int64_t gr = r0 + lr;
    const bool row_ok = gr < r_end;
    const TY* src = Y + (row_ok ? gr : r_begin);
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int c = lc + CSTEP * p;
      const double x = static_cast<double>(src[int64_t(c < k ? c : k - 1) * ldy]);
      v[p] = (row_ok && c < k) ? x : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int c = lc + CSTEP * p;
      if (c < CP) Ys[buf][lr * LDSS + c] = v[p];
    }
  };
  if (r_begin < r_end) {
    load(r_begin);
    store(0);
  }
  __syncthreads();
  int buf = 0;
  for (int64_t r0 = r_begin; r0 < r_end; r0 += kGramRows) {
    const bool more = r0 + kGramRows < r_end;
    if (more) load(r0 + kGramRows);
#pragma unroll
    for (int kk = 0; kk < kGramRows; kk += 4) {
      const double* base = &Ys[buf][(kk + (lane >> 4)) * LDSS + (lane & 15)];
#pragma unroll
      for (int s = 0; s < MAXB; ++s) {
        if (offa[s] >= 0)
          acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(base[offa[s]], base[offb[s]], acc[s], 0,
                                                        0, 0);
      }
    }
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // acc[s][r] is block element (lane/16 + 4r, lane%16), i.e. index lane + 64r
#pragma unroll
  for (int s = 0; s < MAXB; ++s) {
    if (offa[s] < 0) continue;
    double* dst = part + (int64_t(blockIdx.x) * T + wave + NW * s) * 256 + lane;
#pragma unroll
    for (int r = 0; r < 4; ++r) dst[64 * r] = acc[s][r];
  }
}

__global__ __launch_bounds__(256) void gram_reduce_kernel(int k, int T, int nparts,
                                                          const double* __restrict__ part,
                                                          double* __restrict__ G, int64_t ldg) {
  const int q = blockIdx.x, e = threadIdx.x;
  double s = 0.0;
  const double* p = part + int64_t(q) * 256 + e;
#pragma unroll 8
  for (int w = 0; w < nparts; ++w) s += p[int64_t(w) * T * 256];
  int bi, bj;
  tri_block(q, bi, bj);
  const int li = e >> 4, lj = e & 15;
  if (bi == bj && li > lj) return;  // the mirror element writes both
  const int i = bi * 16 + li, j = bj * 16 + lj;
  if (i >= k || j >= k) return;
  G[i + int64_t(j) * ldg] = s;
  G[j + int64_t(i) * ldg] = s;
}

// smallest instantiated block count covering k columns (0: k too large)
static int qr_nblk(int64_t k) {
  for (int nb : {4, 8, 12, 16, 17, 18})
    if (k <= 16 * nb) return nb;
  return 0;
}

static int64_t gram_parts(int64_t m) {
  return std::max<int64_t>(1, std::min<int64_t>(256, ceil_div(m, 128)));
}

size_t gram_f64_part_bytes(int64_t m, int64_t k) {
  const int nb = qr_nblk(k);
  if (!nb) return 0;
  return size_t(gram_parts(m)) * (nb * (nb + 1) / 2) * 256 * sizeof(double);
}

template <class TY>
int gram_f64(rla_ctx* ctx, int64_t m, int64_t k, const TY* Y, int64_t ldy, double* G,
             int64_t ldg, double* part, size_t part_bytes) {
  const int nb = qr_nblk(k);
  if (!nb) return fail(RLA_ENOTSUP, "gram_f64: more than 288 columns");
  if (part_bytes < gram_f64_part_bytes(m, k)) return fail(RLA_EARG, "gram_f64: scratch too small");
  const int64_t P = gram_parts(m);
  const int64_t rows = ceil_div(ceil_div(m, P), kGramRows) * kGramRows;
  const int64_t nparts = ceil_div(m, rows);
  const int T = nb * (nb + 1) / 2;
  const int ki = static_cast<int>(k);
#define RLA_GRAM(NB)                                                                          \
  case NB:                                                                                    \
    gram_f64_kernel<NB, TY><<<dim3(unsigned(nparts)), gram_threads(NB), 0, ctx->stream>>>( \
        m, ki, Y, ldy, rows, part);                                                          \
    break;
  switch (nb) {
    RLA_GRAM(4) RLA_GRAM(8) RLA_GRAM(12) RLA_GRAM(16) RLA_GRAM(17) RLA_GRAM(18)
  }
#undef RLA_GRAM
  RLA_LAUNCH_CHECK("gram_f64");
  gram_reduce_kernel<<<dim3(unsigned(T)), 256, 0, ctx->stream>>>(ki, T, static_cast<int>(nparts),
                                                                 part, G, ldg);
  RLA_LAUNCH_CHECK("gram_f64/reduce");
  return RLA_OK;
}

template int gram_f64<double>(rla_ctx*, int64_t, int64_t, const double*, int64_t, double*,
                              int64_t, double*, size_t);
template int gram_f64<float>(rla_ctx*, int64_t, int64_t, const float*, int64_t, double*,
                             int64_t, double*, size_t);

}  // namespace rla

using namespace rla;

extern "C" int rla_gemm_f64_mfma(rla_ctx* ctx, int trans_a, int trans_b, int64_t m, int64_t n,
                                 int64_t k, double alpha, const double* A, int64_t lda,
                                 const double* B, int64_t ldb, double beta, double* C,
                                 int64_t ldc) {
  RLA_CHECK_CTX(ctx);
  int rc = use_device(ctx);
  if (rc) return rc;
  void* ws = nullptr;
  const size_t pb = gemm_f64_split_bytes(m, n);
  rc = workspace(ctx, pb, &ws);
  if (rc) return rc;
  return gemm_f64(ctx, trans_a != 0, trans_b != 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                  static_cast<double*>(ws), pb);
}
