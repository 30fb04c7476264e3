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
// N <= 288 (the RangeFinder's k columns) runs the pipelined kernel below
// (gemm_bf16_pipe_kernel: all N columns in one workgroup tile, the wavefronts
// split into two column groups, double-buffered B tile, A prefetched one step
// ahead).  Wider N runs gemm_bf16_tn_kernel: workgroup tile 128 rows (4
// wavefronts x 32 rows) x NB*32 columns; per 64-wide K step the B tile (NB*32
// columns x 64) is staged in LDS (padded row stride, conflict-free 16-byte
// fragment reads) and every wavefront runs 4 x NB v_mfma_f32_32x32x16_bf16 on
// its 32 rows.  Both split K when the tiles alone do not fill the CUs; the
// partial tiles are summed in split order by a second kernel, so results are
// deterministic.
#include <algorithm>
#include <mutex>

#include "rla_internal.hpp"

namespace rla {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128;        // rows per workgroup
constexpr int kBK = 64;         // K per step
constexpr int kBstride = 72;    // bf16 per staged B column (64 + 8 pad)

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

// The tiled BF16 copy of a (rows x cols) operand that the pipelined kernel
// streams as its A: tiles of 128 rows x 64 columns (K), tile (t, s) at
// element (t * ceil(cols / 64) + s) * 8192; inside a tile, the 16-byte piece
// of row 32 g + r (r < 32) and columns 16 q + 8 h .. + 8 (h < 2) sits at
// element ((4 g + q) * 64 + r + 32 h) * 8, i.e. in the order one wavefront's
// operand load reads it, so every load is one contiguous kilobyte.  Entries
// past the operand are zero.  The source is read through LDS as in
// to_bf16_rows_kernel; the same rounding (float, then nearest-even bf16).
constexpr int kTileRows = 128;
constexpr int kTileK = 64;
constexpr int kTileElems = kTileRows * kTileK;

__device__ __forceinline__ int64_t tiled_index(int64_t i, int64_t l, int64_t nkt) {
  const int64_t t = i / kTileRows, s = l / kTileK;
  const int r = static_cast<int>(i % kTileRows), c = static_cast<int>(l % kTileK);
  const int g = r >> 5, q = c >> 4, h = (c >> 3) & 1;
  return (t * nkt + s) * kTileElems + ((4 * g + q) * 64 + (r & 31) + 32 * h) * 8 + (c & 7);
}

template <class T>
__global__ __launch_bounds__(256) void to_bf16_tiled_kernel(int64_t rows, int64_t cols,
                                                            const T* __restrict__ src,
                                                            int64_t s_rs, int64_t s_cs,
                                                            uint16_t* __restrict__ dst,
                                                            int64_t nkt) {
  __shared__ float tile[64][65];
  const int64_t i0 = int64_t(blockIdx.y) * 64, l0 = int64_t(blockIdx.x) * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const bool rows_fast = s_rs == 1;
  for (int r = ty; r < 64; r += 4) {
    const int64_t i = rows_fast ? i0 + tx : i0 + r;
    const int64_t l = rows_fast ? l0 + r : l0 + tx;
    float v = 0.f;
    if (i < rows && l < cols) v = static_cast<float>(src[i * s_rs + l * s_cs]);
    if (rows_fast) tile[tx][r] = v; else tile[r][tx] = v;
  }
  __syncthreads();
  // the grid covers the padded extent, so padding is written as zeros
  for (int r = ty; r < 64; r += 4) {
    const int64_t i = i0 + r, l = l0 + tx;
    dst[tiled_index(i, l, nkt)] = f32_to_bf16_bits(tile[r][tx]);
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

// ---- pipelined kernel for N <= 288 -------------------------------------------
//
// Workgroup tile 128 rows x BN = 32*(NB0+NB1) columns (all of N).  The
// wavefronts form a (4/RT) x 2 grid: wavefront (wr, wc) computes rows
// wr*32*RT .. +32*RT (RT 32-row MFMA tiles) times column tiles
// [wc ? NB0 : 0, +NB0 or NB1), so each B fragment read from LDS feeds RT
// MFMAs.  Per 64-wide K step:
//   1. the A fragments of step s+1 are issued into the other half of a
//      two-deep register ring (the loop is unrolled by two so both ring slots
//      are static);
//   2. the step's 4 x NB x RT MFMAs run from the ring slot of step s and the
//      LDS buffer of step s, the B fragments of substep kk+1 read while
//      substep kk multiplies (a scheduling barrier keeps the reads ahead);
//   3. the B tile of step s+1, loaded into registers one step earlier, goes to
//      the other LDS buffer;
//   4. the B tile of step s+2 is issued; one barrier.
// Rows past M and columns past N read clamped (valid) addresses and are never
// stored; only a final partial K step zero-fills, in both operands.
constexpr int kPM = 128;
constexpr int kPipeRT = 1;  // row tiles per wavefront (1: 8 wavefronts, 2 per SIMD; 2: 4)

template <int NB0, int NB1, int RT>
struct PipeCfg {
  static constexpr int NT = 256 * (2 / RT);  // threads: 4 or 8 wavefronts
  static constexpr int WROWS = 4 / RT;       // wavefront rows of the grid
  static constexpr int BN = (NB0 + NB1) * 32;
  static constexpr int CHUNKS = BN * 8;  // 16-byte pieces of one K step of the B tile
  static constexpr int PER_T = (CHUNKS + NT - 1) / NT;
  static constexpr int BUF = BN * kBstride;  // bf16 per LDS buffer
};

template <int NB, int NB0, int NB1, int RT, bool TILED>
__device__ __forceinline__ void pipe_run(int64_t M, int64_t N, const uint16_t* __restrict__ A,
                                         int64_t lda, const uint16_t* __restrict__ B,
                                         int64_t ldb, float* __restrict__ C, int64_t c_rs,
                                         int64_t c_cs, float alpha, float beta, int ksplit,
                                         int split, int64_t kb, int64_t ke, int64_t tm, int wr,
                                         int ct0, float* __restrict__ part, uint16_t* Bs) {
  using P = PipeCfg<NB0, NB1, RT>;
  const int tid = threadIdx.x, lane = tid & 63;
  const int akq = (lane >> 5) * 8;

  f32x16 acc[RT][NB];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[rt][nb][r] = 0.f;

  const int64_t row_w = tm * kPM + wr * 32 * RT;  // first row of this wavefront
  // row layout: row pointers (lda the row stride); tiled layout (lda = the
  // number of 64-wide K tiles): this lane's piece of tile row tm, K tile 0
  const uint16_t* ap[RT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
    ap[rt] = TILED ? A + tm * lda * kTileElems + ((wr * RT + rt) * 4 * 64 + lane) * 8
                   : A + std::min<int64_t>(row_w + rt * 32 + (lane & 31), M - 1) * lda;
  auto a_piece = [&](int rt, int64_t k0, int kk) {
    return TILED ? ap[rt] + (k0 / kTileK) * kTileElems + kk * 64 * 8 : ap[rt] + k0 + kk * 16 + akq;
  };
  // piece u of this thread: column (tid >> 3) + u * NT/8 (clamped to N-1 in
  // global memory), K offset (tid & 7) * 8; LDS offset boff0 + u * NT/8 * kBstride
  const int bcol0 = tid >> 3;
  const uint16_t* bbase = B + (tid & 7) * 8;
  const int boff0 = bcol0 * kBstride + (tid & 7) * 8;
  auto bsrc = [&](int u) {
    return bbase + std::min<int64_t>(bcol0 + (P::NT / 8) * u, N - 1) * ldb;
  };
  auto has_piece = [&](int u) { return u + 1 < P::PER_T || tid + u * P::NT < P::CHUNKS; };

  const int64_t nfull = (ke - kb) / kBK;
  auto step_k = [&](int64_t s) { return kb + std::min<int64_t>(s, nfull - 1) * kBK; };
  auto load_a = [&](int64_t k0, bf16x8(&af)[RT][4]) {
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
        af[rt][kk] = __builtin_bit_cast(bf16x8,
                                        *reinterpret_cast<const u32x4*>(a_piece(rt, k0, kk)));
  };
  auto load_b = [&](int64_t k0, u32x4(&br)[P::PER_T]) {
#pragma unroll
    for (int u = 0; u < P::PER_T; ++u)
      if (has_piece(u)) br[u] = *reinterpret_cast<const u32x4*>(bsrc(u) + k0);
  };
  auto store_b = [&](uint16_t* buf, const u32x4(&br)[P::PER_T]) {
#pragma unroll
    for (int u = 0; u < P::PER_T; ++u)
      if (has_piece(u))
        *reinterpret_cast<u32x4*>(buf + boff0 + u * (P::NT / 8) * kBstride) = br[u];
  };
  const int fbase = (ct0 * 32 + (lane & 31)) * kBstride + akq;
  auto read_frags = [&](const uint16_t* buf, int kk, bf16x8(&bv)[NB]) {
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
      bv[nb] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(
                                              buf + fbase + nb * 32 * kBstride + kk * 16));
  };
  auto compute = [&](const bf16x8(&af)[RT][4], const uint16_t* buf) {
    bf16x8 bv[2][NB];
    read_frags(buf, 0, bv[0]);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      // keep the next substep's LDS reads ahead of this substep's MFMAs (the
      // scheduler would otherwise sink them and serialise on LDS latency)
      if (kk + 1 < 4) read_frags(buf, kk + 1, bv[(kk + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
          acc[rt][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[rt][kk], bv[kk & 1][nb],
                                                                acc[rt][nb], 0, 0, 0);
    }
  };

  if (nfull > 0) {
    bf16x8 r0[RT][4], r1[RT][4];
    u32x4 br[P::PER_T];
    load_b(step_k(0), br);
    load_a(step_k(0), r0);
    store_b(Bs, br);
    load_b(step_k(1), br);
    __syncthreads();
    // step s: prefetch A(s+1) into `nxt`, multiply with `cur`, B(s+1) to LDS,
    // prefetch B(s+2)
    auto step = [&](int64_t s, const bf16x8(&cur)[RT][4], bf16x8(&nxt)[RT][4]) {
      load_a(step_k(s + 1), nxt);
      compute(cur, Bs + (s & 1) * P::BUF);
      store_b(Bs + ((s + 1) & 1) * P::BUF, br);
      load_b(step_k(s + 2), br);
      __syncthreads();
    };
    for (int64_t s = 0; s < nfull; s += 2) {
      step(s, r0, r1);
      if (s + 1 >= nfull) break;
      step(s + 1, r1, r0);
    }
  }
  if (kb + nfull * kBK < ke) {
    // partial last K step: guarded loads, zero past ke in both operands
    const int64_t k0 = kb + nfull * kBK;
    bf16x8 af[RT][4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int64_t gk = k0 + kk * 16 + akq;
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        u32x4 v = {0u, 0u, 0u, 0u};
        if (TILED) v = *reinterpret_cast<const u32x4*>(a_piece(rt, k0, kk));  // zero-padded
        else if (gk + 8 <= ke) v = *reinterpret_cast<const u32x4*>(ap[rt] + gk);
        else if (gk < ke) v = load8_tail(ap[rt] + gk, ke - gk);
        af[rt][kk] = __builtin_bit_cast(bf16x8, v);
      }
    }
    u32x4 br[P::PER_T];
#pragma unroll
    for (int u = 0; u < P::PER_T; ++u) {
      const int64_t gk = k0 + (tid & 7) * 8;
      u32x4 v = {0u, 0u, 0u, 0u};
      if (gk + 8 <= ke) v = *reinterpret_cast<const u32x4*>(bsrc(u) + k0);
      else if (gk < ke) v = load8_tail(bsrc(u) + k0, ke - gk);
      br[u] = v;
    }
    uint16_t* buf = Bs + (nfull & 1) * P::BUF;
    store_b(buf, br);
    __syncthreads();
    compute(af, buf);
  }

  // acc[rt][nb][r]: row row_w + rt*32 + (r/4)*8 + (lane/32)*4 + r%4,
  // column (ct0 + nb)*32 + lane%32
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int64_t j = int64_t(ct0 + nb) * 32 + (lane & 31);
    if (j >= N) continue;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t i = row_w + rt * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (i >= M) continue;
        if (ksplit == 1) {
          float* cp = C + i * c_rs + j * c_cs;
          *cp = beta == 0.f ? alpha * acc[rt][nb][r] : alpha * acc[rt][nb][r] + beta * (*cp);
        } else {
          part[(int64_t(split) * N + j) * M + i] = acc[rt][nb][r];
        }
      }
    }
  }
}

template <int NB0, int NB1, int RT, bool TILED>
__global__ __launch_bounds__(256 * (2 / RT))
__attribute__((amdgpu_waves_per_eu(2 / RT, 2 / RT)))
void gemm_bf16_pipe_kernel(int64_t M, int64_t N, int64_t K, const uint16_t* __restrict__ A,
                           int64_t lda, const uint16_t* __restrict__ B, int64_t ldb,
                           float* __restrict__ C, int64_t c_rs, int64_t c_cs, float alpha,
                           float beta, int ksplit, int64_t k_per_split,
                           float* __restrict__ part) {
  using P = PipeCfg<NB0, NB1, RT>;
  extern __shared__ __align__(16) uint16_t pipe_lds[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int split = blockIdx.x % ksplit;
  const int64_t tm = blockIdx.x / ksplit;
  const int64_t kb = split * k_per_split;
  const int64_t ke = std::min<int64_t>(K, kb + k_per_split);
  if (kb >= ke) return;  // uniform over the workgroup
  const int wr = wave % P::WROWS;
  if (wave / P::WROWS)
    pipe_run<NB1, NB0, NB1, RT, TILED>(M, N, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta, ksplit, split,
                                kb, ke, tm, wr, NB0, part, pipe_lds);
  else
    pipe_run<NB0, NB0, NB1, RT, TILED>(M, N, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta, ksplit, split,
                                kb, ke, tm, wr, 0, part, pipe_lds);
}

template <int NB0, int NB1, int RT, bool TILED = false>
static int launch_pipe(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, const uint16_t* A,
                       int64_t lda, const uint16_t* B, int64_t ldb, float* C, int64_t c_rs,
                       int64_t c_cs, float alpha, float beta) {
  using P = PipeCfg<NB0, NB1, RT>;
  const size_t lds = size_t(2) * P::BUF * sizeof(uint16_t);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&gemm_bf16_pipe_kernel<NB0, NB1, RT, TILED>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  RLA_HIP(attr_err);
  const int64_t tiles = ceil_div(M, kPM);
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
  gemm_bf16_pipe_kernel<NB0, NB1, RT, TILED><<<dim3(static_cast<unsigned>(tiles * ksplit)), P::NT, lds,
                                        ctx->stream>>>(M, N, K, A, lda, B, ldb, C, c_rs, c_cs,
                                                       alpha, beta, ksplit, kps, part);
  RLA_LAUNCH_CHECK("gemm_bf16_pipe");
  if (ksplit > 1) {
    splitk_reduce_kernel<<<dim3(static_cast<unsigned>(ceil_div(M * N, 256))), 256, 0,
                           ctx->stream>>>(M, N, ksplit, part, C, c_rs, c_cs, alpha, beta);
    RLA_LAUNCH_CHECK("gemm_bf16_pipe/splitk");
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
  // N <= 288: the pipelined 2 x 2 kernel, column tiles split between the two
  // wavefront columns
  const int64_t nbk = ceil_div(N, 32);
  switch (nbk) {
    case 2: return launch_pipe<1, 1, kPipeRT>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
    case 3: return launch_pipe<2, 1, kPipeRT>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
    case 4: return launch_pipe<2, 2, kPipeRT>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
    case 5: return launch_pipe<3, 2, kPipeRT>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
    case 6: return launch_pipe<3, 3, kPipeRT>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
    case 7: return launch_pipe<4, 3, kPipeRT>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
    case 8: return launch_pipe<4, 4, kPipeRT>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
    case 9: return launch_pipe<5, 4, kPipeRT>(ctx, M, N, K, A, lda, B, ldb, C, c_rs, c_cs, alpha, beta);
    default: break;
  }
  // otherwise the widest single-wavefront-column tile (256 columns per pass)
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

int64_t bf16_tiled_elems(int64_t rows, int64_t cols) {
  return ceil_div(rows, kTileRows) * ceil_div(cols, kTileK) * kTileElems;
}

template <class T>
int to_bf16_tiled(rla_ctx* ctx, int64_t rows, int64_t cols, const T* src, int64_t s_rs,
                  int64_t s_cs, uint16_t* dst) {
  if (rows <= 0 || cols <= 0) return RLA_OK;
  const int64_t nkt = ceil_div(cols, kTileK);
  dim3 grid(static_cast<unsigned>(nkt), static_cast<unsigned>(ceil_div(rows, kTileRows) * 2));
  to_bf16_tiled_kernel<T><<<grid, 256, 0, ctx->stream>>>(rows, cols, src, s_rs, s_cs, dst, nkt);
  RLA_LAUNCH_CHECK("to_bf16_tiled");
  return RLA_OK;
}

template int to_bf16_tiled<float>(rla_ctx*, int64_t, int64_t, const float*, int64_t, int64_t,
                                  uint16_t*);
template int to_bf16_tiled<double>(rla_ctx*, int64_t, int64_t, const double*, int64_t, int64_t,
                                   uint16_t*);

// gemm_bf16_tn with A in the tiled layout of to_bf16_tiled (M x K, N <= 288)
int gemm_bf16_tn_tiled(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, float alpha,
                       const uint16_t* At, const uint16_t* B, int64_t ldb, float beta, float* C,
                       int64_t c_rs, int64_t c_cs) {
  if (M <= 0 || N <= 0) return RLA_OK;
  if (K <= 0) return fail(RLA_EARG, "gemm_bf16: K must be positive");
  if ((ldb % 8) || (reinterpret_cast<uintptr_t>(At) % 16) || (reinterpret_cast<uintptr_t>(B) % 16))
    return fail(RLA_EARG, "gemm_bf16: operands must be 16-byte aligned with ld a multiple of 8");
  const int64_t nkt = ceil_div(K, kTileK);
#define RLA_TPIPE(A0, A1)                                                                  \
  return launch_pipe<A0, A1, kPipeRT, true>(ctx, M, N, K, At, nkt, B, ldb, C, c_rs, c_cs, \
                                            alpha, beta)
  switch (ceil_div(N, 32)) {
    case 1:
    case 2: RLA_TPIPE(1, 1);
    case 3: RLA_TPIPE(2, 1);
    case 4: RLA_TPIPE(2, 2);
    case 5: RLA_TPIPE(3, 2);
    case 6: RLA_TPIPE(3, 3);
    case 7: RLA_TPIPE(4, 3);
    case 8: RLA_TPIPE(4, 4);
    case 9: RLA_TPIPE(5, 4);
    default: break;
  }
#undef RLA_TPIPE
  return fail(RLA_ENOTSUP, "gemm_bf16 with a tiled A takes N <= 288");
}

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

int rla_to_bf16_tiled_bytes(int64_t rows, int64_t cols, size_t* bytes) {
  if (!bytes) return fail(RLA_EARG, "bytes is NULL");
  if (rows < 0 || cols < 0) return fail(RLA_EARG, "rla_to_bf16_tiled: negative dimensions");
  *bytes = size_t(bf16_tiled_elems(rows, cols)) * sizeof(uint16_t);
  return RLA_OK;
}

int rla_to_bf16_tiled_f32(rla_ctx* ctx, int64_t rows, int64_t cols, const float* src,
                          int64_t s_rs, int64_t s_cs, uint16_t* dst) {
  RLA_CHECK_CTX(ctx);
  if (rows > 0 && cols > 0 && (!src || !dst)) return fail(RLA_EARG, "rla_to_bf16_tiled: NULL");
  int rc = use_device(ctx);
  if (rc) return rc;
  return to_bf16_tiled<float>(ctx, rows, cols, src, s_rs, s_cs, dst);
}

int rla_to_bf16_tiled_f64(rla_ctx* ctx, int64_t rows, int64_t cols, const double* src,
                          int64_t s_rs, int64_t s_cs, uint16_t* dst) {
  RLA_CHECK_CTX(ctx);
  if (rows > 0 && cols > 0 && (!src || !dst)) return fail(RLA_EARG, "rla_to_bf16_tiled: NULL");
  int rc = use_device(ctx);
  if (rc) return rc;
  return to_bf16_tiled<double>(ctx, rows, cols, src, s_rs, s_cs, dst);
}

int rla_gemm_bf16_tiled(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, float alpha,
                        const uint16_t* At, const uint16_t* B, int64_t ldb, float beta, float* C,
                        int64_t c_rs, int64_t c_cs) {
  RLA_CHECK_CTX(ctx);
  int rc = use_device(ctx);
  if (rc) return rc;
  return gemm_bf16_tn_tiled(ctx, M, N, K, alpha, At, B, ldb, beta, C, c_rs, c_cs);
}

}  // extern "C"
