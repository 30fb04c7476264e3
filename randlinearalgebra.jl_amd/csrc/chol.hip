// Cholesky factor of the small k x k Gram matrix of CholeskyQR (k <= 288),
// G = R'R with R upper and a positive diagonal, and the in-place triangular
// solve Y <- Y R^-1 that completes a CholeskyQR pass.
//
// Factorisation: one workgroup of 8 wavefronts.  The upper-triangular
// 16 x 16 blocks of G are spread round-robin over 7 wavefronts and stay in
// their MFMA accumulator registers for the whole factorisation (at most
// 18 x 19 / 2 blocks, 25 per wavefront), so the k^3/3 flops of the trailing
// updates touch no memory.  Block step s:
//   1. the eighth wavefront factors the diagonal block G_ss (in LDS) in
//      registers and inverts R_ss in the same pass (elimination of
//      [G_ss | I]; four lanes per row, each pivot row broadcast through
//      LDS), while the block wavefronts finish the trailing update of step
//      s-1,
//   2. the owners of G_sj (j > s) form R_sj = inv(R_ss)' G_sj on the MFMA
//      and publish it in LDS,
//   3. the owner of G_{s+1,s+1} subtracts R_{s,s+1}' R_{s,s+1} from it first
//      and copies it to LDS (look-ahead: step s+1's factor starts at once);
//      the other trailing blocks are updated during that factor.
// R is written column-major and row-major (the solve's operand layout), with
// the inverses of the diagonal blocks for the solve.
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

// LDS exchanges between the lanes of one wavefront need no fence: the LDS
// executes a wavefront's instructions in order, and these accesses may alias
// so the compiler keeps their order.  The wave barrier only pins the code.
__device__ __forceinline__ void lane_exchange() { __builtin_amdgcn_wave_barrier(); }

// Workgroup barrier for the factorisation steps: the steps exchange data
// through LDS only, so a wavefront waits for its own LDS operations and not
// for its global stores of R, Rt and Dinv (no wavefront reads those during
// the kernel).  __syncthreads() would also wait for every store to reach
// memory.  The memory clobber keeps the compiler from moving LDS accesses
// across the barrier.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

__device__ __forceinline__ void upper_block(int q, int& bi, int& bj) {
  bj = 0;
  while (q > bj) {
    q -= bj + 1;
    ++bj;
  }
  bi = q;
}

// value of v in lane src (src uniform), for every lane
__device__ __forceinline__ double lane_bcast(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// The upper-triangular 16 x 16 blocks (bi, bj) owned by each of BW block
// wavefronts, column by column.  Owner (bi + 2 bj) mod BW spreads every block
// row (the panel of a step) and every trailing region evenly (for BW = 7 no
// wavefront holds more than ceil(T / 7) blocks or more than 3 blocks of a
// panel row).
template <int NBLK, int BW>
struct BlockOwners {
  int bi[BW][NBLK * (NBLK + 1) / 2];
  int bj[BW][NBLK * (NBLK + 1) / 2];
  int n[BW];
  constexpr BlockOwners() : bi(), bj(), n() {
    for (int j = 0; j < NBLK; ++j)
      for (int i = 0; i <= j; ++i) {
        const int w = (i + 2 * j) % BW;
        bi[w][n[w]] = i;
        bj[w][n[w]] = j;
        ++n[w];
      }
  }
};

// Workgroup shape: 512 threads, seven block wavefronts and the factoring one
// (larger workgroups lower the VGPRs per wavefront and spill more).
template <int NBLK>
struct CholShape {
  static constexpr int threads = 512;
  static constexpr int block_waves = threads / 64 - 1;
};

// The trailing-matrix role of block wavefront W: the blocks BlockOwners
// gives it (BW block wavefronts).  W is a template
// parameter so every block coordinate and LDS offset is a compile-time
// constant (no address or index registers per block).
template <int NBLK, int W>
__device__ __forceinline__ void chol_block_role(int k, global_ptr<const double> G, int ldg,
                                                bool gsym,
                                                global_ptr<double> R, int ldr,
                                                global_ptr<double> __restrict__ Rt, int ldt,
                                                global_ptr<double> __restrict__ Dinv,
                                                double (*P)[NBLK * 16 + 1], double (*Ds)[17],
                                                double (*Dv)[17], lds_ptr<volatile int> fail,
                                                int t) {
  constexpr int BW = CholShape<NBLK>::block_waves;
  constexpr BlockOwners<NBLK, BW> own{};
  constexpr int MAXB = own.n[W];  // blocks of this wavefront
  const int lane = t & 63, lr = lane >> 4, lc = lane & 15;
  f64x4 acc[MAXB > 0 ? MAXB : 1];
  // clamped (in-bounds) loads, all issued before any result is used (one
  // wait); the entries outside G are zeroed afterwards.  A symmetric G is read at the
  // mirror element, so the 16 lanes of a row read 128 contiguous bytes.
#pragma unroll
  for (int u = 0; u < MAXB; ++u) {
    const int bi = own.bi[W][u], bj = own.bj[W][u];
    // issued four blocks at a time, so that the addresses of all the loads
    // in flight are not live at once
    if (u % 4 == 0) asm volatile("" ::: "memory");
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = min(bi * 16 + lr + 4 * r, k - 1), j = min(bj * 16 + lc, k - 1);
      acc[u][r] = gsym ? G[j + i * ldg] : G[i + j * ldg];
    }
  }
  // every G load has returned before B0, so after it R may be written where
  // it aliases G
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#pragma unroll
  for (int u = 0; u < MAXB; ++u) {
    const int bi = own.bi[W][u], bj = own.bj[W][u];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (!(bi * 16 + lr + 4 * r < k && bj * 16 + lc < k)) acc[u][r] = 0.0;
  }
  __syncthreads();  // B0: G read into registers
  // zero R below the block diagonal (the diagonal blocks are written whole)
  if (R) {
    const int wv = t >> 6, ln = t & 63;
    for (int j = wv; j < k; j += BW)
      for (int i = ((j >> 4) + 1) * 16 + ln; i < k; i += 64) R[i + j * ldr] = 0.0;
  }
  double* pl = &P[lr][lc];  // this lane's element of a 4 x 16 slab (rows NBLK*16+1 apart)
  constexpr int SR = NBLK * 16 + 1;
  // Barriers per step s: B1 (G_ss in LDS), B2 (R_ss and its inverse in LDS),
  // B3 (panel row s in LDS).  Look-ahead: after B3 the owner of G_{s+1,s+1}
  // updates that block first and publishes it (the next step's B1), so the
  // factor wavefront works on it while the remaining trailing blocks are
  // updated.
  auto update = [&](int u, int bi, int bj) {
    double a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = pl[4 * q * SR + bi * 16];
      b[q] = pl[4 * q * SR + bj * 16];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[q], b[q], acc[u], 0, 0, 0);
  };
  // G_00 to LDS
#pragma unroll
  for (int u = 0; u < MAXB; ++u)
    if (own.bi[W][u] == 0 && own.bj[W][u] == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) Ds[lr + 4 * r][lc] = acc[u][r];
    }
  for (int s = 0; s < NBLK; ++s) {
    const int js = s * 16;
    if (js >= k) break;
    lds_barrier();  // B1: G_ss in LDS; the factor wavefront starts
    if (s > 0) {  // the rest of step s-1's trailing update
#pragma unroll
      for (int u = 0; u < MAXB; ++u) {
        const int bi = own.bi[W][u], bj = own.bj[W][u];
        if (bi > s - 1 && !(bi == s && bj == s)) update(u, bi, bj);
      }
    }
    lds_barrier();  // B2: factored by the last wavefront
    if (*fail) break;
    // panel row R_sj = inv(R_ss)' G_sj by its owners (MFMA); the
    // accumulator layout of G_sj is the operand layout the product needs.
    // The owner of G_ss stores R_ss and inv(R_ss).  The store offsets pass
    // through empty asm: inside `bi == s` the compiler would otherwise know
    // the step and hoist every block's addresses out of the loop (hundreds
    // of registers, spilled).
    int jso = js, ldro = ldr, ldto = ldt;
    asm volatile("" : "+s"(jso), "+s"(ldro), "+s"(ldto));
#pragma unroll
    for (int u = 0; u < MAXB; ++u) {
      const int bi = own.bi[W][u], bj = own.bj[W][u];
      if (bi == s && bj == s) {
        const int c = lc, nb = min(16, k - js);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = lr + 4 * r;
          Dinv[jso * 16 + i * 16 + c] = Dv[i][c];
          if (i < nb && c < nb) {
            if (R) R[(jso + i) + int64_t(jso + c) * ldro] = Ds[i][c];
            Rt[int64_t(jso + i) * ldto + jso + c] = Ds[i][c];
          }
        }
      }
      if (bi == s && bj > s) {
        f64x4 o = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          o = __builtin_amdgcn_mfma_f64_16x16x4f64(Dv[4 * q + lr][lc], acc[u][q], o, 0, 0, 0);
        const int j = bj * 16 + lc;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = lr + 4 * r;
          pl[4 * r * SR + bj * 16] = o[r];
          if (j < k) {
            if (R) R[(jso + i) + int64_t(j) * ldro] = o[r];
            Rt[int64_t(jso + i) * ldto + j] = o[r];
          }
        }
      }
    }
    lds_barrier();  // B3: panel row s in LDS
    // G_{s+1,s+1} first, to LDS for the next B1 (only diagonal blocks qualify)
#pragma unroll
    for (int u = 0; u < MAXB; ++u) {
      const int bi = own.bi[W][u], bj = own.bj[W][u];
      if (bi == bj && bi == s + 1) {
        update(u, bi, bj);
#pragma unroll
        for (int r = 0; r < 4; ++r) Ds[lr + 4 * r][lc] = acc[u][r];
      }
    }
  }
}

template <int NBLK>
__device__ __forceinline__ void chol_factor_role(int k, int check_k, double (*Ds)[17],
                                                 double (*Dv)[17], const double* fl,
                                                 double* rowb, lds_ptr<volatile int> fail,
                                                 int lane) {
  __syncthreads();  // B0: G read into registers
  // Lane (i, g) = (lane / 4, lane % 4) holds columns 4g .. 4g+3 of row i of
  // the diagonal block G_ss (a) and of an identity beside it (x).  Gaussian
  // elimination of [G_ss | I] with each pivot row scaled by 1/sqrt(pivot)
  // leaves [R_ss | R_ss^-T]: the inverse comes out of the same 16 steps.
  // Step j: the pivot is broadcast from its lane, the four lanes of row j
  // scale it and publish it in LDS (rowb: 16 entries of R, 16 of R^-T; every
  // lane reads the same addresses, a conflict-free broadcast), then rows
  // i > j subtract R_ji times row j.  R_ic receives the same operations in
  // the same order as a column-by-column right-looking factorisation.
  const int i = lane >> 2, g = lane & 3;
  for (int s = 0; s < NBLK; ++s) {
    const int js = s * 16;
    if (js >= k) break;
    const int nb = min(16, k - js);
    lds_barrier();  // B1: G_ss in LDS
    // past nb the block is padded with the identity, so the fully unrolled
    // loop has no data-dependent exits
    double a[4], x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 4 * g + q;
      const double gv = Ds[i][c];
      a[q] = (i < nb && c < nb) ? gv : (i == c ? 1.0 : 0.0);
      x[q] = i == c ? 1.0 : 0.0;
    }
    // the block's pivot floors, read before the serial loop
    double flv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) flv[j] = fl[js + j];
    int bad = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      double dj = lane_bcast(a[j & 3], 4 * j + (j >> 2));  // entry (j, j)
      const double floor_j = j < nb ? flv[j] : 0.0;
      const bool ok = (dj > floor_j) && (dj > 0.0) && isfinite(dj);
      if (!bad && !ok && js + j < check_k) bad = js + j + 1;
      if (bad || !ok) dj = 1.0;  // unchecked columns past check_k continue
      // 1/sqrt(dj): reciprocal square-root estimate refined by two Newton
      // steps r += r (1 - dj r^2) / 2; then d = dj r with one correction
      double rd = __builtin_amdgcn_rsq(dj);
      rd = fma(0.5 * rd, fma(-dj * rd, rd, 1.0), rd);
      rd = fma(0.5 * rd, fma(-dj * rd, rd, 1.0), rd);
      double d = dj * rd;
      d = fma(0.5 * rd, fma(-d, d, dj), d);
      if (i == j) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          a[q] = 4 * g + q == j ? d : a[q] * rd;
          x[q] *= rd;
          rowb[4 * g + q] = a[q];
          rowb[16 + 4 * g + q] = x[q];
        }
      }
      lane_exchange();
      const double rji = rowb[i];  // R_ji
      double rj[4], xj[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        rj[q] = rowb[4 * g + q];
        xj[q] = rowb[16 + 4 * g + q];
      }
      lane_exchange();  // the next step's writes follow these reads
      if (i > j) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          a[q] -= rji * rj[q];
          x[q] -= rji * xj[q];
        }
      }
    }
    if (bad) {
      if (lane == 0) *fail = bad;
    } else {
      // R_ss (zero below the diagonal) and inv(R_ss) = (R_ss^-T)', zero
      // outside the nb x nb block
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = 4 * g + q;
        Ds[i][c] = i <= c ? a[q] : 0.0;
        Dv[c][i] = (c <= i && i < nb) ? x[q] : 0.0;
      }
    }
    lds_barrier();  // B2: done
    if (*fail) break;
    lds_barrier();  // B3: panel in LDS
  }
}

// One workgroup factors one matrix; pivots of columns >= check_k are not
// tested (an augmented column whose pivot is a residual may vanish).
template <int NBLK>
__device__ __forceinline__ void chol_blocks_body(int k, int check_k, const double* G, int ldg,
                                                 bool gsym,
                                                 double* R, int ldr, double* __restrict__ Rt,
                                                 int ldt, double* __restrict__ Dinv,
                                                 int* __restrict__ info) {
  constexpr int CP = NBLK * 16;
  __shared__ double P[16][CP + 1];  // panel row R_s* of the current step
  __shared__ double Ds[16][17];     // diagonal block G_ss
  __shared__ double Dv[16][17];     // inv(R_ss)
  __shared__ double fl[CP];         // pivot floors k * eps * G_jj
  __shared__ __attribute__((aligned(16))) double rowb[32];  // pivot row of the factor step
  __shared__ int fail;
  const int t = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  constexpr int BW = CholShape<NBLK>::block_waves;
  const global_ptr<const double> Gg = (global_ptr<const double>)G;
  const lds_ptr<volatile int> failp = (lds_ptr<volatile int>)&fail;
  for (int j = t; j < CP; j += CholShape<NBLK>::threads)
    fl[j] = j < k ? k * 2.220446049250313e-16 * Gg[j + j * ldg] : 0.0;
  if (t == 0) fail = 0;
  // The roles run separate loops with the same barriers (a workgroup barrier
  // counts arrivals, not program points).
  if (wave == BW) {
    chol_factor_role<NBLK>(k, check_k, Ds, Dv, fl, rowb, failp, t & 63);
  } else {
    switch (wave) {
#define RLA_ROLE(W)                                                                        \
  case W:                                                                                  \
    if constexpr (W < BW)                                                                  \
      chol_block_role<NBLK, W>(k, Gg, ldg, gsym, (global_ptr<double>)R, ldr,                  \
                               (global_ptr<double>)Rt, ldt, (global_ptr<double>)Dinv, P, Ds, \
                               Dv, failp, t);                                             \
    break;
      RLA_ROLE(0) RLA_ROLE(1) RLA_ROLE(2) RLA_ROLE(3) RLA_ROLE(4) RLA_ROLE(5) RLA_ROLE(6)
      RLA_ROLE(7) RLA_ROLE(8) RLA_ROLE(9) RLA_ROLE(10) RLA_ROLE(11) RLA_ROLE(12)
      RLA_ROLE(13) RLA_ROLE(14)
#undef RLA_ROLE
      default:
        break;
    }
  }
  if (t == 0) *(global_ptr<int>)info = fail;
}

template <int NBLK>
__global__ __launch_bounds__(CholShape<NBLK>::threads) void chol_blocks_kernel(
    int k, const double* G, int ldg, int gsym, double* R, int ldr, double* __restrict__ Rt,
    int ldt, double* __restrict__ Dinv, int* __restrict__ info) {
  chol_blocks_body<NBLK>(k, k, G, ldg, gsym != 0, R, ldr, Rt, ldt, Dinv, info);
}

// Batched form: workgroup p factors matrix p of a descriptor array.
template <int NBLK>
__global__ __launch_bounds__(CholShape<NBLK>::threads) void chol_blocks_batched_kernel(
    const CholBatchItem* __restrict__ items) {
  const CholBatchItem& it = items[blockIdx.x];
  chol_blocks_body<NBLK>(it.k, it.check_k, it.G, it.k, false, nullptr, 0, it.Rt, it.k, it.Dinv,
                         it.info);
}

// Yout <- Yin R^-1.  Each wavefront owns 16 rows and solves its column
// blocks in order, Ynew_b = (Y_b - sum_{l<b} Ynew_l R_lb) inv(R_bb), keeping
// the solved blocks in registers in the MFMA operand layout.  The panel of
// block column b (-R_lb for l < b, and inv(R_bb)) is staged in LDS once per
// workgroup; the next panel's loads are in flight while the current one is
// used.  A panel block is stored row by row (element (r, c) at 16 r + c), so
// the 64 lanes of an operand read touch 64 consecutive doubles.
// Yin is FP64 or FP32 (widened exactly).  OUT64: the result to Yout (FP64;
// Yout may be Yin itself: a wavefront reads a block of its rows before it
// writes it).  OUTBF: its BF16 rounding f32_to_bf16_bits(float(x)) to Ybf
// (column c at Ybf + c * ld_bf), the operand of a mixed-precision product.
// Nothing is written when any of the ninfo breakdown flags is set.
template <int NBLK, class Tin, bool OUT64, bool OUTBF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void trsm_ru_f64_kernel(
    int64_t m, int k, const Tin* Yin, int64_t ldy, double* Yout, int64_t ldo, uint16_t* Ybf,
    int64_t ld_bf, const double* __restrict__ Rt, int ldt, const double* __restrict__ Dinv,
    const int* __restrict__ info, int ninfo) {
  for (int i = 0; i < ninfo; ++i)
    if (info[i]) return;
  __shared__ double Ps[2][NBLK][256];  // slot l < b: -R_lb; slot NBLK - 1: inv(R_bb)
  __shared__ double Ts[4][16][17];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int lr = lane >> 4, lc = lane & 15;
  const int pr = t >> 4, pc = t & 15;  // panel element of this thread
  const int64_t rbase = int64_t(blockIdx.x) * 64 + wave * 16;
  const int64_t row = rbase + lc;  // operand layout: row lane%16, column 4q + lane/16
  const bool row_ok = row < m;
  // loads are unconditional from clamped (in-bounds) addresses and zeroed
  // afterwards, so they carry no branches and can all be in flight at once
  const Tin* yrow = Yin + (row_ok ? row : m - 1);
  double (*ts)[17] = Ts[wave];
  double ya[NBLK][4];
  double yn[4];
  double pf[NBLK];
  auto load_y = [&](int b) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = b * 16 + 4 * q + lr;
      const double x = static_cast<double>(yrow[int64_t(c < k ? c : k - 1) * ldy]);
      yn[q] = (row_ok && c < k) ? x : 0.0;
    }
  };
  // panel b into registers: pf[l] = -R(16 l + pr, 16 b + pc), l < b (zero
  // past column k), pf[NBLK - 1] = inv(R_bb)(pr, pc)
  auto load_panel = [&](int b) {
    const int c = b * 16 + pc;
    const double* rc = Rt + (c < k ? c : k - 1);
#pragma unroll
    for (int l = 0; l < NBLK - 1; ++l)
      if (l < b) {
        const double x = rc[int64_t(l * 16 + pr) * ldt];
        pf[l] = c < k ? -x : 0.0;
      }
    pf[NBLK - 1] = Dinv[b * 256 + t];
  };
  auto store_panel = [&](int buf, int b) {
#pragma unroll
    for (int l = 0; l < NBLK - 1; ++l)
      if (l < b) Ps[buf][l][t] = pf[l];
    Ps[buf][NBLK - 1][t] = pf[NBLK - 1];
  };
  load_y(0);
  load_panel(0);
  store_panel(0, 0);
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NBLK; ++b) {
    if (b * 16 >= k) break;
    const int buf = b & 1;
    const bool more = (b + 1) * 16 < k;
    if (more) load_panel(b + 1);
    // Y_b into the accumulator layout (through LDS)
#pragma unroll
    for (int q = 0; q < 4; ++q) ts[lc][4 * q + lr] = yn[q];
    lane_exchange();
    f64x4 acc;
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = ts[lr + 4 * r][lc];
    lane_exchange();
    if (more) load_y(b + 1);
    // element (4q + lr, lc) of panel block l: pb[256 l + 64 q]
    const double* pb = &Ps[buf][0][0] + 16 * lr + lc;
    // T = Y_b - sum_{l<b} Ynew_l R_lb
#pragma unroll
    for (int l = 0; l < NBLK - 1; ++l)
      if (l < b) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ya[l][q], pb[256 * l + 64 * q], acc, 0, 0,
                                                     0);
      }
    // Ynew_b = T inv(R_bb)
#pragma unroll
    for (int r = 0; r < 4; ++r) ts[lr + 4 * r][lc] = acc[r];
    lane_exchange();
    double ta[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) ta[q] = ts[lc][4 * q + lr];
    lane_exchange();
    acc = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ta[q], pb[256 * (NBLK - 1) + 64 * q], acc, 0, 0,
                                                 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) ts[lr + 4 * r][lc] = acc[r];
    lane_exchange();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ya[b][q] = ts[lc][4 * q + lr];
      const int c = b * 16 + 4 * q + lr;
      if (row_ok && c < k) {
        if (OUT64) Yout[row + int64_t(c) * ldo] = ya[b][q];
        if (OUTBF) Ybf[row + int64_t(c) * ld_bf] = f32_to_bf16_bits(static_cast<float>(ya[b][q]));
      }
    }
    lane_exchange();
    if (more) {
      // Ps[buf ^ 1] held panel b - 1: every wavefront finished with it before
      // the previous iteration's barrier
      store_panel(buf ^ 1, b + 1);
      __syncthreads();  // panel b + 1 in LDS
    }
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
// g_symmetric: G holds both triangles (faster, coalesced reads); otherwise
// only its upper triangle is read.
int chol_blocks(rla_ctx* ctx, int64_t k, const double* G, int64_t ldg, double* R, int64_t ldr,
                double* Rt, int64_t ldt, double* Dinv, int* info_dev, bool g_symmetric) {
  const int nb = chol_nblk(k);
  if (!nb) return fail(RLA_ENOTSUP, "chol_blocks: k must be in 1..288");
  const int ki = static_cast<int>(k);
#define RLA_CHOL(NB)                                                                       \
  case NB:                                                                                 \
    chol_blocks_kernel<NB><<<1, CholShape<NB>::threads, 0, ctx->stream>>>(                           \
        ki, G, static_cast<int>(ldg), g_symmetric ? 1 : 0, R, static_cast<int>(ldr), Rt,    \
        static_cast<int>(ldt), Dinv, info_dev);                                           \
    break;
  switch (nb) {
    RLA_CHOL(2) RLA_CHOL(4) RLA_CHOL(8) RLA_CHOL(12) RLA_CHOL(16) RLA_CHOL(17) RLA_CHOL(18)
  }
#undef RLA_CHOL
  RLA_LAUNCH_CHECK("chol_blocks");
  return RLA_OK;
}

int chol_blocks_batched(hipStream_t stream, int64_t nitems, const CholBatchItem* items) {
  if (nitems <= 0) return RLA_OK;
  chol_blocks_batched_kernel<17><<<dim3(unsigned(nitems)), CholShape<17>::threads, 0, stream>>>(
      items);
  RLA_LAUNCH_CHECK("chol_blocks_batched");
  return RLA_OK;
}

template <class Tin>
int trsm_ru(rla_ctx* ctx, int64_t m, int64_t k, const Tin* Yin, int64_t ldy, double* Yout,
            int64_t ldo, uint16_t* Ybf, int64_t ld_bf, const double* Rt, int64_t ldt,
            const double* Dinv, const int* info, int ninfo) {
  const int nb = chol_nblk(k);
  if (!nb) return fail(RLA_ENOTSUP, "trsm_ru_f64: k must be in 1..288");
  if ((Yout != nullptr) == (Ybf != nullptr))
    return fail(RLA_EARG, "trsm_ru_f64: exactly one of the FP64 and BF16 outputs");
  if (m <= 0) return RLA_OK;
  if (!info) ninfo = 0;
  const int ki = static_cast<int>(k);
  const unsigned grid = static_cast<unsigned>(ceil_div(m, 64));
#define RLA_TRSM_GO(NB, O64, OBF)                                                            \
  trsm_ru_f64_kernel<NB, Tin, O64, OBF><<<dim3(grid), 256, 0, ctx->stream>>>(                \
      m, ki, Yin, ldy, Yout, ldo, Ybf, ld_bf, Rt, static_cast<int>(ldt), Dinv, info, ninfo)
#define RLA_TRSM(NB)                    \
  case NB:                              \
    if (Yout)                           \
      RLA_TRSM_GO(NB, true, false);     \
    else                                \
      RLA_TRSM_GO(NB, false, true);     \
    break;
  switch (nb) {
    RLA_TRSM(2) RLA_TRSM(4) RLA_TRSM(8) RLA_TRSM(12) RLA_TRSM(16) RLA_TRSM(17) RLA_TRSM(18)
  }
#undef RLA_TRSM
#undef RLA_TRSM_GO
  RLA_LAUNCH_CHECK("trsm_ru_f64");
  return RLA_OK;
}

template int trsm_ru<double>(rla_ctx*, int64_t, int64_t, const double*, int64_t, double*,
                             int64_t, uint16_t*, int64_t, const double*, int64_t, const double*,
                             const int*, int);
template int trsm_ru<float>(rla_ctx*, int64_t, int64_t, const float*, int64_t, double*, int64_t,
                            uint16_t*, int64_t, const double*, int64_t, const double*,
                            const int*, int);

// Y <- Y R^-1 in place; skipped when *info != 0
int trsm_ru_f64(rla_ctx* ctx, int64_t m, int64_t k, double* Y, int64_t ldy, const double* Rt,
                int64_t ldt, const double* Dinv, const int* info) {
  return trsm_ru<double>(ctx, m, k, Y, ldy, Y, ldy, nullptr, 0, Rt, ldt, Dinv, info, 1);
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
  rc = chol_blocks(ctx, k, G, ldg, R, ldr, Rt, k, Dinv, info, false);
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
