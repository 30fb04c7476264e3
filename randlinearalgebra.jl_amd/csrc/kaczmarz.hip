// Block Kaczmarz (src/Solvers/kaczmarz.jl) with the Sampling compressor and
// its Uniform / L2Norm distributions (src/Compressors/sampling.jl,
// Distributions/uniform.jl, Distributions/strohmer_vershynin.jl), the LQ
// sub-solve (src/Solvers/SubSolvers/lq.jl), FullResidual and BasicLogger.
//
// One iteration projects x onto the solution set of a block of s rows,
//   x <- x + alpha * A~' (A~ A~')^-1 (b~ - A~ x),
// the minimum-norm correction the LQ sub-solver computes (lq.jl:35-46).  On the
// device an iteration is one launch of kz_step_kernel: its workgroups apply
// the pending update of the previous block to their columns of x, gather the
// new block's columns from the row-major copy of A and form partial Gram
// matrices (FP64 MFMA on FP32 entries, exact products) and partial A~x; the
// last workgroup to arrive sums the partials, factors the Gram matrix and
// solves for y.  The update of block t is folded into the launch that builds
// block t+1, so A~ is read twice per iteration and x once.
//
// Sampling: the row sets of a whole run of iterations are drawn at once.
// Without replacement the draw is the Efraimidis–Spirakis A-Res scheme
// (key_i = log(u_i) / w_i, keep the s largest), the successive-sampling law
// StatsBase's wsample!(replace=false) implements; with replacement each draw
// inverts the cumulative weights.  Sets are returned sorted ascending
// (ordered=true).  Streams are counter-based Philox, so every set is a pure
// function of (seed, epoch, item).
#include <cmath>

#include "philox.hpp"
#include "rla_internal.hpp"

namespace rla {

constexpr uint32_t DOM_WSAMPLE = 5u;
constexpr int kSortN = 4096;      // candidates per top-k workgroup
constexpr int kSortThreads = 512;
constexpr int kKzCols = 64;       // columns per chunk of the Kaczmarz step
constexpr int kKzMaxBlock = 64;   // rows per Kaczmarz block

// ---- row weights ----------------------------------------------------------

// w_i = sum_j A[i,j]^2 (strohmer_vershynin.jl:82, sum(abs2, A, dims=2)),
// accumulated in FP64 in column order.
template <class T>
__global__ void row_norms2_kernel(int64_t m, int64_t n, const T* __restrict__ A, int64_t a_rs,
                                  int64_t a_cs, double* __restrict__ w) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= m) return;
  const T* row = A + i * a_rs;
  double s = 0.0;
  for (int64_t j = 0; j < n; ++j) {
    const double v = static_cast<double>(row[j * a_cs]);
    s += v * v;
  }
  w[i] = s;
}

// ---- top-k sampling without replacement ------------------------------------

// (key, index) order: larger key first, ties by smaller index
__device__ __forceinline__ bool before(double ka, int ia, double kb, int ib) {
  return ka > kb || (ka == kb && ia < ib);
}

// bitonic sort of N (power of two) LDS pairs into "before" order
__device__ void bitonic_sort(double* key, int* idx, int N) {
  for (int size = 2; size <= N; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int e = threadIdx.x; e < N / 2; e += blockDim.x) {
        const int i = 2 * stride * (e / stride) + (e % stride);
        const int j = i + stride;
        const bool up = (i & size) == 0;  // this half sorts in "before" order
        const bool swap = up ? before(key[j], idx[j], key[i], idx[i])
                             : before(key[i], idx[i], key[j], idx[j]);
        if (swap) {
          const double tk = key[i];
          key[i] = key[j];
          key[j] = tk;
          const int ti = idx[i];
          idx[i] = idx[j];
          idx[j] = ti;
        }
      }
      __syncthreads();
    }
  }
}

// ascending sort of N (power of two) LDS ints
__device__ void bitonic_sort_int(int* v, int N) {
  for (int size = 2; size <= N; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int e = threadIdx.x; e < N / 2; e += blockDim.x) {
        const int i = 2 * stride * (e / stride) + (e % stride);
        const int j = i + stride;
        const bool up = (i & size) == 0;
        if (up ? v[i] > v[j] : v[i] < v[j]) {
          const int t = v[i];
          v[i] = v[j];
          v[j] = t;
        }
      }
      __syncthreads();
    }
  }
}

__host__ __device__ inline int pow2_at_least(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

// Order-preserving code of a key: a larger double gives a larger code.
__device__ __forceinline__ uint64_t topk_code(double k) {
  const uint64_t b = static_cast<uint64_t>(__double_as_longlong(k));
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Pass over candidates (n_rows items in all).  first: the candidates are items
// [0, n_items) with keys drawn here (set r uses epoch0 + r); otherwise they
// are the (key, index) pairs in_key/in_idx[r][0..n_items).  Each workgroup
// takes a chunk of kSortN and keeps its k largest keys, found by a radix
// select on the keys' order-preserving codes: 8-bit digits from the top, a
// histogram of the candidates still tied with the selected prefix per digit,
// until the boundary digit holds exactly the number still to take.  The kept
// set does not depend on timing (keys are continuous draws: equal codes occur
// only for zero weights, key -inf, which the final pass never returns when
// at least k weights are positive); its order in the output does.  final
// (one chunk): the k best are sorted by index and written to out_set[r][0..k).
__global__ __launch_bounds__(kSortThreads) void topk_pass_kernel(
    uint64_t seed, uint64_t epoch0, int64_t n_items, int64_t n_rows, const double* __restrict__ w,
    int k,
    const double* __restrict__ in_key, const int* __restrict__ in_idx, double* __restrict__ out_key,
    int* __restrict__ out_idx, int* __restrict__ out_set, int first, int final_pass) {
  constexpr int E = kSortN / kSortThreads;
  __shared__ int hist[256];
  __shared__ int sel_bin, sel_rem, sel_done, outc, tiec;
  __shared__ int fin[1024];
  const int t = threadIdx.x, lane = t & 63;
  const int r = blockIdx.y, chunk = blockIdx.x;
  const int64_t base = int64_t(chunk) * kSortN;
  const int cnt = static_cast<int>(min<int64_t>(kSortN, n_items - base));
  double kk[E];
  int ii[E];
  uint64_t code[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = t + q * kSortThreads;
    kk[q] = -INFINITY;
    ii[q] = 0x7fffffff;
    if (e < cnt) {
      const int64_t item = base + e;
      if (first) {
        Stream st(seed, epoch0 + r, DOM_WSAMPLE, static_cast<uint64_t>(item));
        const double u = st.open01();
        const double wi = w ? w[item] : 1.0;
        kk[q] = wi > 0.0 ? log(u) / wi : -INFINITY;
        ii[q] = static_cast<int>(item);
      } else {
        kk[q] = in_key[int64_t(r) * n_items + item];
        ii[q] = in_idx[int64_t(r) * n_items + item];
      }
    }
    code[q] = topk_code(kk[q]);
  }
  const int take = min(k, cnt);
  uint64_t prefix = 0, pmask = 0;
  bool all = take == cnt;  // every candidate is kept
  int rem = take;
  if (t == 0) {
    outc = 0;
    tiec = 0;
  }
  __syncthreads();
  for (int shift = 56; shift >= 0 && !all; shift -= 8) {
    if (t < 256) hist[t] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (t + q * kSortThreads < cnt && (code[q] & pmask) == prefix)
        atomicAdd(&hist[int((code[q] >> shift) & 255u)], 1);
    __syncthreads();
    if (t < 64) {
      // suffix sums over bins (lane l: bins 4l .. 4l + 3), from the top
      int h[4], s4 = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) s4 += (h[b] = hist[4 * lane + b]);
      int suf = s4;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_down(suf, d);
        if (lane + d < 64) suf += o;
      }
      int above = __shfl_down(suf, 1);
      if (lane == 63) above = 0;
      if (suf >= rem && above < rem) {  // exactly one lane
        int cum = above;
#pragma unroll
        for (int b = 3; b >= 0; --b) {
          if (cum + h[b] >= rem) {
            const int nr = rem - cum;
            sel_bin = 4 * lane + b;
            sel_rem = nr;
            sel_done = h[b] == nr;
            break;
          }
          cum += h[b];
        }
      }
    }
    __syncthreads();
    prefix |= uint64_t(sel_bin) << shift;
    pmask |= uint64_t(255u) << shift;
    rem = sel_rem;
    all = sel_done != 0;  // the boundary bin is taken whole
  }
  // kept: candidates above the selected prefix, those equal to it (all of
  // them when the boundary bin was taken whole, else rem of the exact ties)
  const bool whole = take == cnt || all;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = t + q * kSortThreads;
    if (e >= cnt) continue;
    const uint64_t c = code[q] & pmask;
    bool keep = take == cnt || c > prefix;
    if (!keep && c == prefix) keep = whole || atomicAdd(&tiec, 1) < rem;
    if (!keep) continue;
    const int slot = atomicAdd(&outc, 1);
    if (!final_pass) {
      const int64_t nout = int64_t(gridDim.x) * k;
      out_key[int64_t(r) * nout + int64_t(chunk) * k + slot] = kk[q];
      out_idx[int64_t(r) * nout + int64_t(chunk) * k + slot] = ii[q];
    } else {
      fin[slot] = ii[q];
    }
  }
  if (!final_pass) {
    // slots past the kept ones (a chunk of fewer than k candidates)
    const int64_t nout = int64_t(gridDim.x) * k;
    for (int e = take + t; e < k; e += kSortThreads) {
      out_key[int64_t(r) * nout + int64_t(chunk) * k + e] = -INFINITY;
      out_idx[int64_t(r) * nout + int64_t(chunk) * k + e] = 0x7fffffff;
    }
    return;
  }
  // k best by index, ascending (ordered = true); entries past take are padding
  const int K = pow2_at_least(k);
  __syncthreads();
  for (int e = take + t; e < K; e += kSortThreads) fin[e] = 0x7fffffff;
  __syncthreads();
  bitonic_sort_int(fin, K);
  // the host guarantees at least k items of positive weight; the clamp only
  // keeps a violated precondition in bounds
  for (int e = t; e < k; e += kSortThreads)
    out_set[int64_t(r) * k + e] = static_cast<int>(min<int64_t>(fin[e], n_rows - 1));
}

// ---- sampling with replacement ----------------------------------------------

// inclusive prefix sums of w in FP64 (one workgroup; chunk per thread in order)
__global__ __launch_bounds__(1024) void cdf_kernel(int64_t n, const double* __restrict__ w,
                                                   double* __restrict__ cdf) {
  __shared__ double part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t b = t * per, e = min<int64_t>(n, b + per);
  double s = 0.0;
  for (int64_t i = b; i < e; ++i) s += w ? w[i] : 1.0;
  part[t] = s;
  __syncthreads();
  if (t == 0)
    for (int q = 1; q < 1024; ++q) part[q] += part[q - 1];
  __syncthreads();
  double run = t ? part[t - 1] : 0.0;
  for (int64_t i = b; i < e; ++i) {
    run += w ? w[i] : 1.0;
    cdf[i] = run;
  }
}

__global__ __launch_bounds__(1024) void replace_sample_kernel(uint64_t seed, uint64_t epoch0,
                                                              int64_t n, const double* __restrict__ cdf,
                                                              int k, int* __restrict__ out_set) {
  __shared__ int v[kSortN];
  const int r = blockIdx.x;
  const double total = cdf[n - 1];
  const int K = pow2_at_least(k);
  for (int e = threadIdx.x; e < K; e += blockDim.x) {
    int pick = 0x7fffffff;
    if (e < k) {
      Stream st(seed, epoch0 + r, DOM_WSAMPLE, static_cast<uint64_t>(n) + e);
      const double target = st.open01() * total;
      int64_t lo = 0, hi = n - 1;  // first i with cdf[i] > target
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cdf[mid] > target) hi = mid; else lo = mid + 1;
      }
      pick = static_cast<int>(lo);
    }
    v[e] = pick;
  }
  __syncthreads();
  bitonic_sort_int(v, K);
  for (int e = threadIdx.x; e < k; e += blockDim.x) out_set[int64_t(r) * k + e] = v[e];
}

// ---- block update ------------------------------------------------------------

// Rows of a block: explicit (element (i, c) at base[i*rs + c*cs]) or gathered
// from a matrix by row index (element (i, c) at base[idx[i]*rs + c*cs]).
struct RowSrc {
  const float* base;
  int64_t rs, cs;
  const int* idx;
  int s;
  __device__ __forceinline__ float at(int i, int64_t c) const {
    const int64_t row = idx ? int64_t(idx[i]) : int64_t(i);
    return base[row * rs + c * cs];
  }
};

typedef double kz_f64x4 __attribute__((ext_vector_type(4)));

constexpr int kKzTiles = 10;              // upper 16 x 16 tiles of a 64 x 64 Gram matrix
constexpr int kKzTs = kKzCols + 4;        // LDS block-tile row stride (floats, 16-byte rows)
constexpr int kKzGs = kKzMaxBlock + 1;    // LDS Gram row stride (doubles)
constexpr int kKzMaxWG = 8;               // workgroups per step: partials the solver sums
constexpr int kKzTileDoubles = 256;       // one accumulator tile: 4 x 64 lanes

// One Kaczmarz iteration (one launch of kz_step_kernel).
struct KzStep {
  int64_t n;
  int chunks_per_wg;     // 64-column chunks each workgroup sweeps
  RowSrc prev;           // the pending block (prev.s == 0: none)
  const double* y_prev;  // its coefficients
  float alpha;
  float* x;
  RowSrc cur;            // the new block (cur.s == 0: apply the pending update only)
  RowSrc bsrc;           // b~ of the new block, element (i, 0)
  double* part_G;        // per workgroup: kKzTiles accumulator tiles
  double* part_d;        // per workgroup: kKzMaxBlock partial sums of A~x
  unsigned* ticket;      // arrival counter; 0 between launches
  double* y_out;         // coefficients of the new block
  int* info;             // sticky: 1-based pivot of a numerically singular block
};

// upper tile b = 0..9 of a 4 x 4 tile grid -> (tile row, tile column)
__device__ __forceinline__ int kz_ti(int b) { return int((0x3221110000ull >> (4 * b)) & 15u); }
__device__ __forceinline__ int kz_tj(int b) { return int((0x3323213210ull >> (4 * b)) & 15u); }

// value of v in lane src (src uniform), in every lane
__device__ __forceinline__ double kz_bcast(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// LDS traffic between lanes of one wavefront needs no fence (the LDS runs a
// wavefront's instructions in order); this only keeps the compiler from
// moving accesses across the exchange point.
__device__ __forceinline__ void kz_lane_order() { __builtin_amdgcn_wave_barrier(); }

// x <- x + alpha A~_prev' y_prev, then for the new block A~ (rows cur, s <= 64)
// y_out = (A~ A~')^-1 (b~ - A~ x) — the minimum-norm correction of the LQ
// sub-solve (lq.jl:35-46, kaczmarz.jl:294-311) — in one launch:
//   * every workgroup sweeps its 64-column chunks (the next chunk's loads in
//     flight while the current one is processed): the pending update of its
//     columns of x, the new block's columns gathered from the row-major copy
//     of A into LDS, the partial A~x, and the partial Gram matrix as ten
//     upper 16 x 16 tiles on the FP64 MFMA (FP32 entries: exact products);
//   * the last workgroup to arrive sums the partials in workgroup order
//     (deterministic; all loads of the sum in flight at once), and one
//     wavefront factors G = L D L' (unit lower L; left-looking over 16-column
//     block columns: the update of a block column from the factored ones on
//     the MFMA, then a right-looking factor of its 16 columns with lane i
//     holding row i, the forward solve L z = r folded in), then solves
//     L' y = D^-1 z.
// A pivot d_j not above 4 s eps G_jj (a numerically dependent row, e.g. a
// row drawn twice with replacement) gets a zero column of L and y_j = 0: the
// correction projects onto the independent rows.  A non-finite pivot gives
// y = 0 and sets *info to its 1-based position; an all-zero block gives
// y = 0 (kaczmarz.jl:300 skips it).
// Gram tiles of wavefront W: b = W, W + 4, W + 8 (tile row, tile column from
// kz_ti / kz_tj), as compile-time lists so the operands stay in registers.
template <int W>
__device__ __forceinline__ void kz_gram_tiles(const float (&op)[4][16], kz_f64x4 (&acc)[3]) {
  constexpr int nt = W < 2 ? 3 : 2;
  constexpr int ti[3] = {0, W == 0 ? 1 : (W == 3 ? 2 : 1), W == 0 ? 2 : 3};
  constexpr int tj[3] = {W, W == 0 ? 1 : (W == 1 ? 2 : (W == 2 ? 3 : 2)), W == 0 ? 3 : 3};
#pragma unroll
  for (int kq = 0; kq < 16; ++kq) {
    double o[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = static_cast<double>(op[r][kq]);
#pragma unroll
    for (int q = 0; q < nt; ++q)
      acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(o[ti[q]], o[tj[q]], acc[q], 0, 0, 0);
  }
}

// LDL' of the (padded) Gram matrix in Gs by one wavefront, lane i holding
// row i: unit lower L left-looking over 16-column block columns (the update
// of a block column from the factored ones on the MFMA, then a right-looking
// factor of its 16 columns).  A pivot d_j not above the floor of its row
// (fl, lane j) is a row dependent on the rows before it: its column of L and
// 1/d_j are zero, so solves project onto the independent rows.  With RHS,
// z (lane i: entry i) is solved forward with unit L on the way.  On return Gs
// holds L strictly below the diagonal, Dv the pivots; lane i has 1/d_i (0
// for a dependent pivot) in invme and isfinite(d_i) in okme.
template <bool RHS>
__device__ __forceinline__ void kz_ldl(double* Gs, double* Dv, int nb, int lane, double fl,
                                       double& z, double& invme, bool& okme) {
  const int i = lane;
  double dme = 1.0;
  // Only entries strictly below the diagonal of L are read back (by the
  // MFMA update and the back-solve); lanes above a pivot compute values
  // that nothing reads.
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    if (kb < nb) {
      const int r0 = 16 * kb;
      // G(r0:, r0:r0+15) -= L(r0:, 0:r0-1) D L(r0:r0+15, 0:r0-1)'
#pragma unroll
      for (int I = kb; I < 4; ++I) {
        if (kb > 0 && I < nb) {
          kz_f64x4 acc4;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            acc4[r] = Gs[(16 * I + (lane >> 4) + 4 * r) * kKzGs + r0 + (lane & 15)];
#pragma unroll
          for (int k0 = 0; k0 < r0; k0 += 4) {
            const int kk = k0 + (lane >> 4);
            const double av_ = -Gs[(16 * I + (lane & 15)) * kKzGs + kk] * Dv[kk];
            const double bv_ = Gs[(r0 + (lane & 15)) * kKzGs + kk];
            acc4 = __builtin_amdgcn_mfma_f64_16x16x4f64(av_, bv_, acc4, 0, 0, 0);
          }
#pragma unroll
          for (int r = 0; r < 4; ++r)
            Gs[(16 * I + (lane >> 4) + 4 * r) * kKzGs + r0 + (lane & 15)] = acc4[r];
        }
      }
      kz_lane_order();
      double a[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) a[c] = Gs[i * kKzGs + r0 + c];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int jj = r0 + j;
        const double d = kz_bcast(a[j], jj);     // pivot d_jj
        const double flj = kz_bcast(fl, jj);     // its floor
        // A pivot not above its floor is a row dependent on the rows before
        // it (a row drawn twice, or a combination of others): its column of
        // L and its coefficient are zero, so the projection is onto the span
        // of the independent rows (exact for a consistent block).
        const bool dep = !(d > flj);
        // 1/d: reciprocal estimate refined by two Newton steps e = 1 - d r, r += r e
        double inv = __builtin_amdgcn_rcp(d);
        inv = fma(inv, fma(-d, inv, 1.0), inv);
        inv = fma(inv, fma(-d, inv, 1.0), inv);
        if (dep) inv = 0.0;
        if (i == jj) {
          okme = isfinite(d);
          dme = d;
          invme = inv;
        }
        const double u = dep ? 0.0 : a[j];  // L_ij d_j
        const double l = u * inv;
        a[j] = l;
        // forward solve with unit L, right-looking
        if constexpr (RHS) {
          const double zj = kz_bcast(z, jj);
          if (i > jj) z = fma(-l, zj, z);
        }
        // trailing columns of this block column: a_ic -= L_ij d_j L_cj
#pragma unroll
        for (int c = j + 1; c < 16; ++c) a[c] = fma(-u, kz_bcast(l, r0 + c), a[c]);
      }
#pragma unroll
      for (int c = 0; c < 16; ++c) Gs[i * kKzGs + r0 + c] = a[c];
      if (i >= r0 && i < r0 + 16) Dv[i] = dme;
      kz_lane_order();
    }
  }
}

template <bool VEC>
__global__ __launch_bounds__(256) void kz_step_kernel(KzStep p) {
  __shared__ __align__(16) float T[kKzMaxBlock * kKzTs];
  __shared__ __align__(16) float Tp[kKzMaxBlock * kKzTs];
  __shared__ float xs[kKzCols];
  __shared__ double ured[4][kKzCols];
  __shared__ double Gs[kKzMaxBlock * kKzGs];
  __shared__ double Dv[kKzMaxBlock];
  __shared__ double rv[kKzMaxBlock];
  __shared__ int64_t prow[kKzMaxBlock], crow[kKzMaxBlock];
  __shared__ double yprev[kKzMaxBlock];
  __shared__ int last;
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int s = p.cur.s, ps = p.prev.s;
  const int nb = (s + 15) >> 4;  // 16-row tiles of the new block
  // Row offsets of both blocks, clamped (rows past a block repeat its last).
  // Loads never depend on which blocks are present: a missing block reads
  // the other block's rows, and its values go unused.
  __shared__ int cidx[kKzMaxBlock];
  const RowSrc P = ps > 0 ? p.prev : p.cur;
  const RowSrc C = s > 0 ? p.cur : p.prev;
  if (t < kKzMaxBlock) {
    const int ip = min(t, P.s - 1), ic = min(t, C.s - 1);
    const int64_t rp = P.idx ? int64_t(P.idx[ip]) : int64_t(ip);
    const int64_t rc = C.idx ? int64_t(C.idx[ic]) : int64_t(ic);
    prow[t] = rp * P.rs;
    crow[t] = rc * C.rs;
    cidx[t] = static_cast<int>(rc);
    yprev[t] = p.y_prev[min(t, max(ps, 1) - 1)];
  }
  __syncthreads();

  kz_f64x4 acc[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) acc[q] = kz_f64x4{0.0, 0.0, 0.0, 0.0};
  double dpart = 0.0;  // lanes < 16: row 16 wv + lane of A~x
  const int64_t nchunks = (p.n + kKzCols - 1) / kKzCols;
  // Loads of a chunk (64 rows x 64 columns of each block): with VEC, float4
  // e = t + 256 k covers row e/16, columns 4 (e%16) ..; otherwise element
  // e = t + 256 k is row e/64, column e%64.  Rows past a block repeat its
  // last row and columns past n are clamped, so every load is in bounds.
  float sp[2][16], sc[2][16];
  float xv[2] = {0.f, 0.f};
  auto load_chunk = [&](int64_t ch, float(&lp)[16], float(&lc)[16], float& lx) {
    const int64_t c0 = ch * kKzCols;
    if constexpr (VEC) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int e = t + 256 * k, row = e >> 4;
        const int64_t c = min<int64_t>(c0 + 4 * (e & 15), p.n - 4);
        const float4 v = *reinterpret_cast<const float4*>(P.base + prow[row] + c);
        const float4 w = *reinterpret_cast<const float4*>(C.base + crow[row] + c);
        lp[4 * k] = v.x, lp[4 * k + 1] = v.y, lp[4 * k + 2] = v.z, lp[4 * k + 3] = v.w;
        lc[4 * k] = w.x, lc[4 * k + 1] = w.y, lc[4 * k + 2] = w.z, lc[4 * k + 3] = w.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int e = t + 256 * k, row = e >> 6;
        const int64_t c = c0 + min<int64_t>(e & 63, p.n - c0 - 1);
        lp[k] = P.base[prow[row] + c * P.cs];
        lc[k] = C.base[crow[row] + c * C.cs];
      }
    }
    lx = p.x[c0 + min<int64_t>(lane, p.n - c0 - 1)];
  };
  auto run_chunk = [&](int64_t ch, const float(&lp)[16], const float(&lc)[16], float lx) {
    const int64_t c0 = ch * kKzCols;
    const int ncol = static_cast<int>(min<int64_t>(kKzCols, p.n - c0));
    // stage both blocks in LDS (new block: rows >= s and columns >= ncol are 0)
    if constexpr (VEC) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int e = t + 256 * k, row = e >> 4, cb = 4 * (e & 15);
        if (ps > 0)
          *reinterpret_cast<float4*>(&Tp[row * kKzTs + cb]) =
              float4{lp[4 * k], lp[4 * k + 1], lp[4 * k + 2], lp[4 * k + 3]};
        if (s > 0) {
          const bool rok = row < s;
          *reinterpret_cast<float4*>(&T[row * kKzTs + cb]) =
              float4{rok && cb < ncol ? lc[4 * k] : 0.f, rok && cb + 1 < ncol ? lc[4 * k + 1] : 0.f,
                     rok && cb + 2 < ncol ? lc[4 * k + 2] : 0.f,
                     rok && cb + 3 < ncol ? lc[4 * k + 3] : 0.f};
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int e = t + 256 * k, row = e >> 6, cc = e & 63;
        if (ps > 0) Tp[row * kKzTs + cc] = lp[k];
        if (s > 0) T[row * kKzTs + cc] = (row < s && cc < ncol) ? lc[k] : 0.f;
      }
    }
    __syncthreads();
    // pending update: rows 16 wv .. 16 wv + 15 of the previous block, column lane
    if (ps > 0) {
      double u = 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (16 * wv + q < ps)
          u += static_cast<double>(Tp[(16 * wv + q) * kKzTs + lane]) * yprev[16 * wv + q];
      ured[wv][lane] = u;
      __syncthreads();
    }
    if (t < kKzCols) {
      float xc = lx;
      if (lane < ncol && ps > 0) {
        const double u = ((ured[0][lane] + ured[1][lane]) + ured[2][lane]) + ured[3][lane];
        xc = p.alpha * static_cast<float>(u) + xc;
        p.x[c0 + lane] = xc;
      }
      xs[lane] = lane < ncol ? xc : 0.f;
    }
    if (s > 0) {
      __syncthreads();
      // MFMA operands of this lane: T[16 rb + lane%16][4 kq + lane/16]
      float op[4][16];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int kq = 0; kq < 16; ++kq)
          op[rb][kq] = T[(16 * rb + (lane & 15)) * kKzTs + 4 * kq + (lane >> 4)];
      // partial A~x of rows 16 wv .. 16 wv + 15: 16 columns per lane, then
      // the four column groups ((g0 + g1) + (g2 + g3))
      {
        double d = 0.0;
#pragma unroll
        for (int kq = 0; kq < 16; ++kq)
          d += static_cast<double>(T[(16 * wv + (lane & 15)) * kKzTs + 4 * kq + (lane >> 4)]) *
               static_cast<double>(xs[4 * kq + (lane >> 4)]);
        d += __shfl_xor(d, 16);
        d += __shfl_xor(d, 32);
        dpart += d;
      }
      switch (wv) {
        case 0: kz_gram_tiles<0>(op, acc); break;
        case 1: kz_gram_tiles<1>(op, acc); break;
        case 2: kz_gram_tiles<2>(op, acc); break;
        default: kz_gram_tiles<3>(op, acc); break;
      }
    }
    __syncthreads();  // T, Tp, xs and ured are rewritten by the next chunk
  };
  // chunks blockIdx.x + ci * gridDim.x, two per round of loads
  for (int ci = 0; ci < p.chunks_per_wg; ci += 2) {
    const int64_t ch0 = int64_t(blockIdx.x) + int64_t(ci) * gridDim.x;
    const int64_t ch1 = ch0 + gridDim.x;
    if (ch0 >= nchunks) break;
    const bool two = ci + 1 < p.chunks_per_wg && ch1 < nchunks;
    load_chunk(ch0, sp[0], sc[0], xv[0]);
    if (two) load_chunk(ch1, sp[1], sc[1], xv[1]);
    run_chunk(ch0, sp[0], sc[0], xv[0]);
    if (two) run_chunk(ch1, sp[1], sc[1], xv[1]);
  }
  if (s == 0) return;

  // ---- partials out; the last workgroup to arrive solves ----
  const int wg = blockIdx.x;
  if (lane < 16) p.part_d[int64_t(wg) * kKzMaxBlock + 16 * wv + lane] = dpart;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int b = wv + 4 * q;
    if (b >= kKzTiles) break;
    double* dst = p.part_G + (int64_t(wg) * kKzTiles + b) * kKzTileDoubles + lane;
#pragma unroll
    for (int r = 0; r < 4; ++r) dst[r * 64] = acc[q][r];
  }
  __threadfence();
  __syncthreads();
  if (t == 0) last = atomicAdd(p.ticket, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();

  // element t of every tile, and row t of A~x, summed over the workgroups
  const int nwg = gridDim.x;
  {
    const double bt = static_cast<double>(p.bsrc.base[int64_t(cidx[t & (kKzMaxBlock - 1)]) * p.bsrc.rs]);
    double g[kKzTiles][kKzMaxWG];
    double dd[kKzMaxWG];
    // unconditional loads from clamped indices (inside the workspace; values
    // of workgroups past nwg and tiles past the block are discarded below)
#pragma unroll
    for (int b = 0; b < kKzTiles; ++b)
#pragma unroll
      for (int v = 0; v < kKzMaxWG; ++v)
        g[b][v] = p.part_G[(int64_t(min(v, nwg - 1)) * kKzTiles + b) * kKzTileDoubles + t];
#pragma unroll
    for (int v = 0; v < kKzMaxWG; ++v)
      dd[v] = p.part_d[int64_t(min(v, nwg - 1)) * kKzMaxBlock + (t & (kKzMaxBlock - 1))];
    // accumulator layout: element t = 64 r + l is row l/16 + 4r, column l%16
    const int row0 = (lane >> 4) + 4 * (t >> 6), col0 = lane & 15;
#pragma unroll
    for (int b = 0; b < kKzTiles; ++b) {
      const int I = kz_ti(b), J = kz_tj(b);
      if (J >= nb) continue;
      double sum = 0.0;
#pragma unroll
      for (int v = 0; v < kKzMaxWG; ++v)
        if (v < nwg) sum += g[b][v];
      const int row = 16 * I + row0, col = 16 * J + col0;
      Gs[row * kKzGs + col] = sum;
      if (I != J) Gs[col * kKzGs + row] = sum;
    }
    if (t < kKzMaxBlock) {
      double d = 0.0;
#pragma unroll
      for (int v = 0; v < kKzMaxWG; ++v)
        if (v < nwg) d += dd[v];
      // b~ - A~ x (kaczmarz.jl:305, mul!(vec_view, mat_view, x, -1.0, 1.0))
      rv[t] = t < s ? bt - d : 0.0;
    }
  }
  __syncthreads();
  if (wv != 0) return;

  // ---- one wavefront: lane i holds row i ----
  // Rows s .. 16 nb - 1 of the Gram matrix are zero; their diagonal is set
  // to 1, so the factor runs over whole 16-column blocks without branches
  // (the padding factors to the identity and leaves rows < s unchanged).
  const int i = lane;
  const int np = 16 * nb;
  if (i >= s && i < np) Gs[i * kKzGs + i] = 1.0;
  const double gii = i < s ? Gs[i * kKzGs + i] : 0.0;
  if (__ballot(gii != 0.0) == 0) {  // A~ == 0: no update
    if (i < s) p.y_out[i] = 0.0;
    if (i == 0) *p.ticket = 0u;
    return;
  }
  const double fl = 4.0 * s * 2.220446049250313e-16 * gii;  // pivot floor of row i
  double z = rv[i];
  double invme = 1.0;  // lane j: 1 / d_j
  bool okme = true;    // lane j: pivot j finite
  kz_ldl<true>(Gs, Dv, nb, lane, fl, z, invme, okme);
  const unsigned long long badm = __ballot(!okme && i < s);
  if (badm) {  // non-finite block: skipped, and recorded (info is sticky)
    if (i < s) p.y_out[i] = 0.0;
    if (i == 0) {
      *p.info = __ffsll(badm);  // 1-based first failing pivot
      *p.ticket = 0u;
    }
    return;
  }
  // L' y = D^-1 z; column j of L' is row j of L (read from LDS by lanes below j)
  z *= invme;
#pragma unroll
  for (int kb = 3; kb >= 0; --kb) {
    if (kb < nb) {
#pragma unroll
      for (int j = 16 * kb + 15; j >= 16 * kb; --j) {
        const double yj = kz_bcast(z, j);
        const double lji = Gs[j * kKzGs + i];
        if (i < j) z = fma(-lji, yj, z);
      }
    }
  }
  if (i < s) p.y_out[i] = z;
  if (i == 0) *p.ticket = 0u;
}

// ---- Kaczmarz over a chunk of pre-drawn row sets ------------------------------
//
// The row sets of a residual interval are drawn before its steps, and a
// block's Gram matrix does not depend on x.  So kz_prep_kernel forms, for
// every set of the chunk at once (one workgroup each), G = A~ A~' (FP64 MFMA
// on FP32 entries: exact products), its LDL' with the dependent-pivot rule of
// kz_ldl and the explicit M = L^-T D^+ L^-1, the operator with y = M r exactly
// the coefficients kz_step_kernel's solve gives.  kz_run_kernel then runs
// the chunk's steps in one workgroup: per step the 64 rows are gathered,
// r = b~ - A~ x (FP64 sums of exact products), y = M r and
// x += alpha A~' y, with nothing exchanged between workgroups.

struct KzPrep {
  const float* A;   // row-major rows (row i at A + i * ld)
  int64_t ld, n;
  const int* sets;  // cnt x s row indices
  int s;
  double* M;        // cnt x 64 x 64 (rows and columns past s zero)
  int* info;        // sticky: 1-based failing pivot of a non-finite block
};

template <bool VEC>
__global__ __launch_bounds__(256) void kz_prep_kernel(KzPrep p) {
  __shared__ __align__(16) float T[kKzMaxBlock * kKzTs];
  __shared__ double Gs[kKzMaxBlock * kKzGs];
  __shared__ double Dv[kKzMaxBlock];
  __shared__ double Iv[kKzMaxBlock];
  __shared__ int64_t crow[kKzMaxBlock];
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int s = p.s, nb = (s + 15) >> 4;
  const int* idx = p.sets + int64_t(blockIdx.x) * s;
  if (t < kKzMaxBlock) crow[t] = int64_t(idx[min(t, s - 1)]) * p.ld;
  __syncthreads();
  kz_f64x4 acc[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) acc[q] = kz_f64x4{0.0, 0.0, 0.0, 0.0};
  const int64_t nchunks = (p.n + kKzCols - 1) / kKzCols;
  // rows past s repeat row s - 1 and columns past n are clamped (in bounds);
  // both are zeroed when staged
  auto load = [&](int64_t ch, float(&lc)[16]) {
    const int64_t c0 = ch * kKzCols;
    if constexpr (VEC) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int e = t + 256 * k, row = e >> 4;
        const int64_t c = min<int64_t>(c0 + 4 * (e & 15), p.n - 4);
        const float4 w = *reinterpret_cast<const float4*>(p.A + crow[row] + c);
        lc[4 * k] = w.x, lc[4 * k + 1] = w.y, lc[4 * k + 2] = w.z, lc[4 * k + 3] = w.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int e = t + 256 * k, row = e >> 6;
        lc[k] = p.A[crow[row] + c0 + min<int64_t>(e & 63, p.n - c0 - 1)];
      }
    }
  };
  float buf[2][16];
  load(0, buf[0]);
  for (int64_t ch = 0; ch < nchunks; ++ch) {
    float(&lc)[16] = buf[ch & 1];
    if (ch + 1 < nchunks) load(ch + 1, buf[(ch + 1) & 1]);
    const int64_t c0 = ch * kKzCols;
    const int ncol = static_cast<int>(min<int64_t>(kKzCols, p.n - c0));
    if constexpr (VEC) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int e = t + 256 * k, row = e >> 4, cb = 4 * (e & 15);
        const bool rok = row < s;
        *reinterpret_cast<float4*>(&T[row * kKzTs + cb]) =
            float4{rok && cb < ncol ? lc[4 * k] : 0.f, rok && cb + 1 < ncol ? lc[4 * k + 1] : 0.f,
                   rok && cb + 2 < ncol ? lc[4 * k + 2] : 0.f,
                   rok && cb + 3 < ncol ? lc[4 * k + 3] : 0.f};
      }
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int e = t + 256 * k, row = e >> 6, cc = e & 63;
        T[row * kKzTs + cc] = (row < s && cc < ncol) ? lc[k] : 0.f;
      }
    }
    __syncthreads();
    float op[4][16];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int kq = 0; kq < 16; ++kq)
        op[rb][kq] = T[(16 * rb + (lane & 15)) * kKzTs + 4 * kq + (lane >> 4)];
    switch (wv) {
      case 0: kz_gram_tiles<0>(op, acc); break;
      case 1: kz_gram_tiles<1>(op, acc); break;
      case 2: kz_gram_tiles<2>(op, acc); break;
      default: kz_gram_tiles<3>(op, acc); break;
    }
    __syncthreads();  // T is rewritten by the next chunk
  }
  // tiles b = wv + 4q of the Gram matrix; accumulator element r of lane l is
  // row l/16 + 4r, column l%16 of the tile
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int b = wv + 4 * q;
    if (b >= kKzTiles) break;
    const int I = kz_ti(b), J = kz_tj(b);
    if (J >= nb) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * I + (lane >> 4) + 4 * r, col = 16 * J + (lane & 15);
      Gs[row * kKzGs + col] = acc[q][r];
      if (I != J) Gs[col * kKzGs + row] = acc[q][r];
    }
  }
  __syncthreads();
  if (wv != 0) return;

  // ---- one wavefront: the factor, then column `lane` of M ----
  const int i = lane;
  const int np = 16 * nb;
  double* Mo = p.M + int64_t(blockIdx.x) * (kKzMaxBlock * kKzMaxBlock);
  if (i >= s && i < np) Gs[i * kKzGs + i] = 1.0;
  const double gii = i < s ? Gs[i * kKzGs + i] : 0.0;
  bool zero = __ballot(gii != 0.0) == 0;  // A~ == 0: no update (kaczmarz.jl:300)
  if (!zero) {
    const double fl = 4.0 * s * 2.220446049250313e-16 * gii;
    double z = 0.0, invme = 1.0;
    bool okme = true;
    kz_ldl<false>(Gs, Dv, nb, lane, fl, z, invme, okme);
    const unsigned long long badm = __ballot(!okme && i < s);
    if (badm) {  // non-finite block: skipped, and recorded (info is sticky)
      if (i == 0) *p.info = __ffsll(badm);
      zero = true;
    } else {
      Iv[i] = invme;
      kz_lane_order();
    }
  }
  if (zero) {
#pragma unroll 8
    for (int r = 0; r < kKzMaxBlock; ++r) Mo[r * kKzMaxBlock + lane] = 0.0;
    return;
  }
  // v = L^-1 e_c (unit lower, forward), v <- D^+ v, v = L^-T v (backward);
  // L_ij is read by every lane from the same LDS word (a broadcast)
  double v[kKzMaxBlock];
#pragma unroll
  for (int j = 0; j < kKzMaxBlock; ++j) v[j] = j == lane ? 1.0 : 0.0;
#pragma unroll
  for (int j = 0; j < kKzMaxBlock; ++j) {
    if (j < np) {
#pragma unroll
      for (int r = j + 1; r < kKzMaxBlock; ++r)
        if (r < np) v[r] = fma(-Gs[r * kKzGs + j], v[j], v[r]);
    }
  }
#pragma unroll
  for (int j = 0; j < kKzMaxBlock; ++j) v[j] *= j < np ? Iv[j] : 0.0;
#pragma unroll
  for (int j = kKzMaxBlock - 1; j >= 0; --j) {
    if (j < np) {
#pragma unroll
      for (int r = 0; r < j; ++r) v[r] = fma(-Gs[j * kKzGs + r], v[j], v[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < kKzMaxBlock; ++r)
    Mo[r * kKzMaxBlock + lane] = (r < s && lane < s) ? v[r] : 0.0;
}

// Swap lanes 32..63 of a with lanes 0..31 of b (v_permlane32_swap), or lanes
// 16..31 / 48..63 of a with lanes 0..15 / 32..47 of b (v_permlane16_swap),
// for both halves of a double.  The builtins let the compiler insert the
// wait states these instructions need after a VALU write of their operands.
template <int ROW>
__device__ __forceinline__ void kz_swap(double& a, double& b) {
  const unsigned al = __double2loint(a), ah = __double2hiint(a);
  const unsigned bl = __double2loint(b), bh = __double2hiint(b);
  if constexpr (ROW == 32) {
    const auto lo = __builtin_amdgcn_permlane32_swap(al, bl, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(ah, bh, false, false);
    a = __hiloint2double(int(hi[0]), int(lo[0]));
    b = __hiloint2double(int(hi[1]), int(lo[1]));
  } else {
    const auto lo = __builtin_amdgcn_permlane16_swap(al, bl, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
    a = __hiloint2double(int(hi[0]), int(lo[0]));
    b = __hiloint2double(int(hi[1]), int(lo[1]));
  }
}

template <int CTRL>
__device__ __forceinline__ double kz_dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// One DPP stage of the transpose-reduce: values q and q + H are paired; lane l
// keeps the one selected by its bit SEL and adds the partner lane's copy
// (partner = l ^ mask of the DPP pattern, which flips bit SEL).
template <int H, int SEL, int CTRL>
__device__ __forceinline__ void kz_tr_stage(double (&v)[16], int lane) {
  const bool hi = (lane >> SEL) & 1;
#pragma unroll
  for (int q = 0; q < H; ++q) {
    const double keep = hi ? v[q + H] : v[q];
    const double send = hi ? v[q] : v[q + H];
    v[q] = keep + kz_dpp<CTRL>(send);
  }
}

struct KzRun {
  const float* A;   // row-major rows
  int64_t ld, n;
  const float* b;
  float* x;
  float alpha;
  const int* sets;  // cnt x s
  int s, cnt;
  const double* M;  // cnt x 64 x 64
};

constexpr int kKzRunThreads = 1024;
constexpr int kKzRunCols = 1024;  // columns one run workgroup holds

// The steps of one chunk in one workgroup (n <= 1024).  Thread t takes rows
// 16 g .. 16 g + 15 (g = t / 256) of the block and columns 4 c .. 4 c + 3
// (c = t % 256): one 16-byte load per row.  A~x: per row the lane's four
// products (exact in FP64), summed over the wavefront by a transpose-reduce
// that leaves row l / 4 of the group in lane l, then over the group's four
// wavefronts in a fixed order.  x += alpha A~' y: per column the lane's
// sixteen rows, then the four groups' partial sums in a fixed order; every
// group applies the same update to its own copy of x (identical bits).
template <bool VEC>
__global__ __launch_bounds__(kKzRunThreads) void kz_run_kernel(KzRun p) {
  constexpr int NW = kKzRunThreads / 64;
  __shared__ int64_t roff[2][kKzMaxBlock];
  __shared__ double bts[2][kKzMaxBlock];
  __shared__ double red[NW][16];
  __shared__ double ypart[NW][kKzMaxBlock];
  __shared__ double rr[kKzMaxBlock];
  __shared__ double yy[kKzMaxBlock];
  __shared__ double upd[4][kKzRunCols];
  __shared__ double Ms[2][kKzMaxBlock * kKzMaxBlock];
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int g = wv >> 2;  // row group
  const int s = p.s;
  const int64_t n = p.n;
  const int c0 = 4 * (t & 255);  // first column of this thread
  bool okc[4];
  float xq[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    okc[e] = c0 + e < n;
    xq[e] = okc[e] ? p.x[c0 + e] : 0.f;
  }
  // clamped columns (in bounds; x = 0 there and never written)
  const int64_t cq = VEC ? min<int64_t>(c0, n - 4) : 0;
  auto fetch = [&](int it, int bf) {
    if (t < kKzMaxBlock) {
      const int r = p.sets[int64_t(it) * s + min(t, s - 1)];
      roff[bf][t] = int64_t(r) * p.ld;
      bts[bf][t] = t < s ? static_cast<double>(p.b[r]) : 0.0;
    }
    const double* Mi = p.M + int64_t(it) * (kKzMaxBlock * kKzMaxBlock);
#pragma unroll
    for (int k = 0; k < kKzMaxBlock * kKzMaxBlock / kKzRunThreads; ++k)
      Ms[bf][t + kKzRunThreads * k] = Mi[t + kKzRunThreads * k];
  };
  fetch(0, 0);
  __syncthreads();
  for (int it = 0; it < p.cnt; ++it) {
    const int bf = it & 1;
    // rows 16 g + j of the block; rows past s repeat row s - 1 (their
    // residual entries are set to zero and M's rows past s are zero)
    float a[16][4];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int64_t ro = roff[bf][16 * g + j];  // the same in every lane
      const int64_t rou = (int64_t(__builtin_amdgcn_readfirstlane(int(ro >> 32))) << 32) |
                          uint32_t(__builtin_amdgcn_readfirstlane(int(ro)));
      const float* row = p.A + rou;
      if constexpr (VEC) {
        const float4 v = *reinterpret_cast<const float4*>(row + cq);
        a[j][0] = v.x, a[j][1] = v.y, a[j][2] = v.z, a[j][3] = v.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) a[j][e] = row[min<int64_t>(c0 + e, n - 1)];
      }
    }
    // the other buffer was last read before the previous step's barriers
    if (it + 1 < p.cnt) fetch(it + 1, bf ^ 1);
    double v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      double d = 0.0;
#pragma unroll
      for (int e = 0; e < 4; ++e) d = fma(static_cast<double>(a[j][e]), static_cast<double>(xq[e]), d);
      v[j] = d;
    }
    // the FP64 copies of the block are not kept for the update (that would
    // hold 128 registers); it converts the floats again
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) asm volatile("" : "+v"(a[j][e]));
    // transpose-reduce: row bits 3, 2 by lane bits 5, 4 (permlane swaps),
    // row bits 1, 0 by lane bits 3, 2 (DPP), then lane bits 1, 0 summed
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      kz_swap<32>(v[j], v[j + 8]);
      v[j] = v[j] + v[j + 8];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      kz_swap<16>(v[j], v[j + 4]);
      v[j] = v[j] + v[j + 4];
    }
    kz_tr_stage<2, 3, 0x128>(v, lane);  // row_ror:8, partner lane ^ 8
    kz_tr_stage<1, 2, 0x141>(v, lane);  // row_half_mirror, partner lane ^ 7
    v[0] = v[0] + kz_dpp<0x4E>(v[0]);   // quad_perm [2,3,0,1], lane ^ 2
    v[0] = v[0] + kz_dpp<0xB1>(v[0]);   // quad_perm [1,0,3,2], lane ^ 1
    if ((lane & 3) == 0) red[wv][lane >> 2] = v[0];
    __syncthreads();
    if (t < kKzMaxBlock) {
      const int w0 = 4 * (t >> 4), j = t & 15;
      const double ax = ((red[w0][j] + red[w0 + 1][j]) + red[w0 + 2][j]) + red[w0 + 3][j];
      rr[t] = t < s ? bts[bf][t] - ax : 0.0;
    }
    __syncthreads();
    // y = M r (M symmetric): thread t adds 4 terms of entry t % 64
    {
      const int q = t & 63, j0 = (t >> 6) * (kKzMaxBlock / NW);
      double y = 0.0;
#pragma unroll
      for (int j = 0; j < kKzMaxBlock / NW; ++j)
        y = fma(Ms[bf][(j0 + j) * kKzMaxBlock + q], rr[j0 + j], y);
      ypart[t >> 6][q] = y;
    }
    __syncthreads();
    if (t < kKzMaxBlock) {
      double y = ypart[0][t];
#pragma unroll
      for (int w = 1; w < NW; ++w) y += ypart[w][t];
      yy[t] = y;
    }
    __syncthreads();
    // x += alpha A~' y (kaczmarz.jl:308-309)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double u = 0.0;
#pragma unroll
      for (int j = 0; j < 16; ++j) u = fma(static_cast<double>(a[j][e]), yy[16 * g + j], u);
      upd[g][c0 + e] = u;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double u = ((upd[0][c0 + e] + upd[1][c0 + e]) + upd[2][c0 + e]) + upd[3][c0 + e];
      if (okc[e]) xq[e] = p.alpha * static_cast<float>(u) + xq[e];
    }
  }
  if (g == 0) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (okc[e]) p.x[c0 + e] = xq[e];
  }
}

__global__ void set_int_kernel(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ---- FP64 residual ||b - A x|| (full_residual.jl:32-41) ----------------------

constexpr int kResRowsPerWave = 4;

// A row-major (row i at A + i*ld): one wavefront per row, FP64 sums of exact
// FP32 products, a fixed shuffle tree; per-workgroup sums of r_i^2.
template <bool VEC>
__global__ __launch_bounds__(256) void residual_rows_kernel(int64_t m, int64_t n,
                                                            const float* __restrict__ A,
                                                            int64_t ld, const float* __restrict__ x,
                                                            const float* __restrict__ b,
                                                            double* __restrict__ part) {
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double r2 = 0.0;
  for (int q = 0; q < kResRowsPerWave; ++q) {
    const int64_t i = (int64_t(blockIdx.x) * 4 + wv) * kResRowsPerWave + q;
    if (i >= m) break;
    const float* row = A + i * ld;
    double acc = 0.0;
    if (VEC) {
      for (int64_t c = int64_t(lane) * 4; c < n; c += 256) {
        const float4 a = *reinterpret_cast<const float4*>(row + c);
        const float4 xv = *reinterpret_cast<const float4*>(x + c);
        acc += static_cast<double>(a.x) * xv.x;
        acc += static_cast<double>(a.y) * xv.y;
        acc += static_cast<double>(a.z) * xv.z;
        acc += static_cast<double>(a.w) * xv.w;
      }
    } else {
      for (int64_t c = lane; c < n; c += 64) acc += static_cast<double>(row[c]) * x[c];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    const double r = static_cast<double>(b[i]) - acc;
    r2 += r * r;
  }
  if (lane == 0) red[wv] = r2;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// A with any strides: one thread per row.
__global__ __launch_bounds__(256) void residual_part_kernel(int64_t m, int64_t n,
                                                            const float* __restrict__ A,
                                                            int64_t a_rs, int64_t a_cs,
                                                            const float* __restrict__ x,
                                                            const float* __restrict__ b,
                                                            double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  double r2 = 0.0;
  if (i < m) {
    double acc = 0.0;
    const float* row = A + i * a_rs;
    for (int64_t j = 0; j < n; ++j)
      acc += static_cast<double>(row[j * a_cs]) * static_cast<double>(x[j]);
    const double r = static_cast<double>(b[i]) - acc;
    r2 = r * r;
  }
  red[threadIdx.x] = r2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// sqrt of the sum of the partials: strided per-thread sums, then a fixed tree
__global__ __launch_bounds__(256) void residual_final_kernel(int64_t nparts,
                                                             const double* __restrict__ part,
                                                             double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t q = threadIdx.x; q < nparts; q += 256) s += part[q];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sqrt(red[0]);
}

// ---- host side ----------------------------------------------------------------

struct Bump2 {
  char* p;
  size_t left;
  size_t need = 0;
  template <class U>
  U* take(size_t n) {
    const size_t b = (n * sizeof(U) + 255) & ~size_t(255);
    need += b;
    if (!p || b > left) {
      left = 0;
      return nullptr;
    }
    U* r = reinterpret_cast<U*>(p);
    p += b;
    left -= b;
    return r;
  }
};

static int topk_chunks(int64_t n) { return static_cast<int>(ceil_div(n, kSortN)); }

// scratch for nsets top-k draws of k from n items
static size_t sample_scratch(int64_t n, int k, int64_t nsets, Bump2& b, double** k1, int** i1,
                             double** k2, int** i2, double** cdf) {
  const int64_t c1 = int64_t(topk_chunks(n)) * k;
  *k1 = b.take<double>(nsets * c1);
  *i1 = b.take<int>(nsets * c1);
  *k2 = b.take<double>(nsets * c1);
  *i2 = b.take<int>(nsets * c1);
  *cdf = b.take<double>(n);
  return b.need;
}

static int draw_sets(rla_ctx* ctx, uint64_t seed, uint64_t epoch0, int64_t n, const double* w,
                     int k, int replace, int64_t nsets, int* out, double* k1, int* i1,
                     double* k2, int* i2, double* cdf, bool cdf_ready) {
  if (replace) {
    if (!cdf_ready) {
      cdf_kernel<<<1, 1024, 0, ctx->stream>>>(n, w, cdf);
      RLA_LAUNCH_CHECK("cdf");
    }
    replace_sample_kernel<<<dim3(unsigned(nsets)), 1024, 0, ctx->stream>>>(seed, epoch0, n, cdf,
                                                                           k, out);
    RLA_LAUNCH_CHECK("replace_sample");
    return RLA_OK;
  }
  int64_t len = n;
  bool first = true;
  double* kin = nullptr;
  int* iin = nullptr;
  double* kout = k1;
  int* iout = i1;
  while (true) {
    const int chunks = topk_chunks(len);
    const bool fin = chunks == 1;
    topk_pass_kernel<<<dim3(unsigned(chunks), unsigned(nsets)), kSortThreads, 0, ctx->stream>>>(
        seed, epoch0, len, n, w, k, kin, iin, kout, iout, out, first ? 1 : 0, fin ? 1 : 0);
    RLA_LAUNCH_CHECK("topk_pass");
    if (fin) return RLA_OK;
    len = int64_t(chunks) * k;
    first = false;
    kin = kout;
    iin = iout;
    kout = (kout == k1) ? k2 : k1;
    iout = (iout == i1) ? i2 : i1;
  }
}

static int check_sample(int64_t n, int64_t k, int replace) {
  if (k < 1 || k > 1024) return fail(RLA_ENOTSUP, "sample size must be in 1..1024");
  if (!replace && k > n) return fail(RLA_EARG, "cannot draw more items than exist without replacement");
  if (n > INT32_MAX - 1) return fail(RLA_ENOTSUP, "more than 2^31-2 items");
  return RLA_OK;
}

// partial sums the residual writes for m rows
static int64_t residual_parts(int64_t m) { return ceil_div(m, 4 * kResRowsPerWave); }

static int residual(rla_ctx* ctx, int64_t m, int64_t n, const float* A, int64_t a_rs,
                    int64_t a_cs, const float* x, const float* b, double* part, double* out) {
  int64_t nb;
  if (a_cs == 1 && a_rs >= n) {  // row-major: one wavefront per row
    nb = ceil_div(m, 4 * kResRowsPerWave);
    const bool vec = n % 4 == 0 && a_rs % 4 == 0 &&
                     ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(x)) & 15) == 0;
    if (vec)
      residual_rows_kernel<true><<<dim3(unsigned(nb)), 256, 0, ctx->stream>>>(m, n, A, a_rs, x, b,
                                                                             part);
    else
      residual_rows_kernel<false><<<dim3(unsigned(nb)), 256, 0, ctx->stream>>>(m, n, A, a_rs, x,
                                                                              b, part);
  } else {
    nb = ceil_div(m, 256);
    residual_part_kernel<<<dim3(unsigned(nb)), 256, 0, ctx->stream>>>(m, n, A, a_rs, a_cs, x, b,
                                                                      part);
  }
  RLA_LAUNCH_CHECK("residual");
  residual_final_kernel<<<1, 256, 0, ctx->stream>>>(nb, part, out);
  RLA_LAUNCH_CHECK("residual/final");
  return RLA_OK;
}

struct KzBufs {
  int* sets;          // every x s row indices
  double* part_G;     // W x kKzTiles accumulator tiles
  double* part_d;     // W x kKzMaxBlock
  double* y;          // 2 x kKzMaxBlock (the pending block's and the new block's)
  unsigned* ticket;   // arrival counter of kz_step_kernel
  int* info;          // sticky breakdown flag
  double* res;        // residual partials
  double* resv;       // residual norm
  double *k1, *k2, *cdf;
  int *i1, *i2;
  double* M;          // kz_prep_kernel output: up to kKzPrepMax x 64 x 64
  int* sets2;         // second row-set buffer (prepared interval ahead)
  double* M2;         // second operator buffer
};

constexpr int64_t kKzPrepMax = 256;  // row sets prepared per kz_prep / kz_run pair

// workgroups of a step over n columns, and the 64-column chunks each sweeps
static void kz_grid(int64_t n, int* wgs, int* chunks_per_wg) {
  const int64_t chunks = ceil_div(n, kKzCols);
  const int64_t per = ceil_div(chunks, kKzMaxWG);
  *chunks_per_wg = static_cast<int>(per);
  *wgs = static_cast<int>(ceil_div(chunks, per));
}

static size_t kz_layout(int64_t m, int64_t n, int64_t s, int64_t every, void* work, size_t bytes,
                        KzBufs* out) {
  Bump2 b{static_cast<char*>(work), bytes};
  KzBufs k{};
  int W, per;
  kz_grid(n, &W, &per);
  k.sets = b.take<int>(every * s);
  k.part_G = b.take<double>(int64_t(W) * kKzTiles * kKzTileDoubles);
  k.part_d = b.take<double>(int64_t(W) * kKzMaxBlock);
  k.y = b.take<double>(2 * kKzMaxBlock);
  k.ticket = b.take<unsigned>(1);
  k.info = b.take<int>(1);
  k.res = b.take<double>(std::max<int64_t>(residual_parts(m), ceil_div(m, 256)));
  k.resv = b.take<double>(1);
  sample_scratch(m, static_cast<int>(s), every, b, &k.k1, &k.i1, &k.k2, &k.i2, &k.cdf);
  if (n <= kKzRunCols) {
    k.M = b.take<double>(std::min(every, kKzPrepMax) * kKzMaxBlock * kKzMaxBlock);
    if (every <= kKzPrepMax) {
      k.sets2 = b.take<int>(every * s);
      k.M2 = b.take<double>(every * kKzMaxBlock * kKzMaxBlock);
    }
  }
  if (out) *out = k;
  return b.need;
}

// one Kaczmarz step: the pending update (prev, y_prev) and the new block's
// coefficients (cur, bsrc -> y_out)
static int kz_step(rla_ctx* ctx, const KzBufs& k, int64_t n, const RowSrc& prev,
                   const double* y_prev, float alpha, float* x, const RowSrc& cur,
                   const RowSrc& bsrc, double* y_out) {
  KzStep st;
  int W;
  kz_grid(n, &W, &st.chunks_per_wg);
  st.n = n;
  st.prev = prev;
  st.y_prev = y_prev;
  st.alpha = alpha;
  st.x = x;
  st.cur = cur;
  st.bsrc = bsrc;
  st.part_G = k.part_G;
  st.part_d = k.part_d;
  st.ticket = k.ticket;
  st.y_out = y_out;
  st.info = k.info;
  auto vec_ok = [&](const RowSrc& r) {
    return r.s == 0 || (r.cs == 1 && r.rs % 4 == 0 && (reinterpret_cast<uintptr_t>(r.base) & 15) == 0);
  };
  if (n % 4 == 0 && vec_ok(prev) && vec_ok(cur))
    kz_step_kernel<true><<<dim3(unsigned(W)), 256, 0, ctx->stream>>>(st);
  else
    kz_step_kernel<false><<<dim3(unsigned(W)), 256, 0, ctx->stream>>>(st);
  RLA_LAUNCH_CHECK("kz_step");
  return RLA_OK;
}

// kz_solve_chunked with the next interval's draws and operators prepared on
// a second stream while the current interval runs (intervals of at most
// kKzPrepMax steps): the draws and kz_prep_kernel depend only on the counter
// stream and A, and kz_run_kernel occupies one CU.  Two buffer pairs
// alternate; a buffer is refilled only after the run that read it.
static int kz_solve_overlapped(rla_ctx* ctx, KzBufs& k, int64_t m, int64_t n, const float* A_rm,
                               int64_t ld_rm, const float* b, float* x, int s, float alpha,
                               const double* w, int replace, uint64_t seed, uint64_t epoch0,
                               int64_t max_it, int64_t every, double threshold, double* hist,
                               int64_t hist_cap, int64_t* n_hist, int64_t* iterations,
                               int* converged) {
  struct Side {
    hipStream_t st = nullptr;
    hipEvent_t prep[2] = {nullptr, nullptr}, run[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    ~Side() {
      if (st) (void)hipStreamSynchronize(st);
      for (int q = 0; q < 2; ++q) {
        if (prep[q]) (void)hipEventDestroy(prep[q]);
        if (run[q]) (void)hipEventDestroy(run[q]);
      }
      if (st) (void)hipStreamDestroy(st);
    }
  } side;
  RLA_HIP(hipStreamCreateWithFlags(&side.st, hipStreamNonBlocking));
  for (int q = 0; q < 2; ++q) {
    RLA_HIP(hipEventCreateWithFlags(&side.prep[q], hipEventDisableTiming));
    RLA_HIP(hipEventCreateWithFlags(&side.run[q], hipEventDisableTiming));
  }
  int* sets[2] = {k.sets, k.sets2};
  double* Ms[2] = {k.M, k.M2};
  const bool vec = n % 4 == 0 && ld_rm % 4 == 0 && (reinterpret_cast<uintptr_t>(A_rm) & 15) == 0;
  bool cdf_ready = false;
  set_int_kernel<<<1, 1, 0, ctx->stream>>>(k.info, 0);
  RLA_LAUNCH_CHECK("set_int");
  // the side stream starts after the work already queued (x, info)
  RLA_HIP(hipEventRecord(side.run[1], ctx->stream));
  RLA_HIP(hipStreamWaitEvent(side.st, side.run[1], 0));
  // row sets of iterations ic .. ic + every - 1 and the operators of the
  // first cnt of them, into buffer q, on the side stream
  auto prepare = [&](int64_t ic, int64_t cnt, int q) -> int {
    if (side.used[q]) RLA_HIP(hipStreamWaitEvent(side.st, side.run[q], 0));
    hipStream_t main_st = ctx->stream;
    ctx->stream = side.st;
    int r = draw_sets(ctx, seed, epoch0 + uint64_t(ic - 1), m, w, s, replace, every, sets[q],
                      k.k1, k.i1, k.k2, k.i2, k.cdf, cdf_ready);
    ctx->stream = main_st;
    if (r) return r;
    cdf_ready = true;
    KzPrep pp{A_rm, ld_rm, n, sets[q], s, Ms[q], k.info};
    if (vec)
      kz_prep_kernel<true><<<dim3(unsigned(cnt)), 256, 0, side.st>>>(pp);
    else
      kz_prep_kernel<false><<<dim3(unsigned(cnt)), 256, 0, side.st>>>(pp);
    RLA_LAUNCH_CHECK("kz_prep");
    RLA_HIP(hipEventRecord(side.prep[q], side.st));
    return RLA_OK;
  };
  int64_t nh = 0, it = 0;
  int conv = 0;
  int rc;
  int q = 0;
  if (max_it > 1) {
    rc = prepare(1, std::min(every, max_it - 1), 0);
    if (rc) return rc;
  }
  for (int64_t i = 1; max_it >= 1;) {
    it = i;
    rc = residual(ctx, m, n, A_rm, ld_rm, 1, x, b, k.res, k.resv);
    if (rc) return rc;
    double err = 0.0;
    int info = 0;
    RLA_HIP(hipMemcpyAsync(&err, k.resv, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RLA_HIP(hipMemcpyAsync(&info, k.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    RLA_HIP(hipStreamSynchronize(ctx->stream));
    if (info > 0)
      return fail(RLA_EDOMAIN, "block Gram matrix is not finite (NaN or Inf in the block or x)");
    if (hist && nh < hist_cap) hist[nh] = err;
    ++nh;
    conv = (i < max_it) ? (err < threshold) : 1;  // update_logger!, basic_logger.jl:93-98
    if (conv) break;
    const int64_t cnt = std::min(every, max_it - i);
    // the run is queued first, so that its one workgroup has a CU before the
    // next interval's draws fill the device
    RLA_HIP(hipStreamWaitEvent(ctx->stream, side.prep[q], 0));
    KzRun pr{A_rm, ld_rm, n, b, x, alpha, sets[q], s, static_cast<int>(cnt), Ms[q]};
    if (vec)
      kz_run_kernel<true><<<1, kKzRunThreads, 0, ctx->stream>>>(pr);
    else
      kz_run_kernel<false><<<1, kKzRunThreads, 0, ctx->stream>>>(pr);
    RLA_LAUNCH_CHECK("kz_run");
    RLA_HIP(hipEventRecord(side.run[q], ctx->stream));
    side.used[q] = true;
    if (i + cnt < max_it) {  // the next interval, prepared while this one runs
      rc = prepare(i + cnt, std::min(every, max_it - i - cnt), q ^ 1);
      if (rc) return rc;
    }
    q ^= 1;
    i += cnt;
  }
  RLA_HIP(hipStreamSynchronize(side.st));
  int info = 0;
  RLA_HIP(hipMemcpyAsync(&info, k.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (info > 0) return fail(RLA_EDOMAIN, "block Gram matrix is not finite (NaN or Inf in the block or x)");
  if (n_hist) *n_hist = nh;
  if (iterations) *iterations = it;
  if (converged) *converged = conv;
  return RLA_OK;
}

// rla_kaczmarz_solve_f32 for n <= 1024: per residual interval, the row sets
// are drawn, kz_prep_kernel forms the operators M of up to kKzPrepMax sets at
// once and kz_run_kernel applies their steps in one workgroup.  The residual
// checks, draws and iteration count are those of the per-step loop.
static int kz_solve_chunked(rla_ctx* ctx, KzBufs& k, int64_t m, int64_t n, const float* A_rm,
                            int64_t ld_rm, const float* b, float* x, int s, float alpha,
                            const double* w, int replace, uint64_t seed, uint64_t epoch0,
                            int64_t max_it, int64_t every, double threshold, double* hist,
                            int64_t hist_cap, int64_t* n_hist, int64_t* iterations,
                            int* converged) {
  set_int_kernel<<<1, 1, 0, ctx->stream>>>(k.info, 0);
  RLA_LAUNCH_CHECK("set_int");
  const bool vec = n % 4 == 0 && ld_rm % 4 == 0 && (reinterpret_cast<uintptr_t>(A_rm) & 15) == 0;
  int64_t nh = 0, it = 0;
  int conv = 0;
  bool cdf_ready = false;
  int rc;
  for (int64_t i = 1; max_it >= 1;) {
    // error and logger at i (i = 1, 1 + every, ... and max_it)
    it = i;
    rc = residual(ctx, m, n, A_rm, ld_rm, 1, x, b, k.res, k.resv);
    if (rc) return rc;
    double err = 0.0;
    int info = 0;
    RLA_HIP(hipMemcpyAsync(&err, k.resv, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RLA_HIP(hipMemcpyAsync(&info, k.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    RLA_HIP(hipStreamSynchronize(ctx->stream));
    if (info > 0)
      return fail(RLA_EDOMAIN, "block Gram matrix is not finite (NaN or Inf in the block or x)");
    if (hist && nh < hist_cap) hist[nh] = err;
    ++nh;
    conv = (i < max_it) ? (err < threshold) : 1;  // update_logger!, basic_logger.jl:93-98
    if (conv) break;
    // row sets of iterations i .. i + every - 1 (epochs epoch0 + i - 1 ...)
    rc = draw_sets(ctx, seed, epoch0 + uint64_t(i - 1), m, w, s, replace, every, k.sets, k.k1,
                   k.i1, k.k2, k.i2, k.cdf, cdf_ready);
    if (rc) return rc;
    cdf_ready = true;
    const int64_t cnt = std::min(every, max_it - i);  // steps before the next check
    for (int64_t o = 0; o < cnt; o += kKzPrepMax) {
      const int c = static_cast<int>(std::min(kKzPrepMax, cnt - o));
      KzPrep pp{A_rm, ld_rm, n, k.sets + o * s, s, k.M, k.info};
      if (vec)
        kz_prep_kernel<true><<<dim3(unsigned(c)), 256, 0, ctx->stream>>>(pp);
      else
        kz_prep_kernel<false><<<dim3(unsigned(c)), 256, 0, ctx->stream>>>(pp);
      RLA_LAUNCH_CHECK("kz_prep");
      KzRun pr{A_rm, ld_rm, n, b, x, alpha, k.sets + o * s, s, c, k.M};
      if (vec)
        kz_run_kernel<true><<<1, kKzRunThreads, 0, ctx->stream>>>(pr);
      else
        kz_run_kernel<false><<<1, kKzRunThreads, 0, ctx->stream>>>(pr);
      RLA_LAUNCH_CHECK("kz_run");
    }
    i += cnt;
  }
  int info = 0;
  RLA_HIP(hipMemcpyAsync(&info, k.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (info > 0) return fail(RLA_EDOMAIN, "block Gram matrix is not finite (NaN or Inf in the block or x)");
  if (n_hist) *n_hist = nh;
  if (iterations) *iterations = it;
  if (converged) *converged = conv;
  return RLA_OK;
}

}  // namespace rla

using namespace rla;

extern "C" {

int rla_row_norms2_f32(rla_ctx* ctx, int64_t m, int64_t n, const float* A, int64_t a_rs,
                       int64_t a_cs, double* w) {
  RLA_CHECK_CTX(ctx);
  if (m < 0 || n < 0) return fail(RLA_EARG, "negative dimension");
  int rc = use_device(ctx);
  if (rc) return rc;
  if (m == 0) return RLA_OK;
  row_norms2_kernel<float><<<dim3(unsigned(ceil_div(m, 256))), 256, 0, ctx->stream>>>(
      m, n, A, a_rs, a_cs, w);
  RLA_LAUNCH_CHECK("row_norms2");
  return RLA_OK;
}

int rla_row_norms2_f64(rla_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t a_rs,
                       int64_t a_cs, double* w) {
  RLA_CHECK_CTX(ctx);
  if (m < 0 || n < 0) return fail(RLA_EARG, "negative dimension");
  int rc = use_device(ctx);
  if (rc) return rc;
  if (m == 0) return RLA_OK;
  row_norms2_kernel<double><<<dim3(unsigned(ceil_div(m, 256))), 256, 0, ctx->stream>>>(
      m, n, A, a_rs, a_cs, w);
  RLA_LAUNCH_CHECK("row_norms2");
  return RLA_OK;
}

int rla_wsample(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n, const double* w,
                int64_t k, int replace, int64_t nsets, int32_t* sets) {
  RLA_CHECK_CTX(ctx);
  int rc = check_sample(n, k, replace);
  if (rc) return rc;
  if (nsets < 1 || nsets > 65535) return fail(RLA_EARG, "nsets must be in 1..65535");
  rc = use_device(ctx);
  if (rc) return rc;
  Bump2 probe{nullptr, 0};
  double *k1, *k2, *cdf;
  int *i1, *i2;
  const size_t need = sample_scratch(n, static_cast<int>(k), nsets, probe, &k1, &i1, &k2, &i2, &cdf);
  void* ws = nullptr;
  rc = workspace(ctx, need, &ws);
  if (rc) return rc;
  Bump2 b{static_cast<char*>(ws), need};
  sample_scratch(n, static_cast<int>(k), nsets, b, &k1, &i1, &k2, &i2, &cdf);
  return draw_sets(ctx, seed, epoch, n, w, static_cast<int>(k), replace, nsets, sets, k1, i1, k2,
                   i2, cdf, false);
}

int rla_residual_norm_f32(rla_ctx* ctx, int64_t m, int64_t n, const float* A, int64_t a_rs,
                          int64_t a_cs, const float* x, const float* b, double* out_host) {
  RLA_CHECK_CTX(ctx);
  if (m < 1 || n < 0 || !out_host) return fail(RLA_EARG, "bad residual arguments");
  int rc = use_device(ctx);
  if (rc) return rc;
  void* ws = nullptr;
  const int64_t np = std::max<int64_t>(residual_parts(m), ceil_div(m, 256));
  rc = workspace(ctx, sizeof(double) * (np + 1), &ws);
  if (rc) return rc;
  double* part = static_cast<double*>(ws);
  double* res = part + np;
  rc = residual(ctx, m, n, A, a_rs, a_cs, x, b, part, res);
  if (rc) return rc;
  RLA_HIP(hipMemcpyAsync(out_host, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  return RLA_OK;
}

int rla_kaczmarz_workspace_bytes(int64_t m, int64_t n, int64_t s, int64_t every, size_t* bytes) {
  if (!bytes || m < 1 || n < 1 || s < 1 || every < 1)
    return fail(RLA_EARG, "rla_kaczmarz_workspace_bytes: bad arguments");
  *bytes = kz_layout(m, n, s, every, nullptr, 0, nullptr);
  return RLA_OK;
}

// One block update x += alpha * At' (At At')^-1 (bt - At x) with an explicit
// s x n block At (strided) and bt (kaczmarz_update_block!, kaczmarz.jl:294-311;
// s = 1 is kaczmarz_update!, :262-276).  Synchronises to report a breakdown.
int rla_kaczmarz_block_update_f32(rla_ctx* ctx, int64_t s, int64_t n, const float* At,
                                  int64_t at_rs, int64_t at_cs, const float* bt, int64_t bt_inc,
                                  float* x, double alpha, void* work, size_t work_bytes) {
  RLA_CHECK_CTX(ctx);
  if (s < 1 || n < 1) return fail(RLA_EARG, "empty Kaczmarz block");
  if (s > kKzMaxBlock) return fail(RLA_ENOTSUP, "Kaczmarz blocks are limited to 64 rows");
  int rc = use_device(ctx);
  if (rc) return rc;
  KzBufs k;
  if (kz_layout(s, n, s, 1, work, work_bytes, &k) > work_bytes)
    return fail(RLA_EARG, "Kaczmarz workspace too small");
  const int si = static_cast<int>(s);
  const RowSrc none{nullptr, 0, 0, nullptr, 0};
  const RowSrc cur{At, at_rs, at_cs, nullptr, si};
  const RowSrc bsrc{bt, bt_inc, 0, nullptr, si};
  set_int_kernel<<<1, 1, 0, ctx->stream>>>(k.info, 0);
  RLA_LAUNCH_CHECK("set_int");
  set_int_kernel<<<1, 1, 0, ctx->stream>>>(reinterpret_cast<int*>(k.ticket), 0);
  RLA_LAUNCH_CHECK("set_int");
  rc = kz_step(ctx, k, n, none, k.y, 0.f, x, cur, bsrc, k.y);
  if (rc) return rc;
  rc = kz_step(ctx, k, n, cur, k.y, static_cast<float>(alpha), x, none, none, nullptr);
  if (rc) return rc;
  int info = 0;
  RLA_HIP(hipMemcpyAsync(&info, k.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (info > 0) return fail(RLA_EDOMAIN, "block Gram matrix is not finite (NaN or Inf in the block or x)");
  return RLA_OK;
}

// rsolve! of a Kaczmarz recipe with a Sampling compressor (kaczmarz.jl:313-368):
// for i = 1..max_it: error and logger at i = 1, 1+every, ... and at max_it
// (threshold_stop: converged when error < threshold; i == max_it always
// stops), then a new row set and a block update.  A_rm (the row-major copy of
// A, row stride ld_rm) is read for the blocks and the FP64 residual; A
// (column-major, lda) is only checked.  w: row weights (NULL: uniform).  hist
// receives the evaluated errors.
int rla_kaczmarz_solve_f32(rla_ctx* ctx, int64_t m, int64_t n, const float* A, int64_t lda,
                           const float* A_rm, int64_t ld_rm, const float* b, float* x, int64_t s,
                           double alpha, const double* w, int replace, uint64_t seed,
                           uint64_t epoch0, int64_t max_it, int64_t every, double threshold,
                           double* hist, int64_t hist_cap, int64_t* n_hist, int64_t* iterations,
                           int* converged, void* work, size_t work_bytes) {
  RLA_CHECK_CTX(ctx);
  if (m < 1 || n < 1 || s < 1 || every < 1 || max_it < 0 || lda < m || ld_rm < n)
    return fail(RLA_EARG, "rla_kaczmarz_solve_f32: bad arguments");
  if (every > 4096) return fail(RLA_EARG, "residual interval above 4096 iterations");
  if (s > kKzMaxBlock) return fail(RLA_ENOTSUP, "Kaczmarz blocks are limited to 64 rows");
  int rc = check_sample(m, s, replace);
  if (rc) return rc;
  rc = use_device(ctx);
  if (rc) return rc;
  KzBufs k;
  if (kz_layout(m, n, s, every, work, work_bytes, &k) > work_bytes)
    return fail(RLA_EARG, "Kaczmarz workspace too small");
  const int si = static_cast<int>(s);
  if (n <= kKzRunCols && every <= kKzPrepMax)
    return kz_solve_overlapped(ctx, k, m, n, A_rm, ld_rm, b, x, si, static_cast<float>(alpha), w,
                               replace, seed, epoch0, max_it, every, threshold, hist, hist_cap,
                               n_hist, iterations, converged);
  if (n <= kKzRunCols)
    return kz_solve_chunked(ctx, k, m, n, A_rm, ld_rm, b, x, si, static_cast<float>(alpha), w,
                            replace, seed, epoch0, max_it, every, threshold, hist, hist_cap,
                            n_hist, iterations, converged);
  const RowSrc none{nullptr, 0, 0, nullptr, 0};
  const float al = static_cast<float>(alpha);
  RowSrc pending = none;
  int yb = 0;  // k.y + yb * kKzMaxBlock holds the pending block's coefficients
  bool cdf_ready = false;
  set_int_kernel<<<1, 1, 0, ctx->stream>>>(k.info, 0);
  RLA_LAUNCH_CHECK("set_int");
  set_int_kernel<<<1, 1, 0, ctx->stream>>>(reinterpret_cast<int*>(k.ticket), 0);
  RLA_LAUNCH_CHECK("set_int");
  int64_t nh = 0, it = 0;
  int conv = 0;
  int64_t set_base = 1;  // iteration whose set is sets[0]
  for (int64_t i = 1; i <= max_it; ++i) {
    it = i;
    if ((i - 1) % every == 0 || i == max_it) {
      if (pending.s) {
        rc = kz_step(ctx, k, n, pending, k.y + yb * kKzMaxBlock, al, x, none, none, nullptr);
        if (rc) return rc;
        pending = none;
      }
      rc = residual(ctx, m, n, A_rm, ld_rm, 1, x, b, k.res, k.resv);
      if (rc) return rc;
      double err = 0.0;
      int info = 0;
      RLA_HIP(hipMemcpyAsync(&err, k.resv, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      RLA_HIP(hipMemcpyAsync(&info, k.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      RLA_HIP(hipStreamSynchronize(ctx->stream));
      if (info > 0)
        return fail(RLA_EDOMAIN, "block Gram matrix is not finite (NaN or Inf in the block or x)");
      if (hist && nh < hist_cap) hist[nh] = err;
      ++nh;
      conv = (i < max_it) ? (err < threshold) : 1;  // update_logger!, basic_logger.jl:93-98
      if (conv) break;
    }
    if ((i - 1) % every == 0) {
      // row sets for iterations i .. i+every-1 (epochs epoch0+i-1 ...).  The
      // set buffer is reused: the pending update of the previous chunk's last
      // set was applied just above, under the same condition.
      rc = draw_sets(ctx, seed, epoch0 + uint64_t(i - 1), m, w, si, replace, every, k.sets, k.k1,
                     k.i1, k.k2, k.i2, k.cdf, cdf_ready);
      if (rc) return rc;
      cdf_ready = true;
      set_base = i;
    }
    const RowSrc cur{A_rm, ld_rm, 1, k.sets + (i - set_base) * s, si};
    const RowSrc bsrc{b, 1, 0, cur.idx, si};
    rc = kz_step(ctx, k, n, pending, k.y + yb * kKzMaxBlock, al, x, cur, bsrc,
                 k.y + (yb ^ 1) * kKzMaxBlock);
    if (rc) return rc;
    yb ^= 1;
    pending = cur;
  }
  if (pending.s) {
    rc = kz_step(ctx, k, n, pending, k.y + yb * kKzMaxBlock, al, x, none, none, nullptr);
    if (rc) return rc;
  }
  int info = 0;
  RLA_HIP(hipMemcpyAsync(&info, k.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (info > 0) return fail(RLA_EDOMAIN, "block Gram matrix is not finite (NaN or Inf in the block or x)");
  if (n_hist) *n_hist = nh;
  if (iterations) *iterations = it;
  if (converged) *converged = conv;
  return RLA_OK;
}

}  // extern "C"
