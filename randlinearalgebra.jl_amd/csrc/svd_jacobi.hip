// FP64 SVD of a small square matrix (order k <= 288) by one-sided Jacobi,
// the `svd(QA)` step of RandSVD (randsvd.jl:183-187) after the range's
// factor has been reduced to k x k.
//
// One-sided (Hestenes) Jacobi: plane rotations of column pairs of M until
// every pair is orthogonal to tolerance; then M V = U diag(S) with
// S_j = ||M(:, j)||, U(:, j) = M(:, j) / S_j and V the product of the
// rotations.  Each rotation makes columns p, q orthogonal: with
// a = |m_p|^2, b = |m_q|^2, g = m_p . m_q (FP64 sums),
//   zeta = (b - a) / (2 g), t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
//   c = 1 / sqrt(1 + t^2), s = c t,
//   m_p <- c m_p - s m_q,  m_q <- s m_p + c m_q   (and likewise V),
// skipped when |g| <= tol sqrt(a b), tol = sqrt(k) eps.
//
// Layout: the columns are cut into blocks of kJB = 16 (zero columns pad the
// last block; a zero column is never rotated).  One cooperative launch holds
// nblk / 2 workgroups of kJB / 4 wavefronts, and M and V in global memory
// (L2).  (Blocks of 8 give twice the workgroups and half the LDS traffic per
// round, but twice the grid barriers per sweep: slower at k = 266.)
// A column pair is worked by one 16-lane DPP row: lane i holds rows
// 2i, 2i + 1, 2i + 32, 2i + 33, ..., so a wavefront works 4 pairs at once,
// moves columns through LDS with 16-byte accesses, and its sums are four DPP
// steps within the row.  A sweep visits every column pair once:
//   * an intra pass: workgroup w loads blocks 2w, 2w+1 into LDS and runs the
//     round-robin schedule over each block's kJB columns (kJB - 1 rounds of
//     kJB / 2 disjoint pairs per block, kJB pairs per round on the kJB rows);
//   * nblk - 1 passes of the round-robin schedule over blocks: workgroup w
//     takes its blocks I, J; row u (wavefront u / 4, row u % 4) keeps
//     column u of I (and of V) in registers for the whole pass and meets
//     column (u + r) mod kJB of J, read from and written back to LDS, in
//     round r (kJB rounds).
// A grid barrier follows every pass.  The sweep's rotation count decides,
// after a grid barrier, whether another sweep runs, so every workgroup
// executes the same barriers.  At most 60 sweeps.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <mutex>

#include "rla_internal.hpp"

namespace rla {

namespace {

namespace cg = cooperative_groups;

constexpr int kJB = 16;            // columns per block (a multiple of 4)
constexpr int kJMaxK = 288;
constexpr int kJT = kJMaxK / 16;   // rows per lane of a 16-lane row
constexpr int kJLd = 304;          // LDS column stride (doubles): the 4 rows of a
                                   // wavefront reading 4 columns fill two bank passes
constexpr int kJThreads = kJB * 16;  // one 16-lane row per pair: kJB pairs per round
constexpr int kJMaxSweeps = 60;

struct JacobiArgs {
  int k;           // rows of M and V, and real columns
  int nblk;        // column blocks (even)
  double* M;       // k x (nblk * kJB), ld k
  double* V;       // k x (nblk * kJB), ld k
  int* rot;        // kJMaxSweeps counters, zero on entry
  int* sweeps;     // out: sweeps run
  double tol;
};

// partner of slot i in round r of the circle schedule over n (even) items:
// slot 0 stays, the other n - 1 rotate; pairs are (pos 0, pos n-1),
// (pos 1, pos n-2), ...
__device__ __forceinline__ int circle_item(int pos, int r, int n) {
  if (pos == 0) return 0;
  return 1 + (pos - 1 + r) % (n - 1);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// row of M held by entry q of lane li of a 16-lane row: rows 2 li, 2 li + 1,
// then the same 32 rows further down, so LDS moves are 16-byte accesses
__device__ __forceinline__ int jrow(int li, int q) { return 2 * li + (q & 1) + 32 * (q >> 1); }

// sum over the 16 lanes of a DPP row, in every lane of the row: quad swaps,
// then the half-row and row mirrors
__device__ __forceinline__ double row_sum16(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);  // row_half_mirror
  v += dpp_f64<0x140>(v);  // row_mirror
  return v;
}

// 1/x and 1/sqrt(x) from the hardware estimates and two Newton steps
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

// the rotation (c, s) of a pair from its sums; false (c = 1, s = 0) when the
// pair is orthogonal to tolerance or has a zero column
__device__ __forceinline__ bool jacobi_rotation(double a, double b, double g, double tol,
                                                double& c, double& tn) {
  c = 1.0;
  tn = 0.0;
  // |g| > tol sqrt(a b), squared and divided by a^2 so nothing leaves the
  // exponent range: (g / a)^2 > tol^2 (b / a)
  if (!(a > 0.0 && b > 0.0)) return false;
  const double ra = rcp_nr(a), gr = g * ra;
  if (!(gr * gr > tol * tol * (b * ra))) return false;
  const double zeta = (b - a) * rcp_nr(2.0 * g);
  const double az = fabs(zeta);
  double tt;
  if (az > 1e150) {
    tt = 0.5 * rcp_nr(az);
  } else {
    const double w = fma(zeta, zeta, 1.0);
    tt = rcp_nr(az + w * rsq_nr(w));
  }
  tn = zeta < 0.0 ? -tt : tt;
  c = rsq_nr(fma(tn, tn, 1.0));
  return true;
}

// Columns x and y of one pair (a lane's 18 rows) are held scaled: the true
// columns are dx * x and dy * y (likewise for V).  A rotation by the angle
// with tangent t and cosine c becomes, on the stored columns,
//   x <- x - t (dy / dx) y,  y <- y + t (dx / dy) x,  dx <- c dx,  dy <- c dy
// (one FMA per entry; the scales are folded back into the columns at the end
// of every pass, after at most 16 rotations, so they stay above 2^-8).
// Every lane of the wavefront calls these (the rows of a wavefront work 4
// different pairs).

// the pair's rotation from its sums (c = 1, t = 0 and false when none)
__device__ __forceinline__ bool pair_rotation(const double (&x)[kJT], const double (&y)[kJT],
                                              double dx, double dy, double tol, double& c,
                                              double& tn) {
  double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0, g0 = 0.0, g1 = 0.0;
#pragma unroll
  for (int t = 0; t < kJT; t += 2) {
    a0 = fma(x[t], x[t], a0);
    b0 = fma(y[t], y[t], b0);
    g0 = fma(x[t], y[t], g0);
    a1 = fma(x[t + 1], x[t + 1], a1);
    b1 = fma(y[t + 1], y[t + 1], b1);
    g1 = fma(x[t + 1], y[t + 1], g1);
  }
  const double a = row_sum16(a0 + a1) * (dx * dx);
  const double b = row_sum16(b0 + b1) * (dy * dy);
  const double g = row_sum16(g0 + g1) * (dx * dy);
  return jacobi_rotation(a, b, g, tol, c, tn);
}

// applies the rotation to the stored M and V columns of the pair
__device__ __forceinline__ void apply_rotation(double (&x)[kJT], double (&vx)[kJT],
                                               double (&y)[kJT], double (&vy)[kJT], double& dx,
                                               double& dy, double c, double tn) {
  const double tp = tn * dy * rcp_nr(dx), tq = tn * dx * rcp_nr(dy);
#pragma unroll
  for (int t = 0; t < kJT; ++t) {
    const double xp = x[t], u = vx[t];
    x[t] = fma(-tp, y[t], xp);
    y[t] = fma(tq, xp, y[t]);
    vx[t] = fma(-tp, vy[t], u);
    vy[t] = fma(tq, u, vy[t]);
  }
  dx *= c;
  dy *= c;
}

__global__ __launch_bounds__(kJThreads) void svd_jacobi_kernel(JacobiArgs g) {
  extern __shared__ __align__(16) double jsm[];
  const int k = g.k;
  double* Ms = jsm;                  // 2 kJB columns of M, column c at c * kJLd
  double* Vs = jsm + 2 * kJB * kJLd;  // 2 kJB columns of V
  __shared__ double dsc[2 * kJB];    // scales of the LDS columns
  cg::grid_group grid = cg::this_grid();
  const int t = threadIdx.x, lane = t & 63;
  const int u = (t >> 6) * 4 + (lane >> 4);  // this row's pair slot 0 .. kJB-1
  const int li = lane & 15;
  const int nblk = g.nblk;
  const int w = blockIdx.x;
  // the kJB columns of block B of M and V <-> LDS columns c0.. (zero rows from
  // k up to 288).  Row u of the workgroup moves column u; every load is
  // issued before the first LDS store.
  auto load_block = [&](int B, int c0) {
    if (li == 0) dsc[c0 + u] = 1.0;
    double lm[kJT], lv[kJT];
    const int64_t base = int64_t(B * kJB + u) * k;
#pragma unroll
    for (int q = 0; q < kJT; ++q) {
      const int i = jrow(li, q);
      lm[q] = i < k ? g.M[base + i] : 0.0;
      lv[q] = i < k ? g.V[base + i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kJT; ++q) {
      Ms[(c0 + u) * kJLd + jrow(li, q)] = lm[q];
      Vs[(c0 + u) * kJLd + jrow(li, q)] = lv[q];
    }
  };
  auto store_block = [&](int B, int c0) {
    const int64_t base = int64_t(B * kJB + u) * k;
    const double d = dsc[c0 + u];
#pragma unroll
    for (int q = 0; q < kJT; ++q) {
      const int i = jrow(li, q);
      if (i < k) {
        g.M[base + i] = d * Ms[(c0 + u) * kJLd + i];
        g.V[base + i] = d * Vs[(c0 + u) * kJLd + i];
      }
    }
  };
  // a lane's 18 entries of LDS column c as 9 16-byte reads and writes
  typedef double f64x2 __attribute__((ext_vector_type(2)));
  auto lds_col = [&](const double* base, int c, double (&x)[kJT]) {
#pragma unroll
    for (int q = 0; q < kJT; q += 2) {
      const f64x2 v = *reinterpret_cast<const f64x2*>(base + c * kJLd + jrow(li, q));
      x[q] = v.x;
      x[q + 1] = v.y;
    }
  };
  auto lds_put = [&](double* base, int c, const double (&x)[kJT]) {
#pragma unroll
    for (int q = 0; q < kJT; q += 2)
      *reinterpret_cast<f64x2*>(base + c * kJLd + jrow(li, q)) = f64x2{x[q], x[q + 1]};
  };
  auto grid_barrier = [&]() {
    __threadfence();
    grid.sync();
  };
  int sweep = 0;
  for (; sweep < kJMaxSweeps; ++sweep) {
    int nrot = 0;  // this lane's rotations (lane 0 of each row counts)
    // intra pass: blocks 2w and 2w+1 in LDS columns 0.. kJB-1 and kJB.. 2kJB-1
    load_block(2 * w, 0);
    load_block(2 * w + 1, kJB);
    __syncthreads();
    {
      const int half = u / (kJB / 2), ap = u % (kJB / 2);
      for (int ir = 0; ir < kJB - 1; ++ir) {
        const int p = kJB * half + circle_item(ap, ir, kJB);
        const int q = kJB * half + circle_item(kJB - 1 - ap, ir, kJB);
        double x[kJT], y[kJT];
        lds_col(Ms, p, x);
        lds_col(Ms, q, y);
        double dx = dsc[p], dy = dsc[q], c, tn;
        const bool r = pair_rotation(x, y, dx, dy, g.tol, c, tn);
        if (__builtin_amdgcn_ballot_w64(r)) {
          double vx[kJT], vy[kJT];
          lds_col(Vs, p, vx);
          lds_col(Vs, q, vy);
          apply_rotation(x, vx, y, vy, dx, dy, c, tn);
          if (r) {
            lds_put(Ms, p, x);
            lds_put(Vs, p, vx);
            lds_put(Ms, q, y);
            lds_put(Vs, q, vy);
            if (li == 0) {
              dsc[p] = dx;
              dsc[q] = dy;
            }
          }
        }
        nrot += r && li == 0;
        __syncthreads();
      }
    }
    store_block(2 * w, 0);
    store_block(2 * w + 1, kJB);
    grid_barrier();
    // cross passes over the round-robin schedule of blocks
    for (int rr = 0; rr < nblk - 1; ++rr) {
      const int I = circle_item(w, rr, nblk), J = circle_item(nblk - 1 - w, rr, nblk);
      // column u of I (and of V) in registers, J in LDS columns 0 .. kJB-1
      double x[kJT], vx[kJT];
#pragma unroll
      for (int q = 0; q < kJT; ++q) {
        const int i = jrow(li, q);
        const int64_t a = int64_t(I * kJB + u) * k + i;
        x[q] = i < k ? g.M[a] : 0.0;
        vx[q] = i < k ? g.V[a] : 0.0;
      }
      double dx = 1.0;
      load_block(J, 0);
      __syncthreads();
      for (int ir = 0; ir < kJB; ++ir) {
        const int q = (u + ir) % kJB;
        double y[kJT];
        lds_col(Ms, q, y);
        double dy = dsc[q], c, tn;
        const bool r = pair_rotation(x, y, dx, dy, g.tol, c, tn);
        if (__builtin_amdgcn_ballot_w64(r)) {
          double vy[kJT];
          lds_col(Vs, q, vy);
          apply_rotation(x, vx, y, vy, dx, dy, c, tn);
          if (r) {
            lds_put(Ms, q, y);
            lds_put(Vs, q, vy);
            if (li == 0) dsc[q] = dy;
          }
        }
        nrot += r && li == 0;
        __syncthreads();
      }
#pragma unroll
      for (int q = 0; q < kJT; ++q) {
        const int i = jrow(li, q);
        if (i < k) {
          const int64_t a = int64_t(I * kJB + u) * k + i;
          g.M[a] = dx * x[q];
          g.V[a] = dx * vx[q];
        }
      }
      store_block(J, 0);
      grid_barrier();
    }
    // one atomic per wavefront
    const int wrot = __builtin_amdgcn_readfirstlane(
        static_cast<int>(__builtin_popcountll(__builtin_amdgcn_ballot_w64(nrot != 0))));
    int tot = nrot;
    tot += __shfl_xor(tot, 16, 64);
    tot += __shfl_xor(tot, 32, 64);
    if (lane == 0 && wrot) atomicAdd(g.rot + sweep, tot);
    grid_barrier();
    const int total = __hip_atomic_load(g.rot + sweep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (total == 0) {
      ++sweep;
      break;
    }
  }
  if (w == 0 && t == 0) *g.sweeps = sweep;
}

// S_j = ||M(:, j)||, ranked descending (ties by index); U(:, r) = M(:, j) / S_j,
// V(:, r) = Vin(:, j) for the column j of rank r.  One workgroup.
__global__ __launch_bounds__(1024) void svd_finish_kernel(int k, const double* __restrict__ M,
                                                          const double* __restrict__ Vin,
                                                          double* __restrict__ U, int64_t ldu,
                                                          double* __restrict__ S,
                                                          double* __restrict__ V, int64_t ldv) {
  __shared__ double sig[kJMaxK];
  __shared__ int dest[kJMaxK];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  for (int j = wave; j < k; j += 16) {
    double a = 0.0;
    for (int i = lane; i < k; i += 64) a = fma(M[int64_t(j) * k + i], M[int64_t(j) * k + i], a);
    a = wave_sum(a);
    if (lane == 0) sig[j] = sqrt(a);
  }
  __syncthreads();
  for (int j = t; j < k; j += 1024) {
    int r = 0;
    const double sj = sig[j];
    for (int l = 0; l < k; ++l) r += (sig[l] > sj) || (sig[l] == sj && l < j);
    dest[j] = r;
    S[r] = sj;
  }
  __syncthreads();
  for (int j = wave; j < k; j += 16) {
    const int r = dest[j];
    const double inv = sig[j] > 0.0 ? 1.0 / sig[j] : 0.0;
    for (int i = lane; i < k; i += 64) {
      U[int64_t(r) * ldu + i] = M[int64_t(j) * k + i] * inv;
      V[int64_t(r) * ldv + i] = Vin[int64_t(j) * k + i];
    }
  }
}

// M <- X' (k x k) into a k x kp buffer with zero padding columns; V <- I
__global__ void svd_init_kernel(int k, int kp, const double* __restrict__ X, int64_t ldx,
                                int transpose, double* __restrict__ M, double* __restrict__ V,
                                int* rot) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e < kJMaxSweeps) rot[e] = 0;
  if (e >= int64_t(k) * kp) return;
  const int j = int(e / k), i = int(e - int64_t(j) * k);
  double m = 0.0;
  if (j < k) m = transpose ? X[int64_t(i) * ldx + j] : X[int64_t(j) * ldx + i];
  M[e] = m;
  V[e] = (i == j) ? 1.0 : 0.0;
}

}  // namespace

size_t svd_small_work_bytes(int64_t k) {
  const int64_t nblk = 2 * ceil_div(k, 2 * kJB);
  return size_t(2) * k * nblk * kJB * sizeof(double) + 256 + kJMaxSweeps * sizeof(int) + 256;
}

// X (k x k, ld ldx; its transpose when transpose != 0) = U diag(S) V'.
int svd_small_f64(rla_ctx* ctx, int64_t k, const double* X, int64_t ldx, int transpose,
                  double* U, int64_t ldu, double* S, double* V, int64_t ldv, void* work,
                  size_t work_bytes, int* sweeps_out) {
  if (k <= 0) return RLA_OK;
  if (k > kJMaxK) return fail(RLA_ENOTSUP, "small SVD: order above 288");
  if (work_bytes < svd_small_work_bytes(k)) return fail(RLA_EARG, "small SVD: workspace too small");
  const int nblk = static_cast<int>(2 * ceil_div(k, 2 * kJB));
  const int kp = nblk * kJB;
  char* p = static_cast<char*>(work);
  double* M = reinterpret_cast<double*>(p);
  double* Vw = M + int64_t(k) * kp;
  int* rot = reinterpret_cast<int*>(reinterpret_cast<uintptr_t>(Vw + int64_t(k) * kp + 31) &
                                    ~uintptr_t(255));
  int* sweeps = rot + kJMaxSweeps;
  const int64_t n = std::max<int64_t>(int64_t(k) * kp, kJMaxSweeps);
  svd_init_kernel<<<dim3(unsigned(ceil_div(n, 256))), 256, 0, ctx->stream>>>(
      int(k), kp, X, ldx, transpose, M, Vw, rot);
  RLA_LAUNCH_CHECK("svd_init");
  const size_t lds = size_t(4 * kJB) * kJLd * sizeof(double);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(&svd_jacobi_kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(lds));
  });
  RLA_HIP(attr_err);
  JacobiArgs a{};
  a.k = static_cast<int>(k);
  a.nblk = nblk;
  a.M = M;
  a.V = Vw;
  a.rot = rot;
  a.sweeps = sweeps;
  // sqrt(k) eps: columns whose cosine is below it count as orthogonal
  a.tol = sqrt(double(k)) * 2.220446049250313e-16;
  void* args[] = {&a};
  RLA_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&svd_jacobi_kernel),
                                     dim3(unsigned(nblk / 2)), dim3(kJThreads), args, lds,
                                     ctx->stream));
  svd_finish_kernel<<<1, 1024, 0, ctx->stream>>>(int(k), M, Vw, U, ldu, S, V, ldv);
  RLA_LAUNCH_CHECK("svd_finish");
  if (sweeps_out) {
    RLA_HIP(hipMemcpyAsync(sweeps_out, sweeps, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    RLA_HIP(hipStreamSynchronize(ctx->stream));
  }
  return RLA_OK;
}

}  // namespace rla

using namespace rla;

extern "C" {

int rla_svd_small_workspace_bytes(int64_t k, size_t* bytes) {
  if (!bytes || k < 0) return fail(RLA_EARG, "rla_svd_small_workspace_bytes: bad arguments");
  *bytes = svd_small_work_bytes(k);
  return RLA_OK;
}

int rla_svd_small_f64(rla_ctx* ctx, int64_t k, const double* X, int64_t ldx, int transpose,
                      double* U, int64_t ldu, double* S, double* V, int64_t ldv, void* work,
                      size_t work_bytes, int* sweeps) {
  RLA_CHECK_CTX(ctx);
  if (k < 0 || ldx < std::max<int64_t>(1, k) || ldu < std::max<int64_t>(1, k) ||
      ldv < std::max<int64_t>(1, k))
    return fail(RLA_EARG, "rla_svd_small_f64: bad sizes");
  if (k > 0 && (!X || !U || !S || !V || !work)) return fail(RLA_EARG, "rla_svd_small_f64: NULL");
  int rc = use_device(ctx);
  if (rc) return rc;
  return svd_small_f64(ctx, k, X, ldx, transpose, U, ldu, S, V, ldv, work, work_bytes, sweeps);
}

}  // extern "C"
