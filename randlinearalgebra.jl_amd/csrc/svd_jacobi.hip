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
// Layout: the columns are cut into blocks of 16 (zero columns pad the last
// block; a zero column is never rotated).  One cooperative launch holds
// nblk / 2 workgroups of 1024 threads, one per CU, and M and V in global
// memory (L2).  A sweep visits every column pair once:
//   * an intra pass: workgroup w loads blocks 2w, 2w+1 into LDS and runs the
//     round-robin schedule over each block's 16 columns (15 rounds of 8
//     disjoint pairs per block, one wavefront per pair);
//   * nblk - 1 rounds of the round-robin schedule over blocks: workgroup w
//     loads its blocks I, J and visits the 16 x 16 cross pairs in 16 rounds
//     (pair (I_a, J_(a+r) mod 16) on wavefront a).
// A grid barrier follows every pass.  A wavefront keeps its pair's entries
// in registers between the dot products and the rotation; the three sums
// are reduced with DPP and permlane swaps.  The sweep's rotation count
// decides, after a grid barrier, whether another sweep runs, so every
// workgroup executes the same barriers.  At most 60 sweeps.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <mutex>

#include "rla_internal.hpp"

namespace rla {

namespace {

namespace cg = cooperative_groups;

constexpr int kJB = 16;            // columns per block
constexpr int kJMaxK = 288;
constexpr int kJThreads = 1024;    // 16 wavefronts: one column pair each
constexpr int kJMaxSweeps = 60;

struct JacobiArgs {
  int k;           // rows of M and V, and real columns
  int nblk;        // column blocks (even)
  double* M;       // k x (nblk * 16), ld k
  double* V;       // k x (nblk * 16), ld k
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

// sum over the 64 lanes, in every lane: within rows of 16 by DPP (quad
// swaps, then the half-row and row mirrors), across rows by swapping a copy
// with itself (v_permlane16_swap / v_permlane32_swap)
__device__ __forceinline__ double wave_sum_fast(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);  // row_half_mirror
  v += dpp_f64<0x140>(v);  // row_mirror
  int alo = __double2loint(v), ahi = __double2hiint(v);
  int blo = alo, bhi = ahi;
  asm("v_permlane16_swap_b32 %0, %1" : "+v"(alo), "+v"(blo));
  asm("v_permlane16_swap_b32 %0, %1" : "+v"(ahi), "+v"(bhi));
  v = __hiloint2double(ahi, alo) + __hiloint2double(bhi, blo);
  alo = __double2loint(v);
  ahi = __double2hiint(v);
  blo = alo;
  bhi = ahi;
  asm("v_permlane32_swap_b32 %0, %1" : "+v"(alo), "+v"(blo));
  asm("v_permlane32_swap_b32 %0, %1" : "+v"(ahi), "+v"(bhi));
  return __hiloint2double(ahi, alo) + __hiloint2double(bhi, blo);
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

constexpr int kJRows = kJMaxK / 64;  // rows per lane (k <= 288)

// rotates columns p and q (LDS, k rows each) of M and V when they are not
// orthogonal to tolerance; returns 1 if it rotated
__device__ __forceinline__ int jacobi_pair(double* Ms, double* Vs, int k, int p, int q, int lane,
                                           double tol) {
  double* mp = Ms + p * k;
  double* mq = Ms + q * k;
  double x[kJRows + 1], y[kJRows + 1];
  double a = 0.0, b = 0.0, gm = 0.0;
#pragma unroll
  for (int r = 0; r <= kJRows; ++r) {
    const int i = lane + 64 * r;
    x[r] = i < k ? mp[i] : 0.0;
    y[r] = i < k ? mq[i] : 0.0;
    a = fma(x[r], x[r], a);
    b = fma(y[r], y[r], b);
    gm = fma(x[r], y[r], gm);
  }
  a = wave_sum_fast(a);
  b = wave_sum_fast(b);
  gm = wave_sum_fast(gm);
  if (!(a > 0.0 && b > 0.0 && fabs(gm) > tol * sqrt(a * b))) return 0;
  // zeta = (b - a) / (2 g), t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
  // c = 1 / sqrt(1 + t^2), s = c t
  const double zeta = (b - a) * rcp_nr(2.0 * gm);
  const double az = fabs(zeta);
  double tt;
  if (az > 1e150) {
    tt = 0.5 / az;
  } else {
    const double w = fma(zeta, zeta, 1.0);
    tt = rcp_nr(az + w * rsq_nr(w));
  }
  const double tn = zeta < 0.0 ? -tt : tt;
  const double c = rsq_nr(fma(tn, tn, 1.0));
  const double sn = c * tn;
  double* vp = Vs + p * k;
  double* vq = Vs + q * k;
#pragma unroll
  for (int r = 0; r <= kJRows; ++r) {
    const int i = lane + 64 * r;
    if (i < k) {
      mp[i] = c * x[r] - sn * y[r];
      mq[i] = sn * x[r] + c * y[r];
      const double u = vp[i], v = vq[i];
      vp[i] = c * u - sn * v;
      vq[i] = sn * u + c * v;
    }
  }
  return 1;
}

__global__ __launch_bounds__(kJThreads) void svd_jacobi_kernel(JacobiArgs g) {
  extern __shared__ __align__(16) double jsm[];
  const int k = g.k;
  double* Ms = jsm;               // 32 columns of M, column c at c * k
  double* Vs = jsm + 32 * k;      // 32 columns of V
  cg::grid_group grid = cg::this_grid();
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int nblk = g.nblk;
  const int w = blockIdx.x;
  // a pass moves 32 columns of M and V (at most 9 elements per thread each);
  // all of a thread's global loads are issued before any is used
  constexpr int kPer = 32 * kJMaxK / kJThreads;
  auto gaddr = [&](int I, int J, int e) {
    const int c = e / k, i = e - c * k;
    return int64_t(c < kJB ? I * kJB + c : J * kJB + c - kJB) * k + i;
  };
  auto load = [&](int I, int J) {
    double lm[kPer], lv[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int e = t + q * kJThreads;
      if (e < 32 * k) {
        const int64_t a = gaddr(I, J, e);
        lm[q] = g.M[a];
        lv[q] = g.V[a];
      }
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int e = t + q * kJThreads;
      if (e < 32 * k) {
        Ms[e] = lm[q];
        Vs[e] = lv[q];
      }
    }
    __syncthreads();
  };
  auto store = [&](int I, int J) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int e = t + q * kJThreads;
      if (e < 32 * k) {
        const int64_t a = gaddr(I, J, e);
        g.M[a] = Ms[e];
        g.V[a] = Vs[e];
      }
    }
    __threadfence();
    grid.sync();
  };
  int sweep = 0;
  for (; sweep < kJMaxSweeps; ++sweep) {
    int nrot = 0;  // this wavefront's rotations (lane 0 counts)
    // intra pass: blocks 2w and 2w+1, 8 wavefronts each
    load(2 * w, 2 * w + 1);
    {
      const int half = wave >> 3, a8 = wave & 7;
      for (int ir = 0; ir < 15; ++ir) {
        const int p = 16 * half + circle_item(a8, ir, 16);
        const int q = 16 * half + circle_item(15 - a8, ir, 16);
        nrot += jacobi_pair(Ms, Vs, k, p, q, lane, g.tol);
        __syncthreads();
      }
    }
    store(2 * w, 2 * w + 1);
    // cross passes over the round-robin schedule of blocks
    for (int r = 0; r < nblk - 1; ++r) {
      const int I = circle_item(w, r, nblk), J = circle_item(nblk - 1 - w, r, nblk);
      load(I, J);
      for (int ir = 0; ir < 16; ++ir) {
        nrot += jacobi_pair(Ms, Vs, k, wave, 16 + ((wave + ir) & 15), lane, g.tol);
        __syncthreads();
      }
      store(I, J);
    }
    if (lane == 0 && nrot) atomicAdd(g.rot + sweep, nrot);
    __threadfence();
    grid.sync();
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
  const size_t lds = size_t(64) * k * sizeof(double);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(&svd_jacobi_kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
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
