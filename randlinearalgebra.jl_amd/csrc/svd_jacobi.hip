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
// nblk / 2 workgroups of 1024 threads, one per CU.  A sweep is nblk - 1
// rounds of the round-robin (circle) schedule over blocks: in each round a
// workgroup loads its two blocks of M and V (k x 32 each) into LDS, runs one
// sweep of the same schedule over its 32 columns (31 inner rounds of 16
// disjoint pairs, one wavefront per pair), and stores them back; a grid
// barrier separates rounds.  The sweep count of rotations decides, after a
// grid barrier, whether another sweep runs, so every workgroup executes the
// same number of barriers.  At most max_sweeps sweeps.
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

__global__ __launch_bounds__(kJThreads) void svd_jacobi_kernel(JacobiArgs g) {
  extern __shared__ __align__(16) double jsm[];
  const int k = g.k;
  double* Ms = jsm;               // 32 columns of M, column c at c * k
  double* Vs = jsm + 32 * k;      // 32 columns of V
  cg::grid_group grid = cg::this_grid();
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int nblk = g.nblk;
  const int w = blockIdx.x;
  int sweep = 0;
  for (; sweep < kJMaxSweeps; ++sweep) {
    int nrot = 0;  // rotations done by this wavefront's lane 0 (one per pair)
    for (int r = 0; r < nblk - 1; ++r) {
      const int I = circle_item(w, r, nblk), J = circle_item(nblk - 1 - w, r, nblk);
      // load blocks I and J: local column c < 16 is global I*16 + c, else J*16 + c - 16
      for (int e = t; e < 32 * k; e += kJThreads) {
        const int c = e / k, i = e - c * k;
        const int gc = (c < kJB ? I * kJB + c : J * kJB + c - kJB);
        Ms[e] = g.M[int64_t(gc) * k + i];
        Vs[e] = g.V[int64_t(gc) * k + i];
      }
      __syncthreads();
      for (int ir = 0; ir < 31; ++ir) {
        const int p = circle_item(wave, ir, 32), q = circle_item(31 - wave, ir, 32);
        double* mp = Ms + p * k;
        double* mq = Ms + q * k;
        double a = 0.0, b = 0.0, gm = 0.0;
        for (int i = lane; i < k; i += 64) {
          const double x = mp[i], y = mq[i];
          a = fma(x, x, a);
          b = fma(y, y, b);
          gm = fma(x, y, gm);
        }
        a = wave_sum(a);
        b = wave_sum(b);
        gm = wave_sum(gm);
        if (a > 0.0 && b > 0.0 && fabs(gm) > g.tol * sqrt(a * b)) {
          const double zeta = (b - a) / (2.0 * gm);
          const double az = fabs(zeta);
          // for huge |zeta|, sqrt(1 + zeta^2) ~ |zeta| (t ~ 1 / (2 zeta))
          const double tt = (az > 1e150 ? 0.5 / az : 1.0 / (az + sqrt(fma(zeta, zeta, 1.0))));
          const double tn = zeta < 0.0 ? -tt : tt;
          const double c = 1.0 / sqrt(fma(tn, tn, 1.0));
          const double s = c * tn;
          double* vp = Vs + p * k;
          double* vq = Vs + q * k;
          for (int i = lane; i < k; i += 64) {
            const double x = mp[i], y = mq[i];
            mp[i] = c * x - s * y;
            mq[i] = s * x + c * y;
            const double u = vp[i], v = vq[i];
            vp[i] = c * u - s * v;
            vq[i] = s * u + c * v;
          }
          if (lane == 0) ++nrot;
        }
        __syncthreads();
      }
      for (int e = t; e < 32 * k; e += kJThreads) {
        const int c = e / k, i = e - c * k;
        const int gc = (c < kJB ? I * kJB + c : J * kJB + c - kJB);
        g.M[int64_t(gc) * k + i] = Ms[e];
        g.V[int64_t(gc) * k + i] = Vs[e];
      }
      __threadfence();
      grid.sync();
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
