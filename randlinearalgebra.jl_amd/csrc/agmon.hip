// Agmon maximal-residual selection (src/Compressors/Distributions/agmon.jl:
// sample_distribution!, the top-k part): among candidate indices c (a list,
// or 0..n-1), the k with the largest |v[c]|, in descending order, ties to the
// smaller index.  v is the residual A x - b for rows (Left) or A'(A x - b)
// for columns (Right), formed by the caller.
//
// One workgroup makes k selection rounds.  Each round every thread scans its
// strided share of the candidates not yet taken, the wavefronts reduce
// (|v|, index) pairs with cross-lane shuffles, the first thread writes the
// winner and marks its candidate position in an LDS bitmap.  The cost is
// k * n_cand / 1024 candidate reads per thread; k is the compression
// dimension of a Kaczmarz-type block, so rounds are few.
#include <cfloat>
#include <mutex>

#include "rla_internal.hpp"

namespace rla {

constexpr int kAgmonThreads = 1024;
constexpr int64_t kAgmonMaxCand = int64_t(1) << 20;  // 128 KiB of LDS bitmap

// (score, index) ordering: larger score first, then smaller index.
__device__ __forceinline__ bool agmon_better(double s, int32_t c, double s0, int32_t c0) {
  return s > s0 || (s == s0 && c < c0);
}

template <class T>
__global__ __launch_bounds__(kAgmonThreads) void agmon_select_kernel(
    int64_t n_items, int64_t n_cand, const int32_t* __restrict__ cand, const T* __restrict__ v,
    int k, int32_t* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* red_s = reinterpret_cast<double*>(smem_raw);             // one per wavefront
  int32_t* red_c = reinterpret_cast<int32_t*>(red_s + kAgmonThreads / 64);
  int32_t* red_p = red_c + kAgmonThreads / 64;
  uint32_t* taken = reinterpret_cast<uint32_t*>(red_p + kAgmonThreads / 64);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t words = (n_cand + 31) / 32;
  for (int64_t w = t; w < words; w += kAgmonThreads) taken[w] = 0u;
  __syncthreads();
  for (int r = 0; r < k; ++r) {
    double bs = -1.0;
    int32_t bc = INT32_MAX;
    int32_t bp = -1;
    for (int64_t p = t; p < n_cand; p += kAgmonThreads) {
      if ((taken[p >> 5] >> (p & 31)) & 1u) continue;
      const int32_t c = cand ? cand[p] : static_cast<int32_t>(p);
      if (c < 0 || c >= n_items) continue;  // a candidate outside 0..n-1 is never chosen
      const double s = fabs(double(v[c]));
      if (bp < 0 || agmon_better(s, c, bs, bc)) {
        bs = s;
        bc = c;
        bp = static_cast<int32_t>(p);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double os = __shfl_xor(bs, o);
      const int32_t oc = __shfl_xor(bc, o);
      const int32_t op = __shfl_xor(bp, o);
      if (op >= 0 && (bp < 0 || agmon_better(os, oc, bs, bc))) {
        bs = os;
        bc = oc;
        bp = op;
      }
    }
    if (lane == 0) {
      red_s[wave] = bs;
      red_c[wave] = bc;
      red_p[wave] = bp;
    }
    __syncthreads();
    if (t == 0) {
      for (int w = 1; w < kAgmonThreads / 64; ++w)
        if (red_p[w] >= 0 && (bp < 0 || agmon_better(red_s[w], red_c[w], bs, bc))) {
          bs = red_s[w];
          bc = red_c[w];
          bp = red_p[w];
        }
      out[r] = bp >= 0 ? bc : -1;  // -1: no valid candidate left
      if (bp >= 0) taken[bp >> 5] |= 1u << (bp & 31);
    }
    __syncthreads();
  }
}

template <class T>
static int agmon_select(rla_ctx* ctx, int64_t n_items, const int32_t* cand, int64_t n_cand,
                        const T* v, int64_t k, int32_t* out) {
  RLA_CHECK_CTX(ctx);
  if (n_items < 1 || n_cand < 1 || k < 1) return fail(RLA_EARG, "agmon select: empty input");
  if (k > n_cand) return fail(RLA_EARG, "agmon select: more outputs than candidates");
  if (!cand && n_cand != n_items)
    return fail(RLA_EARG, "agmon select: without a candidate list the candidates are 0..n-1");
  if (n_cand > kAgmonMaxCand || n_items > INT32_MAX)
    return fail(RLA_ENOTSUP, "agmon select: at most 2^20 candidates");
  if (!v || !out) return fail(RLA_EARG, "agmon select: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  const size_t lds = (kAgmonThreads / 64) * (sizeof(double) + 2 * sizeof(int32_t)) +
                     size_t((n_cand + 31) / 32) * sizeof(uint32_t);
  if (lds > 64 * 1024) {
    static std::once_flag once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(once, [&] {
      attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(&agmon_select_kernel<T>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    });
    RLA_HIP(attr_err);
  }
  agmon_select_kernel<T><<<1, kAgmonThreads, lds, ctx->stream>>>(n_items, n_cand, cand, v,
                                                                 static_cast<int>(k), out);
  RLA_LAUNCH_CHECK("agmon_select");
  return RLA_OK;
}

}  // namespace rla

using namespace rla;

extern "C" {

int rla_agmon_select_f32(rla_ctx* ctx, int64_t n_items, const int32_t* cand, int64_t n_cand,
                         const float* v, int64_t k, int32_t* out) {
  return agmon_select<float>(ctx, n_items, cand, n_cand, v, k, out);
}

int rla_agmon_select_f64(rla_ctx* ctx, int64_t n_items, const int32_t* cand, int64_t n_cand,
                         const double* v, int64_t k, int32_t* out) {
  return agmon_select<double>(ctx, n_items, cand, n_cand, v, k, out);
}

}  // extern "C"
