// Block Kaczmarz (src/Solvers/kaczmarz.jl) with the Sampling compressor and
// its Uniform / L2Norm distributions (src/Compressors/sampling.jl,
// Distributions/uniform.jl, Distributions/strohmer_vershynin.jl), the LQ
// sub-solve (src/Solvers/SubSolvers/lq.jl), FullResidual and BasicLogger.
//
// One iteration projects x onto the solution set of a block of s rows,
//   x <- x + alpha * A~' (A~ A~')^-1 (b~ - A~ x),
// the minimum-norm correction the LQ sub-solver computes (lq.jl:35-46).  On the
// device an iteration is two launches:
//   * kz_fused_kernel, one workgroup per 64 columns: first the pending update
//     of the previous block (x_c += alpha * sum_i A~_prev[i,c] y_i), then the
//     gather of the new block's columns from the row-major copy of A and its
//     partial Gram matrix and partial A~ x over those columns (FP64 sums of
//     FP32 products, which are exact);
//   * kz_solve_kernel, one workgroup: sums the partials in column-chunk order,
//     r = b~ - A~ x, Cholesky of the s x s Gram matrix in FP64, y = G^-1 r.
// The update of block t is folded into the sweep that builds block t+1, so A~
// is read twice per iteration and x once.
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
constexpr int kKzCols = 64;       // columns per fused workgroup
constexpr int kKzMaxBlock = 64;   // rows per Kaczmarz block on the fused path

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

// Pass over candidates (n_rows items in all).  first: the candidates are items
// [0, n_items) with keys
// drawn here (set r uses epoch0 + r); otherwise they are the (key, index) pairs
// in_key/in_idx[r][0..n_items).  Each workgroup sorts a chunk of kSortN and
// keeps its k best.  final (one chunk): the k best are sorted by index and
// written to out_set[r][0..k).
__global__ __launch_bounds__(kSortThreads) void topk_pass_kernel(
    uint64_t seed, uint64_t epoch0, int64_t n_items, int64_t n_rows, const double* __restrict__ w,
    int k,
    const double* __restrict__ in_key, const int* __restrict__ in_idx, double* __restrict__ out_key,
    int* __restrict__ out_idx, int* __restrict__ out_set, int first, int final_pass) {
  __shared__ double key[kSortN];
  __shared__ int idx[kSortN];
  const int r = blockIdx.y, chunk = blockIdx.x;
  const int64_t base = int64_t(chunk) * kSortN;
  const int64_t cnt = min<int64_t>(kSortN, n_items - base);
  const int N = pow2_at_least(static_cast<int>(cnt));
  for (int e = threadIdx.x; e < N; e += blockDim.x) {
    double kk = -INFINITY;
    int ii = 0x7fffffff;
    if (e < cnt) {
      const int64_t item = base + e;
      if (first) {
        Stream st(seed, epoch0 + r, DOM_WSAMPLE, static_cast<uint64_t>(item));
        const double u = st.open01();
        const double wi = w ? w[item] : 1.0;
        kk = wi > 0.0 ? log(u) / wi : -INFINITY;
        ii = static_cast<int>(item);
      } else {
        kk = in_key[int64_t(r) * n_items + item];
        ii = in_idx[int64_t(r) * n_items + item];
      }
    }
    key[e] = kk;
    idx[e] = ii;
  }
  __syncthreads();
  bitonic_sort(key, idx, N);
  if (!final_pass) {
    const int64_t nout = int64_t(gridDim.x) * k;
    for (int e = threadIdx.x; e < k; e += blockDim.x) {
      out_key[int64_t(r) * nout + int64_t(chunk) * k + e] = e < cnt ? key[e] : -INFINITY;
      out_idx[int64_t(r) * nout + int64_t(chunk) * k + e] = e < cnt ? idx[e] : 0x7fffffff;
    }
    return;
  }
  // k best by index, ascending (ordered = true); entries past k are padding
  const int K = pow2_at_least(k);
  for (int e = k + threadIdx.x; e < K; e += blockDim.x) idx[e] = 0x7fffffff;
  __syncthreads();
  bitonic_sort_int(idx, K);
  // the host guarantees at least k items of positive weight; the clamp only
  // keeps a violated precondition in bounds
  for (int e = threadIdx.x; e < k; e += blockDim.x)
    out_set[int64_t(r) * k + e] = static_cast<int>(min<int64_t>(idx[e], n_rows - 1));
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

__global__ __launch_bounds__(256) void kz_fused_kernel(int64_t n, RowSrc prev,
                                                       const double* __restrict__ y, float alpha,
                                                       float* __restrict__ x, RowSrc cur,
                                                       double* __restrict__ part_G,
                                                       double* __restrict__ part_d) {
  __shared__ float T[kKzMaxBlock][kKzCols + 1];
  __shared__ float xs[kKzCols];
  const int t = threadIdx.x, w = blockIdx.x;
  const int64_t c0 = int64_t(w) * kKzCols;
  const int ncol = static_cast<int>(min<int64_t>(kKzCols, n - c0));
  // pending update of the previous block: x_c += alpha * sum_i A~[i,c] y_i
  if (t < ncol) {
    const int64_t c = c0 + t;
    float xc = x[c];
    if (prev.s > 0) {
      double u = 0.0;
      for (int i = 0; i < prev.s; ++i) u += static_cast<double>(prev.at(i, c)) * y[i];
      xc = alpha * static_cast<float>(u) + xc;
      x[c] = xc;
    }
    xs[t] = xc;
  }
  if (cur.s == 0) return;
  // gather the new block's columns c0 .. c0+ncol
  for (int e = t; e < cur.s * kKzCols; e += blockDim.x) {
    const int i = e / kKzCols, cc = e % kKzCols;
    T[i][cc] = cc < ncol ? cur.at(i, c0 + cc) : 0.0f;
  }
  if (t >= ncol && t < kKzCols) xs[t] = 0.0f;
  __syncthreads();
  const int s = cur.s;
  for (int i = t; i < s; i += blockDim.x) {
    double d = 0.0;
    for (int cc = 0; cc < kKzCols; ++cc) d += static_cast<double>(T[i][cc]) * xs[cc];
    part_d[int64_t(w) * s + i] = d;
  }
  for (int e = t; e < s * s; e += blockDim.x) {
    const int i = e / s, j = e % s;
    if (i > j) continue;
    double g = 0.0;
    for (int cc = 0; cc < kKzCols; ++cc)
      g += static_cast<double>(T[i][cc]) * static_cast<double>(T[j][cc]);
    part_G[(int64_t(w) * s + i) * s + j] = g;
  }
}

// y = (A~ A~')^-1 (b~ - A~ x) from the column-chunk partials.  A pivot not
// above 4 s eps G_jj (a numerically dependent row) gives y = 0 and sets *info
// to its 1-based position (never cleared here); an all-zero block gives y = 0
// (kaczmarz.jl:300 skips it).
__global__ __launch_bounds__(256) void kz_solve_kernel(int s, int nparts,
                                                       const double* __restrict__ part_G,
                                                       const double* __restrict__ part_d,
                                                       RowSrc bsrc, double* __restrict__ y,
                                                       int* __restrict__ info) {
  __shared__ double G[kKzMaxBlock][kKzMaxBlock + 1];
  __shared__ double r[kKzMaxBlock];
  __shared__ double fl[kKzMaxBlock];  // pivot floors 4 s eps G_jj
  __shared__ int bad;
  const int t = threadIdx.x;
  for (int e = t; e < s * s; e += blockDim.x) {
    const int i = e / s, j = e % s;
    if (i > j) continue;
    double g = 0.0;
    for (int q = 0; q < nparts; ++q) g += part_G[(int64_t(q) * s + i) * s + j];
    G[i][j] = g;
    G[j][i] = g;
  }
  for (int i = t; i < s; i += blockDim.x) {
    double d = 0.0;
    for (int q = 0; q < nparts; ++q) d += part_d[int64_t(q) * s + i];
    // b~ - A~ x (kaczmarz.jl:305, mul!(vec_view, mat_view, x, -1.0, 1.0))
    r[i] = static_cast<double>(bsrc.at(i, 0)) - d;
  }
  __syncthreads();
  for (int i = t; i < s; i += blockDim.x) fl[i] = 4.0 * s * 2.220446049250313e-16 * G[i][i];
  if (t == 0) {
    bool zero = true;
    for (int i = 0; i < s; ++i) zero = zero && G[i][i] == 0.0;
    bad = zero ? -1 : 0;
  }
  __syncthreads();
  if (bad == -1) {  // A~ == 0: no update
    for (int i = t; i < s; i += blockDim.x) y[i] = 0.0;
    return;
  }
  // Cholesky G = L L' (lower, in place)
  for (int j = 0; j < s; ++j) {
    const double p = G[j][j];
    if (!(p > fl[j]) || !(p > 0.0) || !isfinite(p)) {
      if (t == 0) bad = j + 1;
      break;
    }
    const double d = sqrt(p);
    __syncthreads();
    for (int i = j + t; i < s; i += blockDim.x) G[i][j] = i == j ? d : G[i][j] / d;
    __syncthreads();
    const int rem = s - j - 1;
    for (int e = t; e < rem * rem; e += blockDim.x) {
      const int i = j + 1 + e / rem, c = j + 1 + e % rem;
      if (c <= i) G[i][c] -= G[i][j] * G[c][j];
    }
    __syncthreads();
  }
  __syncthreads();
  if (bad) {  // singular block: skipped, and recorded (info is sticky)
    for (int i = t; i < s; i += blockDim.x) y[i] = 0.0;
    if (t == 0) *info = bad;
    return;
  }
  // L z = r, then L' y = z (in r)
  for (int l = 0; l < s; ++l) {
    const double z = r[l] / G[l][l];
    __syncthreads();
    for (int i = l + 1 + t; i < s; i += blockDim.x) r[i] -= G[i][l] * z;
    if (t == 0) r[l] = z;
    __syncthreads();
  }
  for (int l = s - 1; l >= 0; --l) {
    const double v = r[l] / G[l][l];
    __syncthreads();
    for (int i = t; i < l; i += blockDim.x) r[i] -= G[l][i] * v;
    if (t == 0) r[l] = v;
    __syncthreads();
  }
  for (int i = t; i < s; i += blockDim.x) y[i] = r[i];
}

__global__ void set_int_kernel(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ---- FP64 residual ||b - A x|| (full_residual.jl:32-41) ----------------------

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

__global__ void residual_final_kernel(int64_t nparts, const double* __restrict__ part,
                                      double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int64_t q = 0; q < nparts; ++q) s += part[q];
  *out = sqrt(s);
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

static int residual(rla_ctx* ctx, int64_t m, int64_t n, const float* A, int64_t a_rs,
                    int64_t a_cs, const float* x, const float* b, double* part, double* out) {
  const int64_t nb = ceil_div(m, 256);
  residual_part_kernel<<<dim3(unsigned(nb)), 256, 0, ctx->stream>>>(m, n, A, a_rs, a_cs, x, b,
                                                                    part);
  RLA_LAUNCH_CHECK("residual");
  residual_final_kernel<<<1, 64, 0, ctx->stream>>>(nb, part, out);
  RLA_LAUNCH_CHECK("residual/final");
  return RLA_OK;
}

struct KzBufs {
  int* sets;      // every x s
  double* part_G; // W x s x s
  double* part_d; // W x s
  double* y;      // s
  int* info;      // 1
  double* res;    // residual partials + result
  double* resv;
  double *k1, *k2, *cdf;
  int *i1, *i2;
};

static size_t kz_layout(int64_t m, int64_t n, int64_t s, int64_t every, void* work, size_t bytes,
                        KzBufs* out) {
  Bump2 b{static_cast<char*>(work), bytes};
  KzBufs k{};
  const int64_t W = ceil_div(n, kKzCols);
  k.sets = b.take<int>(every * s);
  k.part_G = b.take<double>(W * s * s);
  k.part_d = b.take<double>(W * s);
  k.y = b.take<double>(s);
  k.info = b.take<int>(1);
  k.res = b.take<double>(ceil_div(m, 256));
  k.resv = b.take<double>(1);
  sample_scratch(m, static_cast<int>(s), every, b, &k.k1, &k.i1, &k.k2, &k.i2, &k.cdf);
  if (out) *out = k;
  return b.need;
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
  rc = workspace(ctx, sizeof(double) * (ceil_div(m, 256) + 1), &ws);
  if (rc) return rc;
  double* part = static_cast<double*>(ws);
  double* res = part + ceil_div(m, 256);
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
  const int64_t W = ceil_div(n, kKzCols);
  const int si = static_cast<int>(s);
  RowSrc none{nullptr, 0, 0, nullptr, 0};
  RowSrc cur{At, at_rs, at_cs, nullptr, si};
  set_int_kernel<<<1, 1, 0, ctx->stream>>>(k.info, 0);
  RLA_LAUNCH_CHECK("set_int");
  kz_fused_kernel<<<dim3(unsigned(W)), 256, 0, ctx->stream>>>(n, none, k.y, 0.f, x, cur,
                                                              k.part_G, k.part_d);
  RLA_LAUNCH_CHECK("kz_fused");
  RowSrc bsrc{bt, bt_inc, 0, nullptr, si};
  kz_solve_kernel<<<1, 256, 0, ctx->stream>>>(si, static_cast<int>(W), k.part_G, k.part_d, bsrc,
                                              k.y, k.info);
  RLA_LAUNCH_CHECK("kz_solve");
  kz_fused_kernel<<<dim3(unsigned(W)), 256, 0, ctx->stream>>>(
      n, cur, k.y, static_cast<float>(alpha), x, none, k.part_G, k.part_d);
  RLA_LAUNCH_CHECK("kz_update");
  int info = 0;
  RLA_HIP(hipMemcpyAsync(&info, k.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (info > 0) return fail(RLA_EDOMAIN, "block Gram matrix is singular (rank-deficient block)");
  return RLA_OK;
}

// rsolve! of a Kaczmarz recipe with a Sampling compressor (kaczmarz.jl:313-368):
// for i = 1..max_it: error and logger at i = 1, 1+every, ... and at max_it
// (threshold_stop: converged when error < threshold; i == max_it always
// stops), then a new row set and a block update.  A (column-major, lda) is read
// for the FP64 residual, A_rm (row-major, ld_rm) for the blocks.  w: row
// weights (NULL: uniform).  hist receives the evaluated errors.
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
  const int64_t W = ceil_div(n, kKzCols);
  const int si = static_cast<int>(s);
  const RowSrc none{nullptr, 0, 0, nullptr, 0};
  RowSrc pending = none;
  bool cdf_ready = false;
  set_int_kernel<<<1, 1, 0, ctx->stream>>>(k.info, 0);
  RLA_LAUNCH_CHECK("set_int");
  int64_t nh = 0, it = 0;
  int conv = 0;
  int64_t set_base = 1;  // iteration whose set is sets[0]
  for (int64_t i = 1; i <= max_it; ++i) {
    it = i;
    if ((i - 1) % every == 0 || i == max_it) {
      if (pending.s) {
        kz_fused_kernel<<<dim3(unsigned(W)), 256, 0, ctx->stream>>>(
            n, pending, k.y, static_cast<float>(alpha), x, none, k.part_G, k.part_d);
        RLA_LAUNCH_CHECK("kz_update");
        pending = none;
      }
      rc = residual(ctx, m, n, A, 1, lda, x, b, k.res, k.resv);
      if (rc) return rc;
      double err = 0.0;
      int info = 0;
      RLA_HIP(hipMemcpyAsync(&err, k.resv, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      RLA_HIP(hipMemcpyAsync(&info, k.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      RLA_HIP(hipStreamSynchronize(ctx->stream));
      if (info > 0)
        return fail(RLA_EDOMAIN, "block Gram matrix is singular (rank-deficient block)");
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
    RowSrc cur{A_rm, ld_rm, 1, k.sets + (i - set_base) * s, si};
    kz_fused_kernel<<<dim3(unsigned(W)), 256, 0, ctx->stream>>>(
        n, pending, k.y, static_cast<float>(alpha), x, cur, k.part_G, k.part_d);
    RLA_LAUNCH_CHECK("kz_fused");
    RowSrc bsrc{b, 1, 0, cur.idx, si};
    kz_solve_kernel<<<1, 256, 0, ctx->stream>>>(si, static_cast<int>(W), k.part_G, k.part_d,
                                                bsrc, k.y, k.info);
    RLA_LAUNCH_CHECK("kz_solve");
    pending = cur;
  }
  if (pending.s) {
    kz_fused_kernel<<<dim3(unsigned(W)), 256, 0, ctx->stream>>>(
        n, pending, k.y, static_cast<float>(alpha), x, none, k.part_G, k.part_d);
    RLA_LAUNCH_CHECK("kz_update");
  }
  int info = 0;
  RLA_HIP(hipMemcpyAsync(&info, k.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (info > 0) return fail(RLA_EDOMAIN, "block Gram matrix is singular (rank-deficient block)");
  if (n_hist) *n_hist = nh;
  if (iterations) *iterations = it;
  if (converged) *converged = conv;
  return RLA_OK;
}

}  // extern "C"
