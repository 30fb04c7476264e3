// Batched sketch-and-solve least squares: for each problem p,
//   x_p = argmin || S_p A_p x - S_p b_p ||,
// S_p a CountSketch (count_sketch.jl:113-213) and the small solve the
// QRSolver sub-solver's `qr!(SA) \ Sb` (src/Solvers/SubSolvers/qr.jl:36-43).
//
// One plan runs a whole batch of problems of uneven m x n as four launches
// captured in a hipGraph:
//   1. ls_sketch_kernel, persistent: workgroups pull tiles (problem, block of
//      CB columns of [A | b]) from a queue ordered longest first; a tile
//      streams its columns through LDS in chunks of R rows and each thread
//      accumulates its sketch row(s) over the bucket members in ascending row
//      order (the order of SparseArrays' sparse x dense loop, so SA and Sb are
//      bit-identical to the reference's mul!).  (CB, R) is picked per plan by
//      timing the candidates on the batch itself.
//   2. ls_gram_kernel: the augmented Gram matrix [SA Sb]'[SA Sb] per problem,
//      16 x 16 upper blocks on the FP64 MFMA (FP32 products are exact in FP64).
//   3. the register-resident one-workgroup Cholesky (chol.hip) per problem on
//      the augmented matrix: R_aug = [R z; 0 rho] with R'z = SA'Sb, so the
//      forward substitution comes for free and rho is the sketched residual;
//      pivots of R are tested, rho's is not.
//   4. ls_backsolve_kernel: x = R^-1 z by 16-row blocks with the inverted
//      diagonal blocks.
// The normal-equation route is taken in FP64 on purpose: its error is
// ~cond(SA)^2 * 1e-16, below the FP32 data precision for cond(SA) < 1e4,
// while an FP32 Householder QR (the reference's qr! on Float32) is ~cond * 6e-8.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "rla_internal.hpp"

namespace rla {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Pointers read from a problem descriptor in memory are generic (flat) to the
// compiler; flat loads also count as LDS traffic, so a wait for a descriptor
// field would drain every load in flight.  Loads through these global-space
// views are plain global loads.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> as_global(const T* p) {
  return (gptr<T>)p;
}

constexpr int kLsThreads = 1024;
constexpr int kLsMaxS = 1024;
constexpr int kLsMaxN = 271;  // n + 1 <= 272 for the batched Cholesky

struct LsDev {
  int64_t m, lda;
  int n, s;
  const float* A;
  const float* b;
  const int* rowptr;  // CSR of S (s x m): bucket r = rows colidx[rowptr[r] .. rowptr[r+1])
  const int* colidx;
  const float* val;
  float* SAb;  // s x (n+1), column-major, ld s
  double* G;   // (n+1) x (n+1), ld n+1
  double* Rt;
  double* Dinv;
  int* info;
  float* x;
};

struct SketchTile {
  int p, cb;
};
struct GramTile {
  int p, bi, bj;
};

__global__ void ls_reset_kernel(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

// Per tile, rows of [A b] stream through a double-buffered LDS chunk: the
// global loads of chunk i+1 are in flight while chunk i is summed.  Each
// thread owns one sketch row (bucket) r and a share of the CB columns, and
// adds its bucket's members in ascending row order; the next KW member
// indices and signs are loaded one chunk ahead into registers, so the
// common case (at most KW members of a bucket per chunk) runs without a
// dependent global load.
constexpr int kLsWin = 8;

template <int CB, int R>
__global__ __launch_bounds__(kLsThreads) void ls_sketch_kernel(const LsDev* __restrict__ probs,
                                                               const SketchTile* __restrict__ tiles,
                                                               int ntiles, int* counter) {
  constexpr int PER = CB * R / kLsThreads;  // chunk entries loaded per thread
  static_assert(PER * kLsThreads == CB * R, "chunk must split evenly over the threads");
  __shared__ float T[2][R][CB + 1];
  __shared__ int tile_id;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) tile_id = atomicAdd(counter, 1);
    __syncthreads();
    const int ti = tile_id;
    __syncthreads();  // every thread has read tile_id before it is reused
    if (ti >= ntiles) break;  // uniform: every wave leaves here once the queue is empty
    const SketchTile tl = tiles[ti];
    const LsDev& P = probs[tl.p];
    const int s = P.s, n = P.n, k = P.n + 1;
    const int64_t m = P.m, lda = P.lda;
    const gptr<float> Ag = as_global(P.A);
    const gptr<float> bg = as_global(P.b);
    const gptr<int> rowptr = as_global(P.rowptr);
    const gptr<int> colidx = as_global(P.colidx);
    const gptr<float> val = as_global(P.val);
    const global_ptr<float> SAb = (global_ptr<float>)P.SAb;
    int nsplit = 1;  // threads per sketch row
    while (nsplit < CB && nsplit * 2 * s <= kLsThreads) nsplit *= 2;
    const int cpt = CB / nsplit;
    const int r = tid / nsplit, g = tid % nsplit;
    const bool active = r < s;
    int cur = active ? rowptr[r] : 0;
    const int end = active ? rowptr[r + 1] : 0;
    float acc[CB];
#pragma unroll
    for (int q = 0; q < CB; ++q) acc[q] = 0.0f;

    // chunk entry i of this thread: column c = e / R, row jj = e % R with
    // e = tid + i * kLsThreads.  Rows past m and columns past k read clamped
    // (valid) addresses: those entries are never summed into a stored value.
    float stage[PER];
    auto load_chunk = [&](int64_t j0) {
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int e = tid + i * kLsThreads;
        const int c = e / R, jj = e % R;
        const int col = min(tl.cb + c, k - 1);
        const int64_t row = min<int64_t>(j0 + jj, m - 1);
        const gptr<float> src = col < n ? Ag + int64_t(col) * lda : bg;
        stage[i] = src[row];
      }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int e = tid + i * kLsThreads;
        T[buf][e % R][e / R] = stage[i];
      }
    };
    // window: members cur .. cur+kLsWin-1 of the bucket (indices clamped to
    // the bucket's last member; only those before `end` are used)
    int wj[kLsWin];
    float ws[kLsWin];
    auto load_window = [&]() {
#pragma unroll
      for (int w = 0; w < kLsWin; ++w) {
        const int pos = active ? min(cur + w, end - 1) : 0;
        wj[w] = active && cur + w < end ? colidx[pos] : INT32_MAX;
        ws[w] = active && cur + w < end ? val[pos] : 0.0f;
      }
    };
    auto add_row = [&](const float* trow, float sg) {
#pragma unroll
      for (int q = 0; q < CB; ++q)
        if (q < cpt) acc[q] = acc[q] + sg * trow[g + nsplit * q];
    };

    load_window();
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    int buf = 0;
    for (int64_t j0 = 0; j0 < m; j0 += R) {
      const bool more = j0 + R < m;
      if (more) load_chunk(j0 + R);
      const int64_t j1 = min<int64_t>(j0 + R, m);
      int used = 0;
#pragma unroll
      for (int w = 0; w < kLsWin; ++w) {
        if (wj[w] < j1 && used == w) {
          add_row(T[buf][wj[w] - j0], ws[w]);
          ++used;
        }
      }
      cur += used;
      if (used == kLsWin) {
        // more than kLsWin members of this bucket in the chunk
        while (cur < end) {
          const int j = colidx[cur];
          if (j >= j1) break;
          add_row(T[buf][j - j0], val[cur]);
          ++cur;
        }
      }
      if (more) {
        load_window();
        store_chunk(buf ^ 1);
      }
      __syncthreads();
      buf ^= 1;
    }
    if (active) {
#pragma unroll
      for (int q = 0; q < CB; ++q) {
        const int c = tl.cb + g + nsplit * q;
        if (q < cpt && c < k) SAb[r + int64_t(c) * s] = acc[q];
      }
    }
  }
}

// one wavefront per 16 x 16 block (bi <= bj) of [SA Sb]'[SA Sb]
__global__ __launch_bounds__(256) void ls_gram_kernel(const LsDev* __restrict__ probs,
                                                      const GramTile* __restrict__ tiles,
                                                      int ntiles) {
  const int lane = threadIdx.x & 63;
  const int ti = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (ti >= ntiles) return;
  const GramTile tl = tiles[ti];
  const LsDev& P = probs[tl.p];
  const int s = P.s, k = P.n + 1;
  const int ca = tl.bi * 16 + (lane & 15), cb = tl.bj * 16 + (lane & 15);
  const bool oka = ca < k, okb = cb < k;
  // a lane of a column past k reads column 0 instead: its products only reach
  // output rows / columns past k, which are not stored
  const gptr<float> pa = as_global(P.SAb) + int64_t(oka ? ca : 0) * s;
  const gptr<float> pb = as_global(P.SAb) + int64_t(okb ? cb : 0) * s;
  f64x4 acc0 = f64x4{0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
  int r0 = 0;
  // 32 rows per iteration: lane (i, kq) loads rows 4kq .. 4kq+3 and
  // 16+4kq .. 16+4kq+3 of its column as two 16-byte pieces (when s is a
  // multiple of 4, as for s = 4n, so the pieces are aligned); MFMA q of a piece takes
  // component q, i.e. rows {q, 4+q, 8+q, 12+q} of the 16: the same row
  // permutation on both operands, so the sum is the Gram product.
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  const int kq = lane >> 4;
  for (; (s & 3) == 0 && r0 + 32 <= s; r0 += 32) {
    const f32x4 a0 = *reinterpret_cast<gptr<f32x4>>(pa + r0 + 4 * kq);
    const f32x4 a1 = *reinterpret_cast<gptr<f32x4>>(pa + r0 + 16 + 4 * kq);
    const f32x4 b0 = *reinterpret_cast<gptr<f32x4>>(pb + r0 + 4 * kq);
    const f32x4 b1 = *reinterpret_cast<gptr<f32x4>>(pb + r0 + 16 + 4 * kq);
#pragma unroll
    for (int q = 0; q < 4; q += 2) {
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(double(a0[q]), double(b0[q]), acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(double(a0[q + 1]), double(b0[q + 1]), acc1, 0,
                                                  0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; q += 2) {
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(double(a1[q]), double(b1[q]), acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(double(a1[q + 1]), double(b1[q + 1]), acc1, 0,
                                                  0, 0);
    }
  }
  for (; r0 < s; r0 += 4) {
    const int r = r0 + (lane >> 4);
    const bool okr = r < s;
    const double a0 = (oka && okr) ? static_cast<double>(pa[r]) : 0.0;
    const double b0 = (okb && okr) ? static_cast<double>(pb[r]) : 0.0;
    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc0, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = tl.bi * 16 + (lane >> 4) + 4 * q, j = tl.bj * 16 + (lane & 15);
    if (i < k && j < k) ((global_ptr<double>)P.G)[i + int64_t(j) * k] = acc0[q] + acc1[q];
  }
}

// x = R^-1 z, R = Rt[0:n, 0:n] (row-major, ld n+1), z = Rt[0:n, n]
__global__ __launch_bounds__(256) void ls_backsolve_kernel(const LsDev* __restrict__ probs) {
  __shared__ double xs[kLsMaxN + 1];
  __shared__ double part[16][17];
  __shared__ double tv[16];
  const LsDev& P = probs[blockIdx.x];
  const int n = P.n, k = n + 1, t = threadIdx.x;
  const gptr<double> Rt = as_global(P.Rt);
  const gptr<double> Dinv = as_global(P.Dinv);
  if (*as_global(P.info)) return;  // breakdown: x is left as it was
  const int nb = (n + 15) / 16;
  for (int b = nb - 1; b >= 0; --b) {
    const int rb = b * 16, w = min(16, n - rb);
    {
      const int i = t >> 4, q = t & 15;
      double sum = 0.0;
      if (i < w)
        for (int c = rb + 16 + q; c < n; c += 16) sum += Rt[int64_t(rb + i) * k + c] * xs[c];
      part[i][q] = sum;
    }
    __syncthreads();
    if (t < 16) {
      double v = 0.0;
      if (t < w) {
        v = Rt[int64_t(rb + t) * k + n];
        for (int q = 0; q < 16; ++q) v -= part[t][q];
      }
      tv[t] = v;
    }
    __syncthreads();
    if (t < w) {
      double v = 0.0;
      for (int j = 0; j < w; ++j) v += Dinv[b * 256 + t * 16 + j] * tv[j];
      xs[rb + t] = v;
    }
    __syncthreads();
  }
  for (int i = t; i < n; i += blockDim.x) ((global_ptr<float>)P.x)[i] = static_cast<float>(xs[i]);
}

}  // namespace rla

using namespace rla;

struct rla_ls_plan {
  rla_ctx* ctx;
  int64_t nprob;
  void* mem;  // one device allocation for descriptors, tiles and per-problem state
  size_t bytes;  // its size, charged to the device's HBM account
  LsDev* probs;
  SketchTile* stiles;
  int nstiles;
  GramTile* gtiles;
  int ngtiles;
  CholBatchItem* items;
  int* counter;
  int* infos;
  int cfg;  // index into the (CB, R) candidates
  float tuned_ms[3];
  hipStream_t cap;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

namespace {

// (CB, R): CB columns per tile, R rows per LDS chunk
constexpr int kCfgCB[3] = {8, 16, 32};
constexpr int kCfgR[3] = {2048, 1024, 512};

int launch_sketch(rla_ls_plan* p, hipStream_t st, int cfg) {
  ls_reset_kernel<<<1, 64, 0, st>>>(p->counter);
  RLA_LAUNCH_CHECK("ls_reset");
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const unsigned grid = static_cast<unsigned>(std::max(1, std::min(p->nstiles, 2 * ncu)));
  switch (cfg) {
    case 0:
      ls_sketch_kernel<8, 2048><<<grid, kLsThreads, 0, st>>>(p->probs, p->stiles, p->nstiles,
                                                             p->counter);
      break;
    case 1:
      ls_sketch_kernel<16, 1024><<<grid, kLsThreads, 0, st>>>(p->probs, p->stiles, p->nstiles,
                                                              p->counter);
      break;
    default:
      ls_sketch_kernel<32, 512><<<grid, kLsThreads, 0, st>>>(p->probs, p->stiles, p->nstiles,
                                                             p->counter);
  }
  RLA_LAUNCH_CHECK("ls_sketch");
  return RLA_OK;
}

int launch_solve(rla_ls_plan* p, hipStream_t st) {
  if (p->ngtiles) {
    ls_gram_kernel<<<dim3(unsigned(ceil_div(p->ngtiles, 4))), 256, 0, st>>>(p->probs, p->gtiles,
                                                                           p->ngtiles);
    RLA_LAUNCH_CHECK("ls_gram");
  }
  int rc = chol_blocks_batched(st, p->nprob, p->items);
  if (rc) return rc;
  ls_backsolve_kernel<<<dim3(unsigned(p->nprob)), 256, 0, st>>>(p->probs);
  RLA_LAUNCH_CHECK("ls_backsolve");
  return RLA_OK;
}

// tiles of the current configuration, longest first
void build_sketch_tiles(const std::vector<rla_ls_problem>& v, int cb,
                        std::vector<SketchTile>& out) {
  out.clear();
  std::vector<int> order(v.size());
  for (size_t i = 0; i < v.size(); ++i) order[i] = static_cast<int>(i);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return v[a].m > v[b].m; });
  for (int p : order)
    for (int c = 0; c < v[p].n + 1; c += cb) out.push_back(SketchTile{p, c});
}

}  // namespace

extern "C" {

int rla_ls_plan_create(rla_ctx* ctx, int64_t nprob, const rla_ls_problem* probs, int autotune,
                       rla_ls_plan** out) {
  RLA_CHECK_CTX(ctx);
  if (!out || !probs || nprob < 1 || nprob > (1 << 20))
    return fail(RLA_EARG, "rla_ls_plan_create: bad arguments");
  *out = nullptr;
  std::vector<rla_ls_problem> v(probs, probs + nprob);
  for (int64_t i = 0; i < nprob; ++i) {
    const auto& q = v[i];
    if (q.m < 1 || q.n < 1 || q.lda < q.m || q.s < 1 || !q.A || !q.b || !q.x || !q.s_rowptr ||
        !q.s_colidx || !q.s_val)
      return fail(RLA_EARG, "rla_ls_plan_create: problem " + std::to_string(i) + " is malformed");
    if (q.n > kLsMaxN) return fail(RLA_ENOTSUP, "least-squares problems are limited to 271 columns");
    if (q.s > kLsMaxS) return fail(RLA_ENOTSUP, "sketches are limited to 1024 rows");
    if (q.s < q.n)
      return fail(RLA_EARG, "problem " + std::to_string(i) + ": sketch has fewer rows than columns");
    if (q.m > INT32_MAX) return fail(RLA_ENOTSUP, "more than 2^31-1 rows");
  }
  int rc = use_device(ctx);
  if (rc) return rc;
  // layout of the single device allocation
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  std::vector<SketchTile> st;
  build_sketch_tiles(v, 8, st);  // the largest tile count (CB = 8)
  std::vector<GramTile> gt;
  for (int64_t p = 0; p < nprob; ++p) {
    const int nbk = static_cast<int>((v[p].n + 1 + 15) / 16);
    for (int bj = 0; bj < nbk; ++bj)
      for (int bi = 0; bi <= bj; ++bi) gt.push_back(GramTile{static_cast<int>(p), bi, bj});
  }
  const size_t o_probs = take(sizeof(LsDev) * nprob);
  const size_t o_st = take(sizeof(SketchTile) * st.size());
  const size_t o_gt = take(sizeof(GramTile) * gt.size());
  const size_t o_items = take(sizeof(CholBatchItem) * nprob);
  const size_t o_counter = take(sizeof(int));
  const size_t o_info = take(sizeof(int) * nprob);
  std::vector<size_t> o_sab(nprob), o_g(nprob), o_rt(nprob), o_dinv(nprob);
  for (int64_t p = 0; p < nprob; ++p) {
    const int64_t k = v[p].n + 1;
    o_sab[p] = take(sizeof(float) * v[p].s * k);
    o_g[p] = take(sizeof(double) * k * k);
    o_rt[p] = take(sizeof(double) * k * k);
    o_dinv[p] = take(sizeof(double) * ceil_div(k, 16) * 256);
  }
  void* mem = nullptr;
  rc = hbm_resize(ctx->device, 0, off, "rla_ls_plan_create");
  if (rc) return rc;
  if (hipMalloc(&mem, off) != hipSuccess) {
    (void)hipGetLastError();
    (void)hbm_resize(ctx->device, off, 0, "rla_ls_plan_create");
    return fail(RLA_ENOMEM, "rla_ls_plan_create: device allocation failed");
  }
  auto* plan = new rla_ls_plan{};
  plan->ctx = ctx;
  plan->bytes = off;
  plan->nprob = nprob;
  plan->mem = mem;
  char* base = static_cast<char*>(mem);
  plan->probs = reinterpret_cast<LsDev*>(base + o_probs);
  plan->stiles = reinterpret_cast<SketchTile*>(base + o_st);
  plan->gtiles = reinterpret_cast<GramTile*>(base + o_gt);
  plan->items = reinterpret_cast<CholBatchItem*>(base + o_items);
  plan->counter = reinterpret_cast<int*>(base + o_counter);
  plan->infos = reinterpret_cast<int*>(base + o_info);
  plan->ngtiles = static_cast<int>(gt.size());
  std::vector<LsDev> hd(nprob);
  std::vector<CholBatchItem> hi(nprob);
  for (int64_t p = 0; p < nprob; ++p) {
    const auto& q = v[p];
    LsDev d{};
    d.m = q.m;
    d.lda = q.lda;
    d.n = static_cast<int>(q.n);
    d.s = static_cast<int>(q.s);
    d.A = q.A;
    d.b = q.b;
    d.rowptr = q.s_rowptr;
    d.colidx = q.s_colidx;
    d.val = q.s_val;
    d.SAb = reinterpret_cast<float*>(base + o_sab[p]);
    d.G = reinterpret_cast<double*>(base + o_g[p]);
    d.Rt = reinterpret_cast<double*>(base + o_rt[p]);
    d.Dinv = reinterpret_cast<double*>(base + o_dinv[p]);
    d.info = plan->infos + p;
    d.x = q.x;
    hd[p] = d;
    hi[p] = CholBatchItem{d.n + 1, d.n, d.G, d.Rt, d.Dinv, d.info};
  }
  auto cleanup = [&](int code) {
    (void)hipFree(mem);
    (void)hbm_resize(ctx->device, off, 0, "rla_ls_plan_create");
    delete plan;
    return code;
  };
  if (hipMemcpy(plan->probs, hd.data(), sizeof(LsDev) * nprob, hipMemcpyHostToDevice) ||
      hipMemcpy(plan->gtiles, gt.data(), sizeof(GramTile) * gt.size(), hipMemcpyHostToDevice) ||
      hipMemcpy(plan->items, hi.data(), sizeof(CholBatchItem) * nprob, hipMemcpyHostToDevice))
    return cleanup(fail(RLA_EHIP, "rla_ls_plan_create: descriptor upload failed"));
  // autotune the sketch tile shape on this batch (min of 3 timed runs each)
  // autotune: 0 = default shape, 1 = time all three, 2..4 = force shape 0..2
  // (a shape from the caller's tile cache: only that shape is timed, for
  // rla_ls_plan_config)
  const int forced = autotune >= 2 && autotune <= 4 ? autotune - 2 : -1;
  if (autotune > 4 || autotune < 0) return cleanup(fail(RLA_EARG, "autotune must be 0..4"));
  const bool tune = autotune == 1;
  plan->cfg = forced >= 0 ? forced : 1;
  for (int c = 0; c < 3; ++c) plan->tuned_ms[c] = -1.f;
  const int ncfg = tune ? 3 : (forced >= 0 ? 1 : 0);
  for (int ci = 0; ci < ncfg; ++ci) {
    const int c = tune ? ci : forced;
    build_sketch_tiles(v, kCfgCB[c], st);
    plan->nstiles = static_cast<int>(st.size());
    if (hipMemcpy(plan->stiles, st.data(), sizeof(SketchTile) * st.size(), hipMemcpyHostToDevice))
      return cleanup(fail(RLA_EHIP, "rla_ls_plan_create: tile upload failed"));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      (void)hipEventRecord(e0, ctx->stream);
      rc = launch_sketch(plan, ctx->stream, c);
      (void)hipEventRecord(e1, ctx->stream);
      if (rc) break;
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return cleanup(rc);
    plan->tuned_ms[c] = best;
  }
  if (tune) {
    int bi = 0;
    for (int c = 1; c < 3; ++c)
      if (plan->tuned_ms[c] < plan->tuned_ms[bi]) bi = c;
    plan->cfg = bi;
  }
  build_sketch_tiles(v, kCfgCB[plan->cfg], st);
  plan->nstiles = static_cast<int>(st.size());
  if (hipMemcpy(plan->stiles, st.data(), sizeof(SketchTile) * st.size(), hipMemcpyHostToDevice))
    return cleanup(fail(RLA_EHIP, "rla_ls_plan_create: tile upload failed"));
  // capture the four launches once; every execute replays the graph
  if (hipStreamCreateWithFlags(&plan->cap, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(RLA_EHIP, "rla_ls_plan_create: stream creation failed"));
  if (hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipStreamDestroy(plan->cap);
    return cleanup(fail(RLA_EHIP, "rla_ls_plan_create: capture failed"));
  }
  rc = launch_sketch(plan, plan->cap, plan->cfg);
  if (!rc) rc = launch_solve(plan, plan->cap);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(plan->cap, &g);
  if (rc || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(plan->cap);
    return cleanup(rc ? rc : fail(RLA_EHIP, "rla_ls_plan_create: capture failed"));
  }
  plan->graph = g;
  if (hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(plan->cap);
    return cleanup(fail(RLA_EHIP, "rla_ls_plan_create: graph instantiation failed"));
  }
  *out = plan;
  return RLA_OK;
}

int rla_ls_plan_execute(rla_ls_plan* plan) {
  if (!plan) return fail(RLA_EARG, "null plan");
  int rc = use_device(plan->ctx);
  if (rc) return rc;
  RLA_HIP(hipGraphLaunch(plan->exec, plan->ctx->stream));
  return RLA_OK;
}

int rla_ls_plan_info(rla_ls_plan* plan, int32_t* info_host) {
  if (!plan || !info_host) return fail(RLA_EARG, "null argument");
  RLA_HIP(hipMemcpyAsync(info_host, plan->infos, sizeof(int) * plan->nprob,
                         hipMemcpyDeviceToHost, plan->ctx->stream));
  RLA_HIP(hipStreamSynchronize(plan->ctx->stream));
  return RLA_OK;
}

int rla_ls_plan_config(rla_ls_plan* plan, int* cols_per_tile, int* rows_per_chunk,
                       float* tuned_ms3) {
  if (!plan) return fail(RLA_EARG, "null plan");
  if (cols_per_tile) *cols_per_tile = kCfgCB[plan->cfg];
  if (rows_per_chunk) *rows_per_chunk = kCfgR[plan->cfg];
  if (tuned_ms3)
    for (int c = 0; c < 3; ++c) tuned_ms3[c] = plan->tuned_ms[c];
  return RLA_OK;
}

int rla_ls_plan_sketch(rla_ls_plan* plan, int64_t p, float* SAb_out_dev) {
  if (!plan || p < 0 || p >= plan->nprob || !SAb_out_dev) return fail(RLA_EARG, "bad arguments");
  LsDev d;
  RLA_HIP(hipMemcpy(&d, plan->probs + p, sizeof(LsDev), hipMemcpyDeviceToHost));
  RLA_HIP(hipMemcpyAsync(SAb_out_dev, d.SAb, sizeof(float) * d.s * (d.n + 1),
                         hipMemcpyDeviceToDevice, plan->ctx->stream));
  return RLA_OK;
}

int rla_ls_plan_destroy(rla_ls_plan* plan) {
  if (!plan) return RLA_OK;
  (void)hipStreamSynchronize(plan->ctx->stream);
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  if (plan->cap) (void)hipStreamDestroy(plan->cap);
  (void)hipFree(plan->mem);
  (void)hbm_resize(plan->ctx->device, plan->bytes, 0, "rla_ls_plan_destroy");
  delete plan;
  return RLA_OK;
}

}  // extern "C"
