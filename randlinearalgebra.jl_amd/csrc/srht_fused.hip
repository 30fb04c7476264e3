// Single-pass SRHT forward operator: Y = alpha*(scale*H*D*pad(X))[op, :] + beta*Y.
//
// One workgroup transforms one column at a time (persistent over columns).
// A column of n_pad = 2^L entries is processed as NCH chunks of 2^LC
// contiguous entries (NCH = 1 or 2).  Inside a chunk every thread holds 32
// entries in registers; the LC butterfly stages run in phases of 5 bits
// (bits [0,5), [5,10), [10,15)), each phase in registers, and between phases
// the chunk is re-distributed through LDS so that the next 5 bits become
// register bits.  The stages are the reference's stages in the reference's
// order (fwht.jl:21-70), so results are bit-identical to it.
//
// Only the sampled outputs are needed.  With two chunks the last stage
// (bit LC) combines entry e of chunk 0 with entry e of chunk 1.  A per-recipe
// plan says, for every thread, which of its 32 registers hold a "needed"
// entry after the last phase (e or e + 2^LC sampled) as a 32-bit mask, and
// where its needed entries start in a compact slot numbering; each slot maps
// to its output rows.  Chunk 0's needed values wait in LDS (one value per
// slot); chunk 1 finishes the last stage for exactly those slots and writes
// the rows.  Global traffic is one read of X and one write of the s rows.
//
// LDS layout: entry e at e + (e >> 5) (one pad word per 32 entries).  Every
// register<->LDS move is then thread base + compile-time offset, and the
// contiguous (phase 0), strided (later phases) and load-order moves all hit
// 32 distinct banks per 32 lanes.  The slot buffer follows the chunk.
#pragma clang fp contract(off)

#include <algorithm>
#include <mutex>

#include "rla_internal.hpp"

namespace rla {

constexpr int kVals = 32;  // entries per thread in a chunk

__device__ __forceinline__ int pad(int e) { return e + (e >> 5); }

// entry index of register v of thread t when the register window starts at bit w
template <int W>
__device__ inline int map_entry(int t, int v) {
  return ((t >> W) << (W + 5)) | (v << W) | (t & ((1 << W) - 1));
}

template <int RB, class T>
__device__ __forceinline__ void stage(T (&x)[kVals]) {
#pragma unroll
  for (int v = 0; v < kVals; ++v) {
    if (!(v & (1 << RB))) {
      const T z = x[v], y = x[v | (1 << RB)];
      x[v] = z + y;
      x[v | (1 << RB)] = z - y;
    }
  }
}

template <class T>
__device__ __forceinline__ void stage_range(T (&x)[kVals], int b0, int b1) {
  // b0/b1 are compile-time constants at every call site, so the branches fold
  if (b0 <= 0 && 0 < b1) stage<0>(x);
  if (b0 <= 1 && 1 < b1) stage<1>(x);
  if (b0 <= 2 && 2 < b1) stage<2>(x);
  if (b0 <= 3 && 3 < b1) stage<3>(x);
  if (b0 <= 4 && 4 < b1) stage<4>(x);
}

// Stage 0 of the transform with the sign flips and scaling, in Float64 as in
// the reference (x[i] * (signs[i] ? scaling : -scaling), then z + y, z - y).
template <class T>
__device__ __forceinline__ void stage0_signed(T (&x)[kVals], uint32_t sw, double sc) {
#pragma unroll
  for (int v = 0; v < kVals; v += 2) {
    const double s0 = ((sw >> v) & 1u) ? sc : -sc;
    const double s1 = ((sw >> (v + 1)) & 1u) ? sc : -sc;
    const double z = double(x[v]) * s0;
    const double y = double(x[v + 1]) * s1;
    x[v] = static_cast<T>(z + y);
    x[v + 1] = static_cast<T>(z - y);
  }
}

template <class T, int LC, int P>
struct Phases {
  static constexpr int NPH = (LC + 4) / 5;
  static constexpr int W = (5 * P < LC - 5) ? 5 * P : LC - 5;
  static constexpr int B0 = 5 * P - W;
  static constexpr int B1 = ((5 * P + 5 < LC) ? 5 * P + 5 : LC) - W;

  static __device__ __forceinline__ void run(T (&x)[kVals], T* lds, int t, uint32_t sw, double sc) {
#pragma unroll
    for (int v = 0; v < kVals; ++v) x[v] = lds[pad(map_entry<W>(t, v))];
    if (P == 0) {
      stage0_signed<T>(x, sw, sc);
      stage_range<T>(x, 1, B1);
    } else {
      stage_range<T>(x, B0, B1);
    }
    if constexpr (P + 1 < NPH) {
      __syncthreads();
#pragma unroll
      for (int v = 0; v < kVals; ++v) lds[pad(map_entry<W>(t, v))] = x[v];
      __syncthreads();
      Phases<T, LC, P + 1>::run(x, lds, t, sw, sc);
    }
  }
};

template <class T, int LC>
struct LastWindow {
  static constexpr int value = ((LC + 4) / 5 - 1) * 5 < LC - 5 ? ((LC + 4) / 5 - 1) * 5 : LC - 5;
};

template <class T>
__device__ inline void emit(T* __restrict__ Y, int64_t y_rs, int64_t ycol, int r, double val,
                            double alpha, double beta) {
  T* yp = Y + r * y_rs + ycol;
  *yp = static_cast<T>(beta == 0.0 ? alpha * val : beta * double(*yp) + alpha * val);
}

template <class T, int LC, int NCH>
__global__ __launch_bounds__(1 << (LC - 5)) void srht_fused_kernel(
    const T* __restrict__ X, int64_t x_cs, int64_t m, int64_t ncols,
    const uint32_t* __restrict__ signs32, double scale, const uint32_t* __restrict__ tmask,
    const uint32_t* __restrict__ tstart, const int32_t* __restrict__ rowmap,
    T* __restrict__ Y, int64_t y_rs, int64_t y_cs, double alpha, double beta, int vec_ok) {
  constexpr int TB = LC - 5;
  constexpr int CH = 1 << LC;
  constexpr int W = LastWindow<T, LC>::value;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  T* tbuf = lds + CH + CH / 32;
  const int t = threadIdx.x;
  const uint32_t my_mask = tmask[t];
  const int my_start = static_cast<int>(tstart[t]);
  T x[kVals];

  for (int64_t col = blockIdx.x; col < ncols; col += gridDim.x) {
    const T* xc = X + col * x_cs;
#pragma unroll 1
    for (int ch = 0; ch < NCH; ++ch) {
      const int64_t base = int64_t(ch) << LC;
      // coalesced load: register (vh, q) holds entry (vh << (TB+2)) | (t << 2) | q
#pragma unroll
      for (int vh = 0; vh < 8; ++vh) {
        const int e0 = (vh << (TB + 2)) | (t << 2);
        const int64_t pos = base + e0;
        if (vec_ok && pos + 3 < m) {
          if constexpr (sizeof(T) == 4) {
            const float4 q4 = *reinterpret_cast<const float4*>(xc + pos);
            x[4 * vh + 0] = q4.x;
            x[4 * vh + 1] = q4.y;
            x[4 * vh + 2] = q4.z;
            x[4 * vh + 3] = q4.w;
          } else {
            const double2 a = *reinterpret_cast<const double2*>(xc + pos);
            const double2 b = *reinterpret_cast<const double2*>(xc + pos + 2);
            x[4 * vh + 0] = a.x;
            x[4 * vh + 1] = a.y;
            x[4 * vh + 2] = b.x;
            x[4 * vh + 3] = b.y;
          }
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) x[4 * vh + q] = (pos + q < m) ? xc[pos + q] : T(0);
        }
      }
      const uint32_t sw = signs32[(base >> 5) + t];  // signs of entries 32t .. 32t+31
      __syncthreads();  // the previous chunk's readers of lds / tbuf are done
#pragma unroll
      for (int vh = 0; vh < 8; ++vh)
#pragma unroll
        for (int q = 0; q < 4; ++q) lds[pad((vh << (TB + 2)) | (t << 2) | q)] = x[4 * vh + q];
      __syncthreads();
      Phases<T, LC, 0>::run(x, lds, t, sw, scale);

      const int64_t ycol = col * y_cs;
      int slot = my_start;
      if constexpr (NCH == 1) {
#pragma unroll
        for (int v = 0; v < kVals; ++v) {
          if ((my_mask >> v) & 1u) {
            emit<T>(Y, y_rs, ycol, rowmap[slot], double(x[v]), alpha, beta);
            ++slot;
          }
        }
      } else {
        if (ch == 0) {
#pragma unroll
          for (int v = 0; v < kVals; ++v) {
            if ((my_mask >> v) & 1u) tbuf[slot++] = x[v];
          }
        } else {
#pragma unroll
          for (int v = 0; v < kVals; ++v) {
            if ((my_mask >> v) & 1u) {
              const T z = tbuf[slot], y = x[v];
              const int r0 = rowmap[2 * slot], r1 = rowmap[2 * slot + 1];
              if (r0 >= 0) emit<T>(Y, y_rs, ycol, r0, double(T(z + y)), alpha, beta);
              if (r1 >= 0) emit<T>(Y, y_rs, ycol, r1, double(T(z - y)), alpha, beta);
              ++slot;
            }
          }
        }
      }
    }
  }
}

// ---- plan ------------------------------------------------------------------
// layout (32-bit words): [0] magic, [1] LC, [2] NCH, [3] s, [4] n_pad lo,
// [5] n_pad hi, [6] elem bytes, [7] reserved; then tmask[2^(LC-5)],
// tstart[2^(LC-5)], rowmap[s*NCH].  tmask[t] bit v: register v of thread t
// holds a needed entry after the last phase; tstart[t]: slot of the first one
// (exclusive scan of the mask popcounts); rowmap[slot*NCH + c]: output row of
// position c*2^LC + e, or -1.

constexpr uint32_t kPlanMagic = 0x52484C50u;
constexpr int kHdr = 8;

static bool fused_shape(int elem, int64_t n_pad, int& LC, int& NCH) {
  const int L = ilog2(n_pad);
  const int lc_max = elem == 4 ? 15 : 14;
  if (L < 11 || L > lc_max + 1) return false;
  LC = std::min(L, lc_max);
  NCH = L > lc_max ? 2 : 1;
  return true;
}

static size_t plan_words(int LC, int NCH, int64_t s) {
  return kHdr + 2 * (size_t(1) << (LC - 5)) + static_cast<size_t>(s) * NCH;
}

// register window of the last phase (matches LastWindow<T, LC>)
static int last_window(int LC) { return std::min(((LC + 4) / 5 - 1) * 5, LC - 5); }

__global__ void plan_init_kernel(uint32_t* plan, int LC, int NCH, int64_t s, int64_t n_pad,
                                 int elem) {
  const int64_t nt = int64_t(1) << (LC - 5);
  const int64_t total = 2 * nt + s * NCH;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    plan[kHdr + i] = i < 2 * nt ? 0u : 0xFFFFFFFFu;  // rowmap = -1
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    plan[0] = kPlanMagic;
    plan[1] = LC;
    plan[2] = NCH;
    plan[3] = static_cast<uint32_t>(s);
    plan[4] = static_cast<uint32_t>(n_pad);
    plan[5] = static_cast<uint32_t>(n_pad >> 32);
    plan[6] = elem;
    plan[7] = 0;
  }
}

__device__ inline void owner_of(int e, int W, int& t, int& v) {
  v = (e >> W) & 31;
  t = ((e >> (W + 5)) << W) | (e & ((1 << W) - 1));
}

__global__ void plan_mark_kernel(uint32_t* tmask, const int32_t* __restrict__ op, int64_t s,
                                 int LC, int W) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r >= s) return;
  int t, v;
  owner_of(op[r] & ((1 << LC) - 1), W, t, v);
  atomicOr(&tmask[t], 1u << v);
}

__global__ __launch_bounds__(1024) void plan_prefix_kernel(const uint32_t* __restrict__ tmask,
                                                           uint32_t* tstart, int nt) {
  __shared__ uint32_t part[1024];
  const int per = (nt + 1023) / 1024;
  const int b = threadIdx.x * per;
  uint32_t s = 0;
  for (int i = b; i < b + per && i < nt; ++i) s += __popc(tmask[i]);
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
  for (int i = b; i < b + per && i < nt; ++i) {
    tstart[i] = run;
    run += __popc(tmask[i]);
  }
}

__global__ void plan_rowmap_kernel(const uint32_t* __restrict__ tmask,
                                   const uint32_t* __restrict__ tstart, int32_t* rowmap,
                                   const int32_t* __restrict__ op, int64_t s, int LC, int NCH,
                                   int W) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r >= s) return;
  const int64_t pos = op[r];
  int t, v;
  owner_of(static_cast<int>(pos & ((1 << LC) - 1)), W, t, v);
  const int slot = static_cast<int>(tstart[t]) + __popc(tmask[t] & ((1u << v) - 1u));
  rowmap[int64_t(slot) * NCH + (pos >> LC)] = static_cast<int32_t>(r);
}

static int build_plan(rla_ctx* ctx, int elem, int64_t n_pad, int64_t s, const int32_t* op,
                      uint32_t* plan) {
  int LC = 0, NCH = 0;
  if (!fused_shape(elem, n_pad, LC, NCH)) return RLA_OK;
  const int nt = 1 << (LC - 5);
  const int W = last_window(LC);
  hipStream_t st = ctx->stream;
  plan_init_kernel<<<64, 256, 0, st>>>(plan, LC, NCH, s, n_pad, elem);
  uint32_t* tmask = plan + kHdr;
  uint32_t* tstart = tmask + nt;
  int32_t* rowmap = reinterpret_cast<int32_t*>(tstart + nt);
  plan_mark_kernel<<<dim3(unsigned(ceil_div(s, 256))), 256, 0, st>>>(tmask, op, s, LC, W);
  plan_prefix_kernel<<<1, 1024, 0, st>>>(tmask, tstart, nt);
  plan_rowmap_kernel<<<dim3(unsigned(ceil_div(s, 256))), 256, 0, st>>>(tmask, tstart, rowmap,
                                                                      op, s, LC, NCH, W);
  RLA_LAUNCH_CHECK("srht_plan");
  return RLA_OK;
}

static int num_cus(int device) {
  static int cache[64] = {0};
  if (device >= 0 && device < 64 && cache[device]) return cache[device];
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 256;
  if (device >= 0 && device < 64) cache[device] = p.multiProcessorCount;
  return p.multiProcessorCount;
}

template <class T, int LC, int NCH>
static int launch_fused(rla_ctx* ctx, const T* X, int64_t x_cs, int64_t m, int64_t ncols,
                        const uint64_t* signs, double scale, const uint32_t* plan, int64_t s,
                        T* Y, int64_t y_rs, int64_t y_cs, double alpha, double beta) {
  const int nt = 1 << (LC - 5);
  const uint32_t* tmask = plan + kHdr;
  const uint32_t* tstart = tmask + nt;
  const int32_t* rowmap = reinterpret_cast<const int32_t*>(tstart + nt);
  const int64_t nslots = std::min<int64_t>(s, int64_t(1) << LC);
  const size_t lds = ((size_t(1) << LC) + (size_t(1) << (LC - 5))) * sizeof(T) +
                     (NCH > 1 ? nslots * sizeof(T) : 0);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(&srht_fused_kernel<T, LC, NCH>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  RLA_HIP(attr_err);
  const int threads = 1 << (LC - 5);
  // workgroups resident per CU, limited by LDS (160 KiB) and by 16 waves / CU
  const int per_cu = std::max<int>(1, std::min<int>(static_cast<int>((160 * 1024) / lds),
                                                    std::max(1, 1024 / threads)));
  const int64_t grid = std::min<int64_t>(ncols, int64_t(num_cus(ctx->device)) * per_cu);
  const int vec_ok = (x_cs % (sizeof(T) == 4 ? 4 : 2) == 0) &&
                     (reinterpret_cast<uintptr_t>(X) % 16 == 0);
  srht_fused_kernel<T, LC, NCH><<<dim3(static_cast<unsigned>(grid)), threads, lds,
                                  ctx->stream>>>(X, x_cs, m, ncols,
                                                 reinterpret_cast<const uint32_t*>(signs), scale,
                                                 tmask, tstart, rowmap, Y, y_rs, y_cs, alpha,
                                                 beta, vec_ok);
  RLA_LAUNCH_CHECK("srht_fused");
  return RLA_OK;
}

template <class T>
int srht_fused_forward(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                       const int32_t* op, const uint64_t* signs, double scale, const void* plan,
                       const T* X, int64_t x_rs, int64_t x_cs, T* Y, int64_t y_rs, int64_t y_cs,
                       double alpha, double beta, bool* handled) {
  *handled = false;
  int LC = 0, NCH = 0;
  if (x_rs != 1 || !fused_shape(sizeof(T), n_pad, LC, NCH)) return RLA_OK;
  // the slot buffer of a 2-chunk column must fit next to the chunk in LDS
  if (NCH == 2 && ((size_t(1) << LC) + (size_t(1) << (LC - 5)) +
                   size_t(std::min<int64_t>(s, 1 << LC))) * sizeof(T) > 160 * 1024)
    return RLA_OK;
  const uint32_t* pl = static_cast<const uint32_t*>(plan);
  if (!pl) {
    void* ws = nullptr;
    int rc = workspace(ctx, plan_words(LC, NCH, s) * 4, &ws);
    if (rc) return rc;
    rc = build_plan(ctx, sizeof(T), n_pad, s, op, static_cast<uint32_t*>(ws));
    if (rc) return rc;
    pl = static_cast<const uint32_t*>(ws);
  }
  *handled = true;
#define RLA_FUSED(LCV, NCHV)                                                              \
  if (LC == LCV && NCH == NCHV)                                                           \
    return launch_fused<T, LCV, NCHV>(ctx, X, x_cs, m, ncols, signs, scale, pl, s, Y, y_rs, \
                                      y_cs, alpha, beta);
  if constexpr (sizeof(T) == 4) {
    RLA_FUSED(11, 1) RLA_FUSED(12, 1) RLA_FUSED(13, 1) RLA_FUSED(14, 1) RLA_FUSED(15, 1)
    RLA_FUSED(15, 2)
  } else {
    RLA_FUSED(11, 1) RLA_FUSED(12, 1) RLA_FUSED(13, 1) RLA_FUSED(14, 1) RLA_FUSED(14, 2)
  }
#undef RLA_FUSED
  *handled = false;
  return RLA_OK;
}

template int srht_fused_forward<float>(rla_ctx*, int64_t, int64_t, int64_t, int64_t,
                                       const int32_t*, const uint64_t*, double, const void*,
                                       const float*, int64_t, int64_t, float*, int64_t, int64_t,
                                       double, double, bool*);
template int srht_fused_forward<double>(rla_ctx*, int64_t, int64_t, int64_t, int64_t,
                                        const int32_t*, const uint64_t*, double, const void*,
                                        const double*, int64_t, int64_t, double*, int64_t,
                                        int64_t, double, double, bool*);

}  // namespace rla

using namespace rla;

extern "C" {

int rla_srht_plan_bytes(int elem_bytes, int64_t n_pad, int64_t s, size_t* bytes) {
  if (!bytes) return fail(RLA_EARG, "rla_srht_plan_bytes: bytes is NULL");
  if (elem_bytes != 4 && elem_bytes != 8) return fail(RLA_EARG, "elem_bytes must be 4 or 8");
  int LC = 0, NCH = 0;
  *bytes = fused_shape(elem_bytes, n_pad, LC, NCH) ? plan_words(LC, NCH, s) * 4 : 0;
  return RLA_OK;
}

int rla_srht_plan(rla_ctx* ctx, int elem_bytes, int64_t n_pad, int64_t s, const int32_t* op,
                  void* plan) {
  RLA_CHECK_CTX(ctx);
  if (elem_bytes != 4 && elem_bytes != 8) return fail(RLA_EARG, "elem_bytes must be 4 or 8");
  if (!is_pow2(n_pad) || s < 1 || s > n_pad) return fail(RLA_EARG, "bad SRHT plan shape");
  int LC = 0, NCH = 0;
  if (!fused_shape(elem_bytes, n_pad, LC, NCH)) return RLA_OK;  // nothing to plan
  if (!plan || !op) return fail(RLA_EARG, "rla_srht_plan: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  return build_plan(ctx, elem_bytes, n_pad, s, op, static_cast<uint32_t*>(plan));
}

}  // extern "C"
