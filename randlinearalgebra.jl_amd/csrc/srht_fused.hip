// Single-pass SRHT forward operator: Y = alpha*(scale*H*D*pad(X))[op, :] + beta*Y.
//
// One workgroup transforms one column at a time (persistent over columns).
// A column of n_pad = 2^L entries is processed as NCH chunks of 2^LC
// contiguous entries (NCH = 1 or 2).  Inside a chunk every thread holds
// R = 2^VB entries in registers; the LC butterfly stages run in phases of VB
// bits, each phase in registers, and between phases the chunk is
// re-distributed through LDS so that the next VB bits become register bits.
// The stages are the reference's stages in the reference's order
// (fwht.jl:21-70), so results are bit-identical to it.
//
// Work items (column, chunk) are processed in order by each workgroup; the
// global loads of the next item are issued into a second register set before
// the current item is transformed, so HBM streaming overlaps the LDS/ALU work.
//
// Only the sampled outputs are needed.  With several chunks the last stages
// (bits LC, LC+1, ...) combine entry e of every chunk.  A per-recipe plan
// lists the "needed" positions e (e + c*2^LC sampled for some chunk c) in
// ascending order as slots, and maps each (slot, c) to its output row.  After
// a chunk's last phase the chunk goes back to LDS and the threads walk the
// slots: earlier chunks park their slot values (in LDS when it fits, else in
// a per-workgroup global scratch that the same thread reads back), the last
// chunk finishes the remaining stages per slot and writes the rows.  Global
// traffic is one read of X, one write of the s rows, plus the parked slots.
//
// LDS layout: entry e at e + (e >> 5) (one pad word per 32 entries).  Every
// register<->LDS move is then thread base + compile-time offset, with at most
// two-way bank conflicts on any of the access patterns.  The slot buffer
// follows the chunk.
#pragma clang fp contract(off)

#include <algorithm>
#include <mutex>

#include "rla_internal.hpp"

namespace rla {

constexpr int kPlanHdr = 8;  // words before the plan's bitmap

__device__ __forceinline__ int pad(int e) { return e + (e >> 5); }

// entry index of register v of thread t when the register window is [W, W+VB)
template <int W, int VB>
__device__ __forceinline__ int map_entry(int t, int v) {
  return ((t >> W) << (W + VB)) | (v << W) | (t & ((1 << W) - 1));
}

// pad(map_entry<W, VB>(t, v)) split as thread base + compile-time offset, so
// every LDS access is one base register plus an immediate.
template <int W, int VB>
__device__ __forceinline__ int lds_base(int t) {
  if constexpr (W == 0) {
    return (t << VB) + (t << (VB - 5));
  } else {
    static_assert(W >= 5, "register windows start at bit 0 or at bit >= 5");
    const int hi = t >> W, lo = t & ((1 << W) - 1);
    return (hi << (W + VB)) + lo + (hi << (W + VB - 5)) + (lo >> 5);
  }
}
template <int W, int VB>
__host__ __device__ constexpr int lds_off(int v) {
  return W == 0 ? v + (v >> 5) : v * ((1 << W) + (1 << (W - 5)));
}
// load-order mapping: entry (vh << (TB+2)) | (t << 2) | q
template <int TB>
__device__ __forceinline__ int load_base(int t) { return (t << 2) + (t >> 3); }
template <int TB>
__host__ __device__ constexpr int load_off(int vh, int q) {
  return vh * ((1 << (TB + 2)) + (1 << (TB - 3))) + q;
}

template <int RB, int R, class T>
__device__ __forceinline__ void stage(T (&x)[R]) {
#pragma unroll
  for (int v = 0; v < R; ++v) {
    if (!(v & (1 << RB))) {
      const T z = x[v], y = x[v | (1 << RB)];
      x[v] = z + y;
      x[v | (1 << RB)] = z - y;
    }
  }
}

template <int B0, int B1, int R, class T>
__device__ __forceinline__ void stage_range(T (&x)[R]) {
  if constexpr (B0 <= 0 && 0 < B1) stage<0, R>(x);
  if constexpr (B0 <= 1 && 1 < B1) stage<1, R>(x);
  if constexpr (B0 <= 2 && 2 < B1) stage<2, R>(x);
  if constexpr (B0 <= 3 && 3 < B1) stage<3, R>(x);
  if constexpr (B0 <= 4 && 4 < B1) stage<4, R>(x);
  if constexpr (B0 <= 5 && 5 < B1) stage<5, R>(x);
}

__device__ __forceinline__ float flip_sign(float a, uint32_t neg) {
  return __uint_as_float(__float_as_uint(a) ^ (neg << 31));
}
__device__ __forceinline__ double flip_sign(double a, uint32_t neg) {
  return __longlong_as_double(__double_as_longlong(a) ^ (static_cast<long long>(neg) << 63));
}

// Stage 0 of the transform with the sign flips and scaling, in Float64 as in
// the reference (x[i] * (signs[i] ? scaling : -scaling), then z + y, z - y).
// The sign is applied to x (exact) so the product is x' * scaling.
template <int R, class T>
__device__ __forceinline__ void stage0_signed(T (&x)[R], uint64_t sw, double sc) {
  const uint64_t neg = ~sw;
#pragma unroll
  for (int v = 0; v < R; v += 2) {
    const double z = double(flip_sign(x[v], uint32_t(neg >> v) & 1u)) * sc;
    const double y = double(flip_sign(x[v + 1], uint32_t(neg >> (v + 1)) & 1u)) * sc;
    x[v] = static_cast<T>(z + y);
    x[v + 1] = static_cast<T>(z - y);
  }
}

template <int LC, int VB>
struct Geometry {
  static constexpr int R = 1 << VB;
  static constexpr int TB = LC - VB;
  static constexpr int NPH = (LC + VB - 1) / VB;
  static constexpr int window(int p) { return VB * p < LC - VB ? VB * p : LC - VB; }
  static constexpr int WLAST = window(NPH - 1);
};

template <class T, int LC, int VB, int P>
struct Phases {
  using G = Geometry<LC, VB>;
  static constexpr int W = G::window(P);
  static constexpr int B0 = VB * P - W;
  static constexpr int B1 = ((VB * P + VB < LC) ? VB * P + VB : LC) - W;

  static __device__ __forceinline__ void run(T (&x)[G::R], T* lds, int t, uint64_t sw,
                                             double sc) {
    T* lb = lds + lds_base<W, VB>(t);
#pragma unroll
    for (int v = 0; v < G::R; ++v) x[v] = lb[lds_off<W, VB>(v)];
    if constexpr (P == 0) {
      stage0_signed<G::R>(x, sw, sc);
      stage_range<1, B1, G::R>(x);
    } else {
      stage_range<B0, B1, G::R>(x);
    }
    if constexpr (P + 1 < G::NPH) {
      __syncthreads();
#pragma unroll
      for (int v = 0; v < G::R; ++v) lb[lds_off<W, VB>(v)] = x[v];
      __syncthreads();
      Phases<T, LC, VB, P + 1>::run(x, lds, t, sw, sc);
    }
  }
};

template <class T>
__device__ __forceinline__ void emit(T* __restrict__ Y, int64_t y_rs, int64_t ycol, int r,
                                     double val, double alpha, double beta) {
  T* yp = Y + r * y_rs + ycol;
  *yp = static_cast<T>(beta == 0.0 ? alpha * val : beta * double(*yp) + alpha * val);
}

// Loads the R entries of chunk `base` of column xc in load order:
// register 4*vh + q holds entry (vh << (TB+2)) | (t << 2) | q.
template <class T, int R, int TB>
__device__ __forceinline__ void load_chunk(T (&x)[R], const T* __restrict__ xc, int64_t base,
                                           int64_t m, int t, int vec_ok) {
  constexpr int64_t CH = int64_t(R) << TB;
  if (vec_ok && base + CH <= m) {
    // whole chunk inside the matrix: unguarded vector loads
    const T* p = xc + base + (t << 2);
#pragma unroll
    for (int vh = 0; vh < R / 4; ++vh) {
      if constexpr (sizeof(T) == 4) {
        typedef float v4f __attribute__((ext_vector_type(4)));
        const v4f q4 = __builtin_nontemporal_load(
            reinterpret_cast<const v4f*>(p + (int64_t(vh) << (TB + 2))));
        x[4 * vh + 0] = q4.x;
        x[4 * vh + 1] = q4.y;
        x[4 * vh + 2] = q4.z;
        x[4 * vh + 3] = q4.w;
      } else {
        const double2 a = *reinterpret_cast<const double2*>(p + (int64_t(vh) << (TB + 2)));
        const double2 b = *reinterpret_cast<const double2*>(p + (int64_t(vh) << (TB + 2)) + 2);
        x[4 * vh + 0] = a.x;
        x[4 * vh + 1] = a.y;
        x[4 * vh + 2] = b.x;
        x[4 * vh + 3] = b.y;
      }
    }
    return;
  }
#pragma unroll
  for (int vh = 0; vh < R / 4; ++vh) {
    const int64_t pos = base + ((vh << (TB + 2)) | (t << 2));
#pragma unroll
    for (int q = 0; q < 4; ++q) x[4 * vh + q] = (pos + q < m) ? xc[pos + q] : T(0);
  }
}

template <class T, int LC, int NCH, int VB, bool PF, bool SCR_LDS>
__global__ __launch_bounds__(1 << (LC - VB)) void srht_fused_kernel(
    const T* __restrict__ X, int64_t x_cs, int64_t m, int64_t ncols,
    const uint32_t* __restrict__ signs32, double scale, const uint32_t* __restrict__ plan,
    T* __restrict__ Y, int64_t y_rs, int64_t y_cs, double alpha, double beta, int vec_ok,
    T* __restrict__ gscr, int nslots_cap) {
  using G = Geometry<LC, VB>;
  constexpr int R = G::R;
  constexpr int TB = G::TB;
  constexpr int NT = 1 << TB;
  constexpr int CH = 1 << LC;
  constexpr int WL = G::WLAST;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const int t = threadIdx.x;
  const int nslots = static_cast<int>(plan[7]);
  const int32_t* __restrict__ slot_pos =
      reinterpret_cast<const int32_t*>(plan + kPlanHdr + 2 * (CH >> 5));
  const int32_t* __restrict__ rowmap = slot_pos + nslots_cap;
  // values parked by chunks 0 .. NCH-2, tb[c * nslots_cap + slot]; a slot is
  // always handled by thread slot % NT, which reads back only its own writes
  T* tb = SCR_LDS ? lds + CH + CH / 32
                  : gscr + int64_t(blockIdx.x) * (NCH > 1 ? NCH - 1 : 0) * nslots_cap;
  T x[R];
  T nx[PF ? R : 1];

  // This workgroup's items: columns blockIdx.x, blockIdx.x + gridDim.x, ...,
  // each as its NCH chunks in order.
  const int64_t my_cols = blockIdx.x < ncols ? (ncols - 1 - blockIdx.x) / gridDim.x + 1 : 0;
  const int64_t my_items = my_cols * NCH;
  auto item_col = [&](int64_t k) { return blockIdx.x + (k / NCH) * int64_t(gridDim.x); };
  if constexpr (PF) {
    if (my_items > 0) load_chunk<T, R, TB>(nx, X + item_col(0) * x_cs, 0, m, t, vec_ok);
  }
  for (int64_t k = 0; k < my_items; ++k) {
    const int64_t col = item_col(k);
    const int ch = static_cast<int>(k % NCH);
    const int64_t base = int64_t(ch) << LC;
    if constexpr (PF) {
#pragma unroll
      for (int v = 0; v < R; ++v) x[v] = nx[v];
    } else {
      load_chunk<T, R, TB>(x, X + col * x_cs, base, m, t, vec_ok);
    }
    // signs of entries R*t .. R*t+R-1 (phase-0 ownership)
    uint64_t sw;
    if constexpr (R == 64) {
      const uint32_t* sp = signs32 + (base >> 5) + 2 * t;
      sw = uint64_t(sp[0]) | (uint64_t(sp[1]) << 32);
    } else {
      sw = signs32[(base >> 5) + t];
    }
    __syncthreads();  // the previous item's readers of lds are done
    {
      T* lb = lds + load_base<TB>(t);
#pragma unroll
      for (int vh = 0; vh < R / 4; ++vh)
#pragma unroll
        for (int q = 0; q < 4; ++q) lb[load_off<TB>(vh, q)] = x[4 * vh + q];
    }
    if constexpr (PF) {
      if (k + 1 < my_items)
        load_chunk<T, R, TB>(nx, X + item_col(k + 1) * x_cs, int64_t((k + 1) % NCH) << LC, m, t,
                             vec_ok);
    }
    __syncthreads();
    Phases<T, LC, VB, 0>::run(x, lds, t, sw, scale);
    // the transformed chunk back to LDS, then slot-parallel output
    __syncthreads();
    {
      T* lb = lds + lds_base<WL, VB>(t);
#pragma unroll
      for (int v = 0; v < R; ++v) lb[lds_off<WL, VB>(v)] = x[v];
    }
    __syncthreads();
    const int64_t ycol = col * y_cs;
    if constexpr (NCH == 1) {
      for (int sl = t; sl < nslots; sl += NT)
        emit<T>(Y, y_rs, ycol, rowmap[sl], double(lds[pad(slot_pos[sl])]), alpha, beta);
    } else {
      if (ch < NCH - 1) {
        T* tc = tb + int64_t(ch) * nslots_cap;
        for (int sl = t; sl < nslots; sl += NT) tc[sl] = lds[pad(slot_pos[sl])];
      } else {
        for (int sl = t; sl < nslots; sl += NT) {
          T a[NCH];
#pragma unroll
          for (int c = 0; c < NCH - 1; ++c) a[c] = tb[int64_t(c) * nslots_cap + sl];
          a[NCH - 1] = lds[pad(slot_pos[sl])];
          // the stages of bits LC, LC+1, ... pair chunks c and c|h, in order
#pragma unroll
          for (int h = 1; h < NCH; h <<= 1)
#pragma unroll
            for (int c = 0; c < NCH; ++c)
              if (!(c & h)) {
                const T z = a[c], y = a[c | h];
                a[c] = z + y;
                a[c | h] = z - y;
              }
#pragma unroll
          for (int c = 0; c < NCH; ++c) {
            const int r = rowmap[int64_t(sl) * NCH + c];
            if (r >= 0) emit<T>(Y, y_rs, ycol, r, double(a[c]), alpha, beta);
          }
        }
      }
    }
  }
}

// The same transform for NCH > 1 with the parked slot values in registers:
// thread t handles slots t, t + NT, ..., t + (MS-1)*NT in every chunk of every
// column, so the values chunks 0 .. NCH-2 leave for the cross-chunk stages are
// only ever read back by the thread that parked them.  The slot positions are
// read from the plan once per kernel, the output rows once per column.
// Needs nslots <= MS * NT.
template <class T, int LC, int NCH, int VB, int MS>
__global__ __launch_bounds__(1 << (LC - VB)) __attribute__((amdgpu_waves_per_eu(4, 4)))
void srht_fused_reg_kernel(
    const T* __restrict__ X, int64_t x_cs, int64_t m, int64_t ncols,
    const uint32_t* __restrict__ signs32, double scale, const uint32_t* __restrict__ plan,
    T* __restrict__ Y, int64_t y_rs, int64_t y_cs, double alpha, double beta, int vec_ok,
    int nslots_cap) {
  using G = Geometry<LC, VB>;
  constexpr int R = G::R;
  constexpr int TB = G::TB;
  constexpr int NT = 1 << TB;
  constexpr int WL = G::WLAST;
  static_assert(NCH > 1, "single-chunk columns need no parking");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const int t = threadIdx.x;
  const int nslots = static_cast<int>(plan[7]);
  const int32_t* __restrict__ slot_pos =
      reinterpret_cast<const int32_t*>(plan + kPlanHdr + 2 * ((1 << LC) >> 5));
  const int32_t* __restrict__ rowmap = slot_pos + nslots_cap;
  int spad[MS];  // padded LDS index of each of this thread's slots (-1: none)
#pragma unroll
  for (int j = 0; j < MS; ++j) {
    const int sl = t + j * NT;
    spad[j] = sl < nslots ? pad(slot_pos[sl]) : -1;
  }
  T park[NCH - 1][MS];
  T x[R], nx[R];
  const int64_t my_cols = blockIdx.x < ncols ? (ncols - 1 - blockIdx.x) / gridDim.x + 1 : 0;
  const int64_t my_items = my_cols * NCH;
  auto item_col = [&](int64_t k) { return blockIdx.x + (k / NCH) * int64_t(gridDim.x); };
  if (my_items > 0) load_chunk<T, R, TB>(nx, X + item_col(0) * x_cs, 0, m, t, vec_ok);
  for (int64_t k = 0; k < my_items; ++k) {
    const int64_t col = item_col(k);
    const int ch = static_cast<int>(k % NCH);
    const int64_t base = int64_t(ch) << LC;
#pragma unroll
    for (int v = 0; v < R; ++v) x[v] = nx[v];
    uint64_t sw;
    if constexpr (R == 64) {
      const uint32_t* sp = signs32 + (base >> 5) + 2 * t;
      sw = uint64_t(sp[0]) | (uint64_t(sp[1]) << 32);
    } else {
      sw = signs32[(base >> 5) + t];
    }
    __syncthreads();  // the previous item's readers of lds are done
    {
      T* lb = lds + load_base<TB>(t);
#pragma unroll
      for (int vh = 0; vh < R / 4; ++vh)
#pragma unroll
        for (int q = 0; q < 4; ++q) lb[load_off<TB>(vh, q)] = x[4 * vh + q];
    }
    if (k + 1 < my_items)
      load_chunk<T, R, TB>(nx, X + item_col(k + 1) * x_cs, int64_t((k + 1) % NCH) << LC, m, t,
                           vec_ok);
    __syncthreads();
    Phases<T, LC, VB, 0>::run(x, lds, t, sw, scale);
    __syncthreads();
    {
      T* lb = lds + lds_base<WL, VB>(t);
#pragma unroll
      for (int v = 0; v < R; ++v) lb[lds_off<WL, VB>(v)] = x[v];
    }
    __syncthreads();
    if (ch < NCH - 1) {
      // park this chunk's slot values (selects on the uniform chunk index)
#pragma unroll
      for (int j = 0; j < MS; ++j) {
        const T v = spad[j] >= 0 ? lds[spad[j]] : T(0);
#pragma unroll
        for (int c = 0; c < NCH - 1; ++c)
          if (c == ch) park[c][j] = v;
      }
    } else {
      const int64_t ycol = col * y_cs;
#pragma unroll
      for (int j = 0; j < MS; ++j) {
        if (spad[j] < 0) continue;
        T a[NCH];
#pragma unroll
        for (int c = 0; c < NCH - 1; ++c) a[c] = park[c][j];
        a[NCH - 1] = lds[spad[j]];
        // the stages of bits LC, LC+1, ... pair chunks c and c|h, in order
#pragma unroll
        for (int h = 1; h < NCH; h <<= 1)
#pragma unroll
          for (int c = 0; c < NCH; ++c)
            if (!(c & h)) {
              const T z = a[c], y = a[c | h];
              a[c] = z + y;
              a[c | h] = z - y;
            }
        const int32_t* rmj = rowmap + int64_t(t + j * NT) * NCH;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          const int r = rmj[c];
          if (r >= 0) emit<T>(Y, y_rs, ycol, r, double(a[c]), alpha, beta);
        }
      }
    }
  }
}

// ---- wave-local Float32 transform ---------------------------------------------
// Chunks of 2^14 Float32 entries, 8 wavefronts of 64 lanes x 32 registers.
// Wavefront w owns entries w*2^11 .. w*2^11 + 2^11 - 1 of the chunk and runs
// stages 0..8 on them without LDS:
//   * registers r0, r1 hold entry bits e0, e1 (one float4 per load), r2..r4
//     hold e6, e7, e8 (the load index);
//   * lane bits 0..3 hold e2..e5 through a linear map chosen so that the
//     partner of entry bit e(2+i) is lane ^ {1, 2, 7, 15}[i]: the DPP patterns
//     quad_perm, row_half_mirror and row_mirror;
//   * lane bits 4, 5 hold e9, e10 and the wavefront index e11..e13.
// One DPP stage is one v_fmac_f32 per value: w = c * w[partner] + w with c =
// +1 on the lane that keeps z + y and -1 on the lane that keeps z - y, which
// then holds -(z - y).  A lane's values therefore carry the sign
// (-1)^popcount(e2..e5), the same for both entries of every later butterfly;
// it is removed when the sampled rows are written.  fl(-a + b) = -fl(a - b),
// so values are those of the reference order (an exact zero may change sign).
// Stages 9..13 run in registers after one LDS exchange (each thread then
// holds e9..e13); the cross-chunk stages and the parked slots are those of
// srht_fused_reg_kernel.
//
// LDS layout: entry e at e + (e >> 5) + 2 * (e >> 9).  The extra two words per
// 512 entries put e9 = 1 sixteen-plus-two banks away, so the exchange's
// stores (lanes spread over e2..e5 and e9) and loads (lanes over e0..e4) are
// free of bank conflicts; the map is linear in the entry bits, so every access
// is a thread base plus a compile-time offset.

// entry bits e2..e5 held by lane nibble l (inverse of the DPP lane map)
__device__ __forceinline__ int wl_nibble(int l) {
  const int l0 = l & 1, l1 = (l >> 1) & 1, l2 = (l >> 2) & 1, l3 = (l >> 3) & 1;
  return (l0 ^ l2) | ((l1 ^ l2) << 1) | ((l2 ^ l3) << 2) | (l3 << 3);
}

// w[v] = c * w[v](partner lane) + w[v] for 8 registers.  The leading s_nop
// gives the two wait states a DPP read needs after a VALU write of the same
// register; inside the block every register is written long before it is
// read again through DPP.
#define RLA_DPP_FMAC8(CTRL, p, c)                                                       \
  asm("s_nop 1\n\t"                                                                      \
      "v_fmac_f32_dpp %0, %0, %8 " CTRL "\n\t"                                           \
      "v_fmac_f32_dpp %1, %1, %8 " CTRL "\n\t"                                           \
      "v_fmac_f32_dpp %2, %2, %8 " CTRL "\n\t"                                           \
      "v_fmac_f32_dpp %3, %3, %8 " CTRL "\n\t"                                           \
      "v_fmac_f32_dpp %4, %4, %8 " CTRL "\n\t"                                           \
      "v_fmac_f32_dpp %5, %5, %8 " CTRL "\n\t"                                           \
      "v_fmac_f32_dpp %6, %6, %8 " CTRL "\n\t"                                           \
      "v_fmac_f32_dpp %7, %7, %8 " CTRL                                                  \
      : "+v"(p[0]), "+v"(p[1]), "+v"(p[2]), "+v"(p[3]), "+v"(p[4]), "+v"(p[5]), "+v"(p[6]), \
        "+v"(p[7])                                                                       \
      : "v"(c))

template <int I>
__device__ __forceinline__ void wl_dpp_stage(float (&x)[32], float c) {
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    float* p = x + 8 * g;
    if constexpr (I == 0) RLA_DPP_FMAC8("quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf", p, c);
    if constexpr (I == 1) RLA_DPP_FMAC8("quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf", p, c);
    if constexpr (I == 2) RLA_DPP_FMAC8("row_half_mirror row_mask:0xf bank_mask:0xf", p, c);
    if constexpr (I == 3) RLA_DPP_FMAC8("row_mirror row_mask:0xf bank_mask:0xf", p, c);
  }
}
#undef RLA_DPP_FMAC8

// chunk-local entry offset of register 4k+q of thread t at load time is
// wl_load_off(t) + 64k + q
__device__ __forceinline__ int wl_load_off(int t) {
  const int w = t >> 6, l = t & 63;
  return (w << 11) | (((l >> 5) & 1) << 10) | (((l >> 4) & 1) << 9) | (wl_nibble(l & 15) << 2);
}
__device__ __forceinline__ int wl_pad(int e) { return e + (e >> 5) + 2 * (e >> 9); }
// LDS words of a chunk in the wave-local layout
constexpr int kWaveLdsWords = 16384 + 512 + 64;

// Loads of one chunk in the wave-local order.  MODE 0: the chunk lies inside
// the matrix and columns are 16-byte aligned; MODE 1: aligned with m % 4 == 0,
// float4 loads clamped into the column and zeroed past row m; MODE 2: scalar
// loads, clamped and zeroed.  Every load stays inside the column.
template <int MODE>
__device__ __forceinline__ void wl_load(float (&x)[32], const float* __restrict__ xc,
                                        int64_t base, int64_t m, int off) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  const int e0 = static_cast<int>(base) + off;
  if constexpr (MODE == 0) {
    const float* p = xc + e0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const v4f q4 = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p + 64 * k));
      x[4 * k + 0] = q4.x;
      x[4 * k + 1] = q4.y;
      x[4 * k + 2] = q4.z;
      x[4 * k + 3] = q4.w;
    }
  } else if constexpr (MODE == 1) {
    const int mi = static_cast<int>(m);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int e = e0 + 64 * k;
      const v4f q4 = __builtin_nontemporal_load(
          reinterpret_cast<const v4f*>(xc + (e < mi ? e : mi - 4)));
      const bool in = e < mi;
      x[4 * k + 0] = in ? q4.x : 0.0f;
      x[4 * k + 1] = in ? q4.y : 0.0f;
      x[4 * k + 2] = in ? q4.z : 0.0f;
      x[4 * k + 3] = in ? q4.w : 0.0f;
    }
  } else {
    const int mi = static_cast<int>(m);
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = e0 + 64 * k + q;
        const float v = xc[e < mi ? e : mi - 1];
        x[4 * k + q] = e < mi ? v : 0.0f;
      }
  }
}

template <int NCH, int MS, int MODE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4)))
void srht_wave_kernel(
    const float* __restrict__ X, int64_t x_cs, int64_t m, int64_t ncols,
    const uint32_t* __restrict__ signs32, double scale, const uint32_t* __restrict__ plan,
    float* __restrict__ Y, int64_t y_rs, int64_t y_cs, double alpha, double beta,
    int nslots_cap) {
  constexpr int LC = 14;
  constexpr int NT = 512;
  static_assert(NCH == 2 || NCH == 4, "columns of 2 or 4 chunks");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* lds = reinterpret_cast<float*>(smem_raw);
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int nslots = static_cast<int>(plan[7]);
  const int32_t* __restrict__ slot_pos =
      reinterpret_cast<const int32_t*>(plan + kPlanHdr + 2 * ((1 << LC) >> 5));
  const int32_t* __restrict__ rowmap = slot_pos + nslots_cap;
  int spad[MS];      // padded LDS index of each of this thread's slots (-1: none)
  bool sneg[MS];     // slot carries the lane sign (-1)^popcount(e2..e5)
#pragma unroll
  for (int j = 0; j < MS; ++j) {
    const int sl = t + j * NT;
    const int e = sl < nslots ? slot_pos[sl] : 0;
    spad[j] = sl < nslots ? wl_pad(e) : -1;
    sneg[j] = __builtin_popcount((e >> 2) & 15) & 1;
  }
  // DPP coefficients: -1 where this lane keeps the difference
  const int nib = wl_nibble(lane & 15);
  const float c0 = (nib & 1) ? -1.0f : 1.0f, c1 = (nib & 2) ? -1.0f : 1.0f;
  const float c2 = (nib & 4) ? -1.0f : 1.0f, c3 = (nib & 8) ? -1.0f : 1.0f;
  const int off = wl_load_off(t);
  // sign bits of this thread's 32 entries in each chunk: bit 4k+q <- entry off+64k+q
  uint32_t swc[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    uint32_t sw = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      sw |= ((signs32[(ch << (LC - 5)) + (off >> 5) + 2 * k] >> (off & 31)) & 15u) << (4 * k);
    swc[ch] = sw;
  }
  float* lb_a = lds + wl_pad(off);  // exchange stores: register v at + (v & 3) + 66 (v >> 2)
  float* lb_b = lds + wl_pad(t);    // exchange loads: register v holds entry t + (v << 9)

  float park[NCH - 1][MS];
  float xa[32], xb[32];
  const int64_t my_cols = blockIdx.x < ncols ? (ncols - 1 - blockIdx.x) / gridDim.x + 1 : 0;
  if (my_cols > 0) wl_load<MODE>(xa, X + int64_t(blockIdx.x) * x_cs, 0, m, off);

  auto chunk = [&](auto chc, float(&x)[32], float(&nx)[32], int64_t col, bool more) {
    constexpr int ch = decltype(chc)::value;
    // next item: the next chunk of this column, or chunk 0 of the next column.
    // The offset is opaque per chunk so that the per-load clamps and masks of
    // the guarded modes are not hoisted out of the column loop.
    int offo = off;
    asm volatile("" : "+v"(offo));
    if (ch + 1 < NCH)
      wl_load<MODE>(nx, X + col * x_cs, int64_t(ch + 1) << LC, m, offo);
    else if (more)
      wl_load<MODE>(nx, X + (col + gridDim.x) * x_cs, 0, m, offo);
    // opaque per chunk, so the 32 per-value sign masks are not hoisted out of
    // the column loop (they would not fit in registers for every chunk)
    uint32_t sw = swc[ch];
    asm volatile("" : "+v"(sw));
    stage0_signed<32>(x, sw, scale);
    stage<1, 32>(x);
    wl_dpp_stage<0>(x, c0);
    wl_dpp_stage<1>(x, c1);
    wl_dpp_stage<2>(x, c2);
    wl_dpp_stage<3>(x, c3);
    stage<2, 32>(x);
    stage<3, 32>(x);
    stage<4, 32>(x);
    __syncthreads();  // the previous chunk's slot readers are done
#pragma unroll
    for (int v = 0; v < 32; ++v) lb_a[(v & 3) + 66 * (v >> 2)] = x[v];
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 32; ++v) x[v] = lb_b[530 * v];
    stage_range<0, 5, 32>(x);
#pragma unroll
    for (int v = 0; v < 32; ++v) lb_b[530 * v] = x[v];
    __syncthreads();
    if constexpr (ch < NCH - 1) {
#pragma unroll
      for (int j = 0; j < MS; ++j) park[ch][j] = spad[j] >= 0 ? lds[spad[j]] : 0.0f;
    } else {
      const int64_t ycol = col * y_cs;
#pragma unroll
      for (int j = 0; j < MS; ++j) {
        if (spad[j] < 0) continue;
        float a[NCH];
#pragma unroll
        for (int c = 0; c < NCH - 1; ++c) a[c] = park[c][j];
        a[NCH - 1] = lds[spad[j]];
#pragma unroll
        for (int h = 1; h < NCH; h <<= 1)
#pragma unroll
          for (int c = 0; c < NCH; ++c)
            if (!(c & h)) {
              const float z = a[c], y = a[c | h];
              a[c] = z + y;
              a[c | h] = z - y;
            }
        const double aj = sneg[j] ? -alpha : alpha;
        // opaque base: keeps the 4*MS row-map addresses out of the column loop
        const int32_t* rm = rowmap;
        asm volatile("" : "+s"(rm));
        const int32_t* rmj = rm + (t + j * NT) * NCH;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          const int r = rmj[c];
          if (r >= 0) emit<float>(Y, y_rs, ycol, r, double(a[c]), aj, beta);
        }
      }
    }
  };

  for (int64_t kc = 0; kc < my_cols; ++kc) {
    const int64_t col = blockIdx.x + kc * int64_t(gridDim.x);
    const bool more = kc + 1 < my_cols;
    // even chunks use xa, odd chunks xb; each prefetches into the other set
    chunk(std::integral_constant<int, 0>{}, xa, xb, col, more);
    chunk(std::integral_constant<int, 1>{}, xb, xa, col, more);
    if constexpr (NCH == 4) {
      chunk(std::integral_constant<int, 2 % NCH>{}, xa, xb, col, more);
      chunk(std::integral_constant<int, 3 % NCH>{}, xb, xa, col, more);
    }
  }
}

// ---- pruned Float32 SRHT: the tolerance path --------------------------------
// Y = alpha * (scale * H * D * pad(X))[op, :] + beta * Y for Float32 columns of
// n_pad = 2^L >= 2^14 entries, within the Float32 tolerance of the reference
// (the stages run in another order and the sampled outputs are summed
// differently, so bits may differ).  Only the s sampled rows of H D x are
// formed: with entry e = (h, l), h = e >> 11 its 2048-entry block and
// l = e & 2047, and a sampled row rho = (rho_h, rho_l),
//   (H D x)[rho] = sum_h (-1)^popcount(rho_h & h) * T_h[rho_l],
//   T_h = H_2048 D_h x_h  (the transform of block h alone).
// A wavefront transforms one block completely in registers (11 stages):
// registers hold entry bits e0, e1 (one float4 load) and e6..e8 (the load
// index), lanes hold e2..e5 through the DPP lane map of srht_wave_kernel and
// e9, e10; the signs are applied with one v_cndmask per value from 64-bit lane
// masks (scalar loads from the plan); e2..e5 are DPP stages, and e9, e10 are
// exchanged with the finished register bits e8, e7 by v_permlane16_swap /
// v_permlane32_swap before their stages.  The 8 wavefronts of a workgroup
// store blocks 8i .. 8i+7 of a column to LDS (round i), and each thread adds
// the 8-term signed sums for its sampled rows into its accumulators.  A
// column takes 2^(L-14) rounds; each wavefront's next block (or the first
// block of its next column) is loaded during the current one.  The scale,
// alpha and beta are applied in Float64 when the rows are written.
// Per 2048 entries a wavefront issues about 416 VALU instructions for the
// transform and one ds_write_b128 per 4 values, against ~700 in
// srht_wave_kernel's bit-exact order.

constexpr int kPrRegion = 2048 + 256;  // LDS words per block: 4 pad words per 32 entries

__device__ __forceinline__ int pr_pad(int e) { return e + ((e >> 5) << 2); }

// exchange lane bit 4 (ROW = 16) or 5 (ROW = 32) with register bit RB; the
// swap rewrites both registers in place
template <int RB, int ROW>
__device__ __forceinline__ void pr_swap(float (&x)[32]) {
#pragma unroll
  for (int v = 0; v < 32; ++v) {
    if (!(v & (1 << RB))) {
      // the builtins let the compiler insert the wait states these
      // instructions need after a VALU write of their operands
      const unsigned a = __float_as_uint(x[v]), b = __float_as_uint(x[v | (1 << RB)]);
      if constexpr (ROW == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        x[v] = __uint_as_float(r[0]);
        x[v | (1 << RB)] = __uint_as_float(r[1]);
      } else {
        const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        x[v] = __uint_as_float(r[0]);
        x[v | (1 << RB)] = __uint_as_float(r[1]);
      }
    }
  }
}

typedef __attribute__((address_space(4))) const uint64_t* pr_mask_ptr;

// mask[h * 32 + v] bit l: the entry held by register v of lane l at load
// time in block h has a negative sign (its sign bit is 0: srht.jl's
// `signs[i] ? scaling : -scaling`)
__global__ void prune_masks_kernel(const uint32_t* __restrict__ signs32, int nb,
                                   uint64_t* __restrict__ masks) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nb * 32) return;
  const int h = idx >> 5, v = idx & 31;
  uint64_t mk = 0;
  for (int l = 0; l < 64; ++l) {
    const int e = (h << 11) | (((l >> 5) & 1) << 10) | (((l >> 4) & 1) << 9) |
                  (wl_nibble(l & 15) << 2) | ((v >> 2) << 6) | (v & 3);
    if (!((signs32[e >> 5] >> (e & 31)) & 1u)) mk |= 1ull << l;
  }
  masks[idx] = mk;
}

template <int MS, int MODE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4)))
void srht_prune_kernel(const float* __restrict__ X, int64_t x_cs, int64_t m, int64_t ncols,
                       int nr, const uint64_t* __restrict__ masks_g,
                       const int32_t* __restrict__ op, const int32_t* __restrict__ ord, int s,
                       double scale, float* __restrict__ Y, int64_t y_rs, int64_t y_cs,
                       double alpha, double beta) {
  // the masks are read through the scalar cache (they are wave-uniform)
  const pr_mask_ptr masks = (pr_mask_ptr)masks_g;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* lds = reinterpret_cast<float*>(smem_raw);
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  // this thread's sampled rows ord[t + 512 j] (ord lists the rows by rho_l,
  // so the lanes of a wavefront read nearly consecutive LDS words; NULL: the
  // rows in order): LDS index of rho_l, and rho_h with the lane sign
  // (-1)^popcount(e2..e5) of the DPP stages in bit 30
  int sidx[MS], shi[MS];
#pragma unroll
  for (int j = 0; j < MS; ++j) {
    const int r = t + 512 * j;
    sidx[j] = -1;
    shi[j] = 0;
    if (r < s) {
      const int rho = op[ord ? ord[r] : r];
      const int lo = rho & 2047;
      sidx[j] = pr_pad(lo);
      shi[j] = (rho >> 11) | ((__builtin_popcount((lo >> 2) & 15) & 1) << 30);
    }
  }
  const int nib = wl_nibble(lane & 15);
  const float c0 = (nib & 1) ? -1.0f : 1.0f, c1 = (nib & 2) ? -1.0f : 1.0f;
  const float c2 = (nib & 4) ? -1.0f : 1.0f, c3 = (nib & 8) ? -1.0f : 1.0f;
  const int off = (((lane >> 5) & 1) << 10) | (((lane >> 4) & 1) << 9) | (nib << 2);
  // after the transform: registers hold e0, e1, e6, e10, e9 (bits 0..4), lanes
  // hold e2..e5, e8 (lane bit 4) and e7 (lane bit 5)
  float* wbase = lds + w * kPrRegion +
                 pr_pad((nib << 2) | (((lane >> 5) & 1) << 7) | (((lane >> 4) & 1) << 8));
  float acc[MS];
#pragma unroll
  for (int j = 0; j < MS; ++j) acc[j] = 0.0f;

  const int64_t my_cols = blockIdx.x < ncols ? (ncols - 1 - blockIdx.x) / gridDim.x + 1 : 0;
  const int64_t nitems = my_cols * nr;
  float xa[32], xb[32];
  // item it = (column it / nr, round it % nr), tracked by counters
  auto load_item = [&](float(&x)[32], int64_t col, int i) {
    const int h = 8 * i + w;
    // opaque per item: keeps the guarded modes' clamps inside the loop
    int offo = off;
    asm volatile("" : "+v"(offo));
    wl_load<MODE>(x, X + col * x_cs, int64_t(h) << 11, m, offo);
  };
  if (nitems > 0) load_item(xa, blockIdx.x, 0);
  int64_t col = blockIdx.x;
  int i = 0;

  auto item = [&](float(&x)[32], float(&nx)[32], int64_t it) {
    if (it + 1 < nitems) {
      if (i + 1 < nr) load_item(nx, col, i + 1);
      else load_item(nx, col + gridDim.x, 0);
    }
    const int h = 8 * i + w;
    pr_mask_ptr mk = masks + int64_t(h) * 32;
#pragma unroll
    for (int v = 0; v < 32; ++v) {
      float o;
      asm("v_cndmask_b32_e64 %0, %1, -%1, %2" : "=v"(o) : "v"(x[v]), "s"(mk[v]));
      x[v] = o;
    }
    stage<0, 32>(x);
    stage<1, 32>(x);
    wl_dpp_stage<0>(x, c0);
    wl_dpp_stage<1>(x, c1);
    wl_dpp_stage<2>(x, c2);
    wl_dpp_stage<3>(x, c3);
    stage<2, 32>(x);        // e6
    stage<3, 32>(x);        // e7
    stage<4, 32>(x);        // e8
    pr_swap<4, 16>(x);      // lane bit 4 <-> register bit 4: e9 <-> e8
    stage<4, 32>(x);        // e9
    pr_swap<3, 32>(x);      // lane bit 5 <-> register bit 3: e10 <-> e7
    stage<3, 32>(x);        // e10
    __syncthreads();        // the previous round's readers are done
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      const int eo = ((g & 1) << 6) | (((g >> 1) & 1) << 10) | (((g >> 2) & 1) << 9);
      typedef float v4f __attribute__((ext_vector_type(4)));
      *reinterpret_cast<v4f*>(wbase + pr_pad(eo)) =
          v4f{x[4 * g], x[4 * g + 1], x[4 * g + 2], x[4 * g + 3]};
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < MS; ++j) {
      if (sidx[j] < 0) continue;
      const float* p = lds + sidx[j];
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = p[q * kPrRegion];
      const int hi = shi[j];
      const float s0 = (hi & 1) ? -1.0f : 1.0f, s1 = (hi & 2) ? -1.0f : 1.0f;
      const float s2 = (hi & 4) ? -1.0f : 1.0f;
      const float a0 = fmaf(s0, v[1], v[0]), a1 = fmaf(s0, v[3], v[2]);
      const float a2 = fmaf(s0, v[5], v[4]), a3 = fmaf(s0, v[7], v[6]);
      const float b0 = fmaf(s1, a1, a0), b1 = fmaf(s1, a3, a2);
      const float c = fmaf(s2, b1, b0);
      const int neg = (__builtin_popcount((hi >> 3) & i) ^ (hi >> 30)) & 1;
      acc[j] += neg ? -c : c;
    }
    if (i == nr - 1) {
#pragma unroll
      for (int j = 0; j < MS; ++j) {
        if (sidx[j] < 0) continue;
        const int r = ord ? ord[t + 512 * j] : t + 512 * j;
        float* yp = Y + int64_t(r) * y_rs + col * y_cs;
        const double val = double(acc[j]) * scale;
        *yp = static_cast<float>(beta == 0.0 ? alpha * val : beta * double(*yp) + alpha * val);
        acc[j] = 0.0f;
      }
      i = 0;
      col += gridDim.x;
    } else {
      ++i;
    }
  };

  for (int64_t it = 0; it < nitems; it += 2) {
    item(xa, xb, it);
    if (it + 1 < nitems) item(xb, xa, it + 1);
  }
}

// Full transform of 2^14-entry chunks (stages 0..13, all outputs) for the
// generic FWHT: item i is chunk i % nch of column i / nch.  The same
// wave-local stages and exchange as srht_wave_kernel; after stage 13 each
// thread holds entries t + 512 v of the chunk and stores them, lane sign
// removed, to P in natural order (a wavefront stores 64 consecutive floats
// per instruction).  The stages across chunks (14, 15, ...) are later passes.
// src == nullptr transforms P in place: every chunk is read whole before it
// is written, by its own workgroup.
template <int MODE>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4)))
void fwht_chunk_wave_kernel(const float* __restrict__ src, int64_t src_cs, int64_t m_valid,
                            float* P, int64_t ldp, int64_t ncols, int64_t nch,
                            const uint32_t* __restrict__ signs32, double scale) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* lds = reinterpret_cast<float*>(smem_raw);
  const int t = threadIdx.x;
  const int nib = wl_nibble(t & 15);
  const float c0 = (nib & 1) ? -1.0f : 1.0f, c1 = (nib & 2) ? -1.0f : 1.0f;
  const float c2 = (nib & 4) ? -1.0f : 1.0f, c3 = (nib & 8) ? -1.0f : 1.0f;
  const int off = wl_load_off(t);
  const bool neg_b = __builtin_popcount((t >> 2) & 15) & 1;  // sign carried by entries t + 512 v
  float* lb_a = lds + wl_pad(off);
  float* lb_b = lds + wl_pad(t);
  const int64_t items = ncols * nch;
  auto col_of = [&](int64_t it) { return it / nch; };
  auto base_of = [&](int64_t it) { return (it % nch) << 14; };
  auto xc_of = [&](int64_t it) -> const float* {
    return src ? src + col_of(it) * src_cs : P + col_of(it) * ldp;
  };
  float xa[32], xb[32];
  if (blockIdx.x < items) wl_load<MODE>(xa, xc_of(blockIdx.x), base_of(blockIdx.x), m_valid, off);

  auto body = [&](float(&x)[32], float(&nx)[32], int64_t it) {
    const int64_t nit = it + gridDim.x;
    int offo = off;
    asm volatile("" : "+v"(offo));
    if (nit < items) wl_load<MODE>(nx, xc_of(nit), base_of(nit), m_valid, offo);
    const int64_t cb = base_of(it);
    uint32_t sw = 0xffffffffu;  // no signs: every entry scaled by +scale
    if (signs32) {
      sw = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        sw |= ((signs32[(cb >> 5) + (off >> 5) + 2 * k] >> (off & 31)) & 15u) << (4 * k);
    }
    asm volatile("" : "+v"(sw));
    stage0_signed<32>(x, sw, scale);
    stage<1, 32>(x);
    wl_dpp_stage<0>(x, c0);
    wl_dpp_stage<1>(x, c1);
    wl_dpp_stage<2>(x, c2);
    wl_dpp_stage<3>(x, c3);
    stage<2, 32>(x);
    stage<3, 32>(x);
    stage<4, 32>(x);
    __syncthreads();  // the previous item's readers of LDS are done
#pragma unroll
    for (int v = 0; v < 32; ++v) lb_a[(v & 3) + 66 * (v >> 2)] = x[v];
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 32; ++v) x[v] = lb_b[530 * v];
    stage_range<0, 5, 32>(x);
    float* pc = P + col_of(it) * ldp + cb + t;
#pragma unroll
    for (int v = 0; v < 32; ++v) pc[v << 9] = neg_b ? -x[v] : x[v];
  };
  for (int64_t it = blockIdx.x; it < items; it += 2 * int64_t(gridDim.x)) {
    body(xa, xb, it);
    if (it + gridDim.x < items) body(xb, xa, it + gridDim.x);
  }
}

static int num_cus(int device);

template <int MODE>
static int launch_fwht_chunks(rla_ctx* ctx, const float* src, int64_t src_cs, int64_t m_valid,
                              float* P, int64_t ldp, int64_t ncols, int64_t nch,
                              const uint64_t* signs, double scale) {
  const size_t lds = size_t(kWaveLdsWords) * sizeof(float);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(&fwht_chunk_wave_kernel<MODE>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  RLA_HIP(attr_err);
  const int64_t items = ncols * nch;
  const int64_t grid = std::min<int64_t>(items, int64_t(num_cus(ctx->device)) * 2);
  fwht_chunk_wave_kernel<MODE><<<dim3(static_cast<unsigned>(grid)), 512, lds, ctx->stream>>>(
      src, src_cs, m_valid, P, ldp, ncols, nch, reinterpret_cast<const uint32_t*>(signs), scale);
  RLA_LAUNCH_CHECK("fwht_chunk_wave");
  return RLA_OK;
}

// Stages 0..13 of the FWHT of ncols columns of length n_pad = 2^L (L >= 14),
// Float32: read from src (column stride src_cs, unit row stride, the first
// m_valid rows, zeros past them) or in place from P; written to P (ldp).
int fwht_chunks_f32(rla_ctx* ctx, const float* src, int64_t src_cs, int64_t m_valid, float* P,
                    int64_t ldp, int64_t ncols, int L, const uint64_t* signs, double scale) {
  const int64_t nch = int64_t(1) << (L - 14);
  const float* base = src ? src : P;
  const int64_t cs = src ? src_cs : ldp;
  const bool aligned = (reinterpret_cast<uintptr_t>(base) % 16 == 0) && (cs % 4 == 0);
  if (aligned && m_valid == (int64_t(1) << L))
    return launch_fwht_chunks<0>(ctx, src, src_cs, m_valid, P, ldp, ncols, nch, signs, scale);
  if (aligned && m_valid % 4 == 0)
    return launch_fwht_chunks<1>(ctx, src, src_cs, m_valid, P, ldp, ncols, nch, signs, scale);
  return launch_fwht_chunks<2>(ctx, src, src_cs, m_valid, P, ldp, ncols, nch, signs, scale);
}

// ---- shape selection ---------------------------------------------------------

struct FusedShape {
  int LC, NCH, VB;
};

// Chunks of up to 64 KiB (2^14 Float32 / 2^13 Float64 entries, plus padding)
// so that two workgroups share a CU; columns of up to 8 chunks.
static bool fused_shape(int elem, int64_t n_pad, FusedShape& f) {
  const int L = ilog2(n_pad);
  const int lc_max = elem == 4 ? 14 : 13;
  if (L < 11 || L > lc_max + 3) return false;
  f.LC = std::min(L, lc_max);
  f.NCH = 1 << (L - f.LC);
  f.VB = 5;
  return true;
}


// ---- plan ------------------------------------------------------------------
// layout (32-bit words): [0] magic, [1] LC, [2] NCH, [3] s, [4] n_pad lo,
// [5] n_pad hi, [6] elem bytes, [7] nslots; [8 .. 8+W) needed-position
// bitmap, [8+W .. 8+2W) its exclusive per-word prefix counts (W = 2^LC/32);
// then slot_pos[nslots_cap] (position in the chunk of each slot, ascending)
// and rowmap[nslots_cap * NCH] (output row of position c*2^LC + slot_pos, or
// -1).  nslots_cap = min(s, 2^LC).

constexpr uint32_t kPlanMagic = 0x52484C50u;

static size_t plan_words(const FusedShape& f, int64_t s) {
  const size_t w = size_t(1) << (f.LC - 5);
  const size_t cap = static_cast<size_t>(std::min<int64_t>(s, int64_t(1) << f.LC));
  return kPlanHdr + 2 * w + cap * (1 + f.NCH);
}

// The pruned Float32 kernel: n_pad = 2^L with 14 <= L <= 24, at most 8
// sampled rows per thread (s <= 4096).  Its plan part is the sign-mask table
// (n_pad / 2048 blocks x 32 registers x 64-bit lane masks), after the fused
// kernel's part, 256-byte aligned.
static bool prune_ok(int elem, int64_t n_pad, int64_t s) {
  const int L = ilog2(n_pad);
  return elem == 4 && is_pow2(n_pad) && L >= 14 && L <= 24 && s >= 1 && s <= 4096;
}
static size_t prune_mask_words(int64_t n_pad) { return size_t(n_pad / 2048) * 32 * 2; }
static size_t prune_mask_offset_words(int elem, int64_t n_pad, int64_t s) {
  FusedShape f;
  const size_t fused = fused_shape(elem, n_pad, f) ? plan_words(f, s) : 0;
  return (fused + 63) & ~size_t(63);
}
// after the masks: the row order (s words), the rows sorted by op & 2047
static size_t prune_order_offset_words(int elem, int64_t n_pad, int64_t s) {
  return prune_mask_offset_words(elem, n_pad, s) + prune_mask_words(n_pad);
}
static size_t total_plan_words(int elem, int64_t n_pad, int64_t s) {
  FusedShape f;
  const size_t fused = fused_shape(elem, n_pad, f) ? plan_words(f, s) : 0;
  if (!prune_ok(elem, n_pad, s)) return fused;
  return prune_order_offset_words(elem, n_pad, s) + size_t(s);
}
// ord = the rows 0..s-1 sorted by (op[r] & 2047, r): a bitonic sort of the
// keys (op[r] & 2047) << 12 | r in LDS by one workgroup (s <= 4096)
__global__ __launch_bounds__(1024) void prune_order_kernel(const int32_t* __restrict__ op,
                                                           int s, int32_t* __restrict__ ord) {
  __shared__ uint32_t key[4096];
  int n = 1;
  while (n < s) n <<= 1;
  for (int i = threadIdx.x; i < n; i += 1024)
    key[i] = i < s ? (uint32_t(op[i] & 2047) << 12) | uint32_t(i) : 0xFFFFFFFFu;
  __syncthreads();
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += 1024) {
        const int p = i ^ j;
        if (p > i) {
          const uint32_t a = key[i], b = key[p];
          if (((i & k) == 0) == (a > b)) {
            key[i] = b;
            key[p] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < s; i += 1024) ord[i] = static_cast<int32_t>(key[i] & 4095u);
}

static int build_prune_masks(rla_ctx* ctx, int64_t n_pad, const uint64_t* signs,
                             uint64_t* masks) {
  const int nb = static_cast<int>(n_pad / 2048);
  prune_masks_kernel<<<dim3(unsigned(ceil_div(int64_t(nb) * 32, 256))), 256, 0, ctx->stream>>>(
      reinterpret_cast<const uint32_t*>(signs), nb, masks);
  RLA_LAUNCH_CHECK("srht_prune_masks");
  return RLA_OK;
}

__global__ void plan_init_kernel(uint32_t* plan, int LC, int NCH, int64_t s, int64_t n_pad,
                                 int elem, int64_t cap) {
  const int64_t w = int64_t(1) << (LC - 5);
  const int64_t total = 2 * w + cap * (1 + NCH);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    // slot_pos and rowmap start at -1
    plan[kPlanHdr + i] = i < 2 * w ? 0u : 0xFFFFFFFFu;
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

__global__ void plan_mark_kernel(uint32_t* bitmap, const int32_t* __restrict__ op, int64_t s,
                                 int LC) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r >= s) return;
  const int e = op[r] & ((1 << LC) - 1);
  atomicOr(&bitmap[e >> 5], 1u << (e & 31));
}

__global__ __launch_bounds__(1024) void plan_prefix_kernel(const uint32_t* __restrict__ bitmap,
                                                           uint32_t* prefix, int words,
                                                           uint32_t* nslots_out) {
  __shared__ uint32_t part[1024];
  const int per = (words + 1023) / 1024;
  const int b = threadIdx.x * per;
  uint32_t s = 0;
  for (int i = b; i < b + per && i < words; ++i) s += __popc(bitmap[i]);
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
  for (int i = b; i < b + per && i < words; ++i) {
    prefix[i] = run;
    run += __popc(bitmap[i]);
  }
  if (threadIdx.x == 1023) *nslots_out = part[1023];
}

__global__ void plan_rowmap_kernel(const uint32_t* __restrict__ bitmap,
                                   const uint32_t* __restrict__ prefix, int32_t* slot_pos,
                                   int32_t* rowmap, const int32_t* __restrict__ op, int64_t s,
                                   int LC, int NCH) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r >= s) return;
  const int64_t pos = op[r];
  const int e = static_cast<int>(pos & ((1 << LC) - 1));
  const int slot =
      static_cast<int>(prefix[e >> 5]) + __popc(bitmap[e >> 5] & ((1u << (e & 31)) - 1u));
  slot_pos[slot] = e;
  rowmap[int64_t(slot) * NCH + (pos >> LC)] = static_cast<int32_t>(r);
}

static int build_plan(rla_ctx* ctx, int elem, int64_t n_pad, int64_t s, const int32_t* op,
                      const uint64_t* signs, uint32_t* plan) {
  if (signs && prune_ok(elem, n_pad, s)) {
    int rc = build_prune_masks(
        ctx, n_pad, signs,
        reinterpret_cast<uint64_t*>(plan + prune_mask_offset_words(elem, n_pad, s)));
    if (rc) return rc;
    prune_order_kernel<<<1, 1024, 0, ctx->stream>>>(
        op, static_cast<int>(s),
        reinterpret_cast<int32_t*>(plan + prune_order_offset_words(elem, n_pad, s)));
    RLA_LAUNCH_CHECK("srht_prune_order");
  }
  FusedShape f;
  if (!fused_shape(elem, n_pad, f)) return RLA_OK;
  const int w = 1 << (f.LC - 5);
  const int64_t cap = std::min<int64_t>(s, int64_t(1) << f.LC);
  hipStream_t st = ctx->stream;
  plan_init_kernel<<<64, 256, 0, st>>>(plan, f.LC, f.NCH, s, n_pad, elem, cap);
  uint32_t* bitmap = plan + kPlanHdr;
  uint32_t* prefix = bitmap + w;
  int32_t* slot_pos = reinterpret_cast<int32_t*>(prefix + w);
  int32_t* rowmap = slot_pos + cap;
  plan_mark_kernel<<<dim3(unsigned(ceil_div(s, 256))), 256, 0, st>>>(bitmap, op, s, f.LC);
  plan_prefix_kernel<<<1, 1024, 0, st>>>(bitmap, prefix, w, plan + 7);
  plan_rowmap_kernel<<<dim3(unsigned(ceil_div(s, 256))), 256, 0, st>>>(bitmap, prefix, slot_pos,
                                                                      rowmap, op, s, f.LC, f.NCH);
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

static int64_t nslots_cap(const FusedShape& f, int64_t s) {
  return std::min<int64_t>(s, int64_t(1) << f.LC);
}

// slot values of the first NCH-1 chunks live in LDS when that still lets
// two workgroups share a CU, otherwise in a per-workgroup global scratch
static bool scratch_in_lds(const FusedShape& f, int elem, int64_t s) {
  const size_t chunk = ((size_t(1) << f.LC) + (size_t(1) << (f.LC - 5))) * elem;
  return f.NCH > 1 && 2 * (chunk + (f.NCH - 1) * nslots_cap(f, s) * elem) <= 160 * 1024;
}

static size_t fused_lds_bytes(const FusedShape& f, int elem, int64_t s) {
  const size_t chunk = ((size_t(1) << f.LC) + (size_t(1) << (f.LC - 5))) * elem;
  return chunk + (scratch_in_lds(f, elem, s) ? (f.NCH - 1) * nslots_cap(f, s) * elem : 0);
}

template <class T, int LC, int NCH, int VB, bool PF, bool SCR_LDS>
static int launch_fused(rla_ctx* ctx, const T* X, int64_t x_cs, int64_t m, int64_t ncols,
                        const uint64_t* signs, double scale, const uint32_t* plan, int64_t s,
                        T* Y, int64_t y_rs, int64_t y_cs, double alpha, double beta,
                        T* gscr) {
  const int nt = 1 << (LC - VB);
  const int64_t cap = std::min<int64_t>(s, int64_t(1) << LC);
  const FusedShape f{LC, NCH, VB};
  const size_t lds = fused_lds_bytes(f, sizeof(T), s);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&srht_fused_kernel<T, LC, NCH, VB, PF, SCR_LDS>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  RLA_HIP(attr_err);
  // workgroups resident per CU, limited by LDS (160 KiB) and by 16 waves / CU
  const int per_cu = std::max<int>(1, static_cast<int>((160 * 1024) / lds));
  const int64_t grid = std::min<int64_t>(ncols, int64_t(num_cus(ctx->device)) * per_cu);
  const int vec_ok = (x_cs % (sizeof(T) == 4 ? 4 : 2) == 0) &&
                     (reinterpret_cast<uintptr_t>(X) % 16 == 0);
  srht_fused_kernel<T, LC, NCH, VB, PF, SCR_LDS><<<dim3(static_cast<unsigned>(grid)), nt, lds,
                                                   ctx->stream>>>(
      X, x_cs, m, ncols, reinterpret_cast<const uint32_t*>(signs), scale, plan, Y, y_rs, y_cs,
      alpha, beta, vec_ok, gscr, static_cast<int>(cap));
  RLA_LAUNCH_CHECK("srht_fused");
  return RLA_OK;
}

template <class T, int LC, int NCH, int VB, int MS>
static int launch_fused_reg(rla_ctx* ctx, const T* X, int64_t x_cs, int64_t m, int64_t ncols,
                            const uint64_t* signs, double scale, const uint32_t* plan, int64_t s,
                            T* Y, int64_t y_rs, int64_t y_cs, double alpha, double beta) {
  const int nt = 1 << (LC - VB);
  const int64_t cap = std::min<int64_t>(s, int64_t(1) << LC);
  const size_t lds = ((size_t(1) << LC) + (size_t(1) << (LC - 5))) * sizeof(T);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&srht_fused_reg_kernel<T, LC, NCH, VB, MS>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  RLA_HIP(attr_err);
  const int per_cu = std::max<int>(1, static_cast<int>((160 * 1024) / lds));
  const int64_t grid = std::min<int64_t>(ncols, int64_t(num_cus(ctx->device)) * per_cu);
  const int vec_ok = (x_cs % (sizeof(T) == 4 ? 4 : 2) == 0) &&
                     (reinterpret_cast<uintptr_t>(X) % 16 == 0);
  srht_fused_reg_kernel<T, LC, NCH, VB, MS><<<dim3(static_cast<unsigned>(grid)), nt, lds,
                                              ctx->stream>>>(
      X, x_cs, m, ncols, reinterpret_cast<const uint32_t*>(signs), scale, plan, Y, y_rs, y_cs,
      alpha, beta, vec_ok, static_cast<int>(cap));
  RLA_LAUNCH_CHECK("srht_fused_reg");
  return RLA_OK;
}

template <int NCH, int MODE>
static int launch_wave(rla_ctx* ctx, const float* X, int64_t x_cs, int64_t m, int64_t ncols,
                       const uint64_t* signs, double scale, const uint32_t* plan, int64_t s,
                       float* Y, int64_t y_rs, int64_t y_cs, double alpha, double beta) {
  const int64_t cap = std::min<int64_t>(s, int64_t(1) << 14);
  const size_t lds = size_t(kWaveLdsWords) * sizeof(float);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(&srht_wave_kernel<NCH, 4, MODE>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  RLA_HIP(attr_err);
  const int per_cu = std::max<int>(1, static_cast<int>((160 * 1024) / lds));
  const int64_t grid = std::min<int64_t>(ncols, int64_t(num_cus(ctx->device)) * per_cu);
  srht_wave_kernel<NCH, 4, MODE><<<dim3(static_cast<unsigned>(grid)), 512, lds, ctx->stream>>>(
      X, x_cs, m, ncols, reinterpret_cast<const uint32_t*>(signs), scale, plan, Y, y_rs, y_cs,
      alpha, beta, static_cast<int>(cap));
  RLA_LAUNCH_CHECK("srht_wave");
  return RLA_OK;
}

// n_pad = NCH * 2^14 rows: picks the load mode from the column alignment and m
template <int NCH>
static int launch_wave_any(rla_ctx* ctx, const float* X, int64_t x_cs, int64_t m, int64_t ncols,
                           const uint64_t* signs, double scale, const uint32_t* plan, int64_t s,
                           float* Y, int64_t y_rs, int64_t y_cs, double alpha, double beta) {
  const bool aligned = (x_cs % 4 == 0) && (reinterpret_cast<uintptr_t>(X) % 16 == 0);
  if (aligned && m == (int64_t(NCH) << 14))
    return launch_wave<NCH, 0>(ctx, X, x_cs, m, ncols, signs, scale, plan, s, Y, y_rs, y_cs,
                               alpha, beta);
  if (aligned && m % 4 == 0)
    return launch_wave<NCH, 1>(ctx, X, x_cs, m, ncols, signs, scale, plan, s, Y, y_rs, y_cs,
                               alpha, beta);
  return launch_wave<NCH, 2>(ctx, X, x_cs, m, ncols, signs, scale, plan, s, Y, y_rs, y_cs, alpha,
                             beta);
}

template <int MS, int MODE>
static int launch_prune_ms(rla_ctx* ctx, const float* X, int64_t x_cs, int64_t m, int64_t n_pad,
                           int64_t ncols, const uint64_t* masks, const int32_t* op,
                           const int32_t* ord, int64_t s, double scale, float* Y, int64_t y_rs, int64_t y_cs, double alpha,
                           double beta) {
  const size_t lds = size_t(8) * kPrRegion * sizeof(float);
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [&] {
    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(&srht_prune_kernel<MS, MODE>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  RLA_HIP(attr_err);
  const int per_cu = std::max<int>(1, static_cast<int>((160 * 1024) / lds));
  const int64_t grid = std::min<int64_t>(ncols, int64_t(num_cus(ctx->device)) * per_cu);
  const int nr = static_cast<int>(n_pad >> 14);
  srht_prune_kernel<MS, MODE><<<dim3(static_cast<unsigned>(grid)), 512, lds, ctx->stream>>>(
      X, x_cs, m, ncols, nr, masks, op, ord, static_cast<int>(s),
      scale, Y, y_rs, y_cs, alpha, beta);
  RLA_LAUNCH_CHECK("srht_prune");
  return RLA_OK;
}

template <int MODE>
static int launch_prune_mode(rla_ctx* ctx, const float* X, int64_t x_cs, int64_t m,
                             int64_t n_pad, int64_t ncols, const uint64_t* masks,
                             const int32_t* op, const int32_t* ord, int64_t s, double scale, float* Y, int64_t y_rs,
                             int64_t y_cs, double alpha, double beta) {
  if (s <= 512)
    return launch_prune_ms<1, MODE>(ctx, X, x_cs, m, n_pad, ncols, masks, op, ord, s, scale, Y, y_rs,
                                    y_cs, alpha, beta);
  if (s <= 1024)
    return launch_prune_ms<2, MODE>(ctx, X, x_cs, m, n_pad, ncols, masks, op, ord, s, scale, Y, y_rs,
                                    y_cs, alpha, beta);
  if (s <= 2048)
    return launch_prune_ms<4, MODE>(ctx, X, x_cs, m, n_pad, ncols, masks, op, ord, s, scale, Y, y_rs,
                                    y_cs, alpha, beta);
  return launch_prune_ms<8, MODE>(ctx, X, x_cs, m, n_pad, ncols, masks, op, ord, s, scale, Y, y_rs,
                                  y_cs, alpha, beta);
}

static int launch_prune(rla_ctx* ctx, const float* X, int64_t x_cs, int64_t m, int64_t n_pad,
                        int64_t ncols, const uint64_t* masks, const int32_t* op,
                        const int32_t* ord, int64_t s,
                        double scale, float* Y, int64_t y_rs, int64_t y_cs, double alpha,
                        double beta) {
  const bool aligned = (x_cs % 4 == 0) && (reinterpret_cast<uintptr_t>(X) % 16 == 0);
  if (aligned && m == n_pad)
    return launch_prune_mode<0>(ctx, X, x_cs, m, n_pad, ncols, masks, op, ord, s, scale, Y, y_rs,
                                y_cs, alpha, beta);
  if (aligned && m % 4 == 0)
    return launch_prune_mode<1>(ctx, X, x_cs, m, n_pad, ncols, masks, op, ord, s, scale, Y, y_rs,
                                y_cs, alpha, beta);
  return launch_prune_mode<2>(ctx, X, x_cs, m, n_pad, ncols, masks, op, ord, s, scale, Y, y_rs, y_cs,
                              alpha, beta);
}

template <class T>
int srht_fused_forward(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                       const int32_t* op, const uint64_t* signs, double scale, const void* plan,
                       const T* X, int64_t x_rs, int64_t x_cs, T* Y, int64_t y_rs, int64_t y_cs,
                       double alpha, double beta, bool exact, bool* handled) {
  *handled = false;
  if (x_rs != 1) return RLA_OK;
  if constexpr (sizeof(T) == 4) {
    if (!exact && prune_ok(4, n_pad, s)) {
      const uint64_t* masks = nullptr;
      const int32_t* ord = nullptr;  // no plan: the rows in order
      if (plan) {
        masks = reinterpret_cast<const uint64_t*>(static_cast<const uint32_t*>(plan) +
                                                  prune_mask_offset_words(4, n_pad, s));
        ord = reinterpret_cast<const int32_t*>(static_cast<const uint32_t*>(plan) +
                                               prune_order_offset_words(4, n_pad, s));
      } else {
        void* ws = nullptr;
        int rc = workspace(ctx, prune_mask_words(n_pad) * 4, &ws);
        if (rc) return rc;
        rc = build_prune_masks(ctx, n_pad, signs, static_cast<uint64_t*>(ws));
        if (rc) return rc;
        masks = static_cast<const uint64_t*>(ws);
      }
      *handled = true;
      return launch_prune(ctx, X, x_cs, m, n_pad, ncols, masks, op, ord, s, scale, Y, y_rs, y_cs,
                          alpha, beta);
    }
  }
  FusedShape f;
  if (!fused_shape(sizeof(T), n_pad, f)) return RLA_OK;
  if (fused_lds_bytes(f, sizeof(T), s) > 160 * 1024) return RLA_OK;
  if (s > (int64_t(1) << 30)) return RLA_OK;
  const bool lds_scr = scratch_in_lds(f, sizeof(T), s);
  const int per_cu = std::max<int>(1, static_cast<int>((160 * 1024) / fused_lds_bytes(f, sizeof(T), s)));
  const int64_t grid = std::min<int64_t>(ncols, int64_t(num_cus(ctx->device)) * per_cu);
  const size_t scr_bytes =
      (f.NCH > 1 && !lds_scr) ? size_t(grid) * (f.NCH - 1) * nslots_cap(f, s) * sizeof(T) : 0;
  const size_t plan_bytes = plan ? 0 : plan_words(f, s) * 4;
  void* ws = nullptr;
  if (scr_bytes + plan_bytes > 0) {
    int rc = workspace(ctx, ((plan_bytes + 255) & ~size_t(255)) + scr_bytes, &ws);
    if (rc) return rc;
  }
  const uint32_t* pl = static_cast<const uint32_t*>(plan);
  if (!pl) {
    int rc = build_plan(ctx, sizeof(T), n_pad, s, op, nullptr, static_cast<uint32_t*>(ws));
    if (rc) return rc;
    pl = static_cast<const uint32_t*>(ws);
  }
  T* gscr = scr_bytes ? reinterpret_cast<T*>(static_cast<char*>(ws) +
                                             ((plan_bytes + 255) & ~size_t(255)))
                      : nullptr;
  *handled = true;
  // Float32 with 2 or 4 chunks of 2^14: the wave-local transform, slot values
  // parked in registers, up to 4 slots per thread
  if constexpr (sizeof(T) == 4) {
    if (f.LC == 14 && (f.NCH == 2 || f.NCH == 4) && nslots_cap(f, s) <= 4 * 512)
      return f.NCH == 2 ? launch_wave_any<2>(ctx, X, x_cs, m, ncols, signs, scale, pl, s, Y,
                                             y_rs, y_cs, alpha, beta)
                        : launch_wave_any<4>(ctx, X, x_cs, m, ncols, signs, scale, pl, s, Y,
                                             y_rs, y_cs, alpha, beta);
  }
  if (sizeof(T) == 4 && f.NCH > 1 && nslots_cap(f, s) <= 4 * (int64_t(1) << (f.LC - 5))) {
#define RLA_FUSED_REG(LCV, NCHV)                                                               \
    if (f.LC == LCV && f.NCH == NCHV)                                                          \
      return launch_fused_reg<T, LCV, NCHV, 5, 4>(ctx, X, x_cs, m, ncols, signs, scale, pl, s, \
                                                  Y, y_rs, y_cs, alpha, beta);
    if constexpr (sizeof(T) == 4) {
      RLA_FUSED_REG(14, 2) RLA_FUSED_REG(14, 4)
    }
#undef RLA_FUSED_REG
  }
#define RLA_FUSED(LCV, NCHV, SCRV, PFV)                                                       \
  if (f.LC == LCV && f.NCH == NCHV && lds_scr == SCRV)                                          \
    return launch_fused<T, LCV, NCHV, 5, PFV, SCRV>(ctx, X, x_cs, m, ncols, signs, scale, pl, s,   \
                                                    Y, y_rs, y_cs, alpha, beta, gscr);
  if constexpr (sizeof(T) == 4) {
    RLA_FUSED(11, 1, false, false) RLA_FUSED(12, 1, false, false)
    RLA_FUSED(13, 1, false, false) RLA_FUSED(14, 1, false, false)
    RLA_FUSED(14, 2, true, true) RLA_FUSED(14, 2, false, true) RLA_FUSED(14, 4, false, true)
    RLA_FUSED(14, 8, false, true)
  } else {
    RLA_FUSED(11, 1, false, false) RLA_FUSED(12, 1, false, false)
    RLA_FUSED(13, 1, false, false) RLA_FUSED(13, 2, true, false)
    RLA_FUSED(13, 2, false, false) RLA_FUSED(13, 4, false, false)
    RLA_FUSED(13, 8, false, false)
  }
#undef RLA_FUSED
  *handled = false;
  return RLA_OK;
}

template int srht_fused_forward<float>(rla_ctx*, int64_t, int64_t, int64_t, int64_t,
                                       const int32_t*, const uint64_t*, double, const void*,
                                       const float*, int64_t, int64_t, float*, int64_t, int64_t,
                                       double, double, bool, bool*);
template int srht_fused_forward<double>(rla_ctx*, int64_t, int64_t, int64_t, int64_t,
                                        const int32_t*, const uint64_t*, double, const void*,
                                        const double*, int64_t, int64_t, double*, int64_t,
                                        int64_t, double, double, bool, bool*);

}  // namespace rla

using namespace rla;

extern "C" {

int rla_srht_plan_bytes(int elem_bytes, int64_t n_pad, int64_t s, size_t* bytes) {
  if (!bytes) return fail(RLA_EARG, "rla_srht_plan_bytes: bytes is NULL");
  if (elem_bytes != 4 && elem_bytes != 8) return fail(RLA_EARG, "elem_bytes must be 4 or 8");
  *bytes = is_pow2(n_pad) ? total_plan_words(elem_bytes, n_pad, s) * 4 : 0;
  return RLA_OK;
}

int rla_srht_plan(rla_ctx* ctx, int elem_bytes, int64_t n_pad, int64_t s, const int32_t* op,
                  const uint64_t* sign_chunks, void* plan) {
  RLA_CHECK_CTX(ctx);
  if (elem_bytes != 4 && elem_bytes != 8) return fail(RLA_EARG, "elem_bytes must be 4 or 8");
  if (!is_pow2(n_pad) || s < 1 || s > n_pad) return fail(RLA_EARG, "bad SRHT plan shape");
  if (total_plan_words(elem_bytes, n_pad, s) == 0) return RLA_OK;  // nothing to plan
  if (!plan || !op || !sign_chunks) return fail(RLA_EARG, "rla_srht_plan: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  return build_plan(ctx, elem_bytes, n_pad, s, op, sign_chunks, static_cast<uint32_t*>(plan));
}

}  // extern "C"
