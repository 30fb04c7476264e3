// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the derived
// draws used by the recipe generators.  Everything here is __host__
// __device__ integer arithmetic, so a recipe generated on the GPU is
// bit-identical to the same recipe generated on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rla {

struct u32x4 {
  uint32_t v[4];
};

__host__ __device__ inline void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = static_cast<uint64_t>(a) * static_cast<uint64_t>(b);
  hi = static_cast<uint32_t>(p >> 32);
  lo = static_cast<uint32_t>(p);
}

// Ten rounds of the Philox S-box/P-box network with the published
// multipliers and Weyl key increments.
__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(M0, c.v[0], hi0, lo0);
    mulhilo32(M1, c.v[2], hi1, lo1);
    u32x4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Draw domains keep the index, sign and Gaussian streams of one recipe
// independent of each other.
enum : uint32_t {
  DOM_INDEX = 1u,
  DOM_SIGN = 2u,
  DOM_GAUSS = 3u,
  DOM_PERM = 4u,
};

// A sequential stream of 32-bit words for one (seed, epoch, domain, item):
// counter = {block, item_lo, item_hi ^ (domain << 24), epoch}, key = seed.
struct Stream {
  uint32_t k0, k1;
  u32x4 ctr;
  u32x4 buf;
  int pos;

  __host__ __device__ Stream(uint64_t seed, uint64_t epoch, uint32_t domain, uint64_t item) {
    k0 = static_cast<uint32_t>(seed);
    k1 = static_cast<uint32_t>(seed >> 32);
    ctr.v[0] = 0u;
    ctr.v[1] = static_cast<uint32_t>(item);
    ctr.v[2] = static_cast<uint32_t>(item >> 32) ^ (domain << 24);
    ctr.v[3] = static_cast<uint32_t>(epoch) ^ static_cast<uint32_t>(epoch >> 32);
    pos = 4;
  }

  __host__ __device__ uint32_t next() {
    if (pos == 4) {
      buf = philox4x32_10(ctr, k0, k1);
      ctr.v[0] += 1u;
      pos = 0;
    }
    return buf.v[pos++];
  }

  // Uniform integer in [0, n), n >= 1, by Lemire's multiply-and-reject
  // ("Fast random integer generation in an interval", TOMACS 2019).
  __host__ __device__ uint32_t below(uint32_t n) {
    uint32_t x = next();
    uint64_t m = static_cast<uint64_t>(x) * n;
    uint32_t l = static_cast<uint32_t>(m);
    if (l < n) {
      uint32_t t = (0u - n) % n;
      while (l < t) {
        x = next();
        m = static_cast<uint64_t>(x) * n;
        l = static_cast<uint32_t>(m);
      }
    }
    return static_cast<uint32_t>(m >> 32);
  }

  // Uniform double in (0, 1) with 53 random bits.
  __host__ __device__ double open01() {
    uint32_t a = next() >> 5;  // 27 bits
    uint32_t b = next() >> 6;  // 26 bits
    double u = (static_cast<double>(a) * 67108864.0 + static_cast<double>(b));
    return (u + 0.5) * (1.0 / 9007199254740992.0);
  }
};

// Sign bit j of group `item` (one 32-bit word per 32 signs).
__host__ __device__ inline uint32_t sign_word(uint64_t seed, uint64_t epoch, uint64_t word) {
  Stream st(seed, epoch, DOM_SIGN, word);
  return st.next();
}

// Floyd's algorithm for a uniform k-subset of [0, n), kept sorted ascending
// in out[0..k) (insertion into a sorted array).  Matches the contract of
// StatsBase sample!(1:n, x; replace=false, ordered=true).
template <class Idx>
__host__ __device__ inline void floyd_sorted(Stream& st, uint32_t n, int k, Idx* out) {
  int len = 0;
  for (uint32_t t = n - static_cast<uint32_t>(k); t < n; ++t) {
    uint32_t v = st.below(t + 1u);
    // membership test in the sorted prefix
    int lo = 0, hi = len;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (static_cast<uint32_t>(out[mid]) < v) lo = mid + 1; else hi = mid;
    }
    if (lo < len && static_cast<uint32_t>(out[lo]) == v) {
      v = t;  // t is larger than every element drawn so far: append
      lo = len;
    }
    for (int q = len; q > lo; --q) out[q] = out[q - 1];
    out[lo] = static_cast<Idx>(v);
    ++len;
  }
}

}  // namespace rla
