// Internal helpers shared by the librla_mi355x translation units: context,
// error reporting, workspace, launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/rla.h"

struct rla_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* ws = nullptr;       // grow-only device workspace
  size_t ws_bytes = 0;
  int* flag = nullptr;      // small device word for kernel-side checks
  int* flag_host = nullptr; // pinned mirror of flag
  void* blas = nullptr;     // rocblas_handle, created on first use
  hipEvent_t order_ev = nullptr;  // orders work across a change of stream
};

namespace rla {

// one matrix of a batched factorisation (chol.hip): G (k x k, ld k) -> Rt
// (row-major, ld k) and the inverses of its 16 x 16 diagonal blocks; pivots of
// columns >= check_k are not tested
// Address-space views for device code.  A pointer read from memory (a
// descriptor) or a volatile pointer stays generic to the compiler, and generic
// (flat) accesses are counted both as memory and as LDS traffic, so a wait
// for one drains the other queue too.  Accesses through these views compile
// to global_* / ds_* instructions.
template <class T>
using global_ptr = __attribute__((address_space(1))) T*;
template <class T>
using lds_ptr = __attribute__((address_space(3))) T*;

// Round-to-nearest-even float -> bf16 bits (NaN kept quiet).  The BF16
// operands of the mixed-precision products are made from FP64 values as
// f32_to_bf16_bits(float(x)) wherever they are written.
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return uint16_t((u >> 16) | 0x40);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

struct CholBatchItem {
  int k, check_k;
  const double* G;
  double* Rt;
  double* Dinv;
  int* info;
};
// items: device array; every k <= 272
int chol_blocks_batched(hipStream_t stream, int64_t nitems, const CholBatchItem* items);

// Thread-local last-error text behind rla_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// HIP call check: on error records the message and returns RLA_EHIP.
#define RLA_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess)                                                    \
      return ::rla::fail(RLA_EHIP, std::string(#call " failed: ") +          \
                                       hipGetErrorString(_e));               \
  } while (0)

// After a launch: picks up configuration errors without synchronising.
#define RLA_LAUNCH_CHECK(name)                                               \
  do {                                                                       \
    hipError_t _e = hipGetLastError();                                       \
    if (_e != hipSuccess)                                                    \
      return ::rla::fail(RLA_EHIP, std::string(name " launch failed: ") +    \
                                       hipGetErrorString(_e));               \
  } while (0)

#define RLA_CHECK_CTX(ctx)                                                   \
  do {                                                                       \
    if (!(ctx)) return ::rla::fail(RLA_EARG, "null rla_ctx");                \
  } while (0)

// Releases the context's rocBLAS handle (linalg.hip).
void destroy_blas(rla_ctx* ctx);

// Moves a device allocation of the library from old_bytes to new_bytes on the
// device's HBM account (0 -> n charges, n -> 0 releases).  Growth past a
// non-zero budget fails with RLA_EADMIT and leaves the account unchanged.
int hbm_resize(int device, size_t old_bytes, size_t new_bytes, const char* what);

// Makes ctx->device current for the calling thread.
int use_device(rla_ctx* ctx);

// Returns a device workspace of at least `bytes` (grow-only, owned by ctx).
// Growth synchronises the context stream before freeing the old block.
int workspace(rla_ctx* ctx, size_t bytes, void** out);

// In-place inclusive prefix sum of a[0..n) on the context's stream.
int scan_inclusive_i32(rla_ctx* ctx, int32_t* a, int64_t n);

// Row-group tile apply of a Float32 X (sparse.hip, rla_sparse_apply_tile_f32)
// with the ranges' sums in `part` (sparse_tile_part_bytes bytes), or in the
// context workspace when part is NULL.
int sparse_apply_tile_part(rla_ctx* ctx, int64_t rows, int64_t in_rows, int vals,
                           int64_t rec_words, const uint32_t* tiles, double magnitude,
                           int64_t n_rhs, const float* X, int64_t x_rs, int64_t x_cs, float* C,
                           int64_t c_rs, int64_t c_cs, double alpha, double beta, float* part);
size_t sparse_tile_part_bytes(int64_t rows, int64_t in_rows, int64_t n_rhs);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

inline int ilog2(int64_t n) {
  int l = 0;
  while ((int64_t(1) << l) < n) ++l;
  return l;
}

// gemm_mfma.hip: C = alpha A B + beta C on strided views, FP32 / FP64 MFMA
template <class T>
int gemm_mfma(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, double alpha, const T* A,
              int64_t a_rs, int64_t a_cs, const T* B, int64_t b_rs, int64_t b_cs, double beta,
              T* C, int64_t c_rs, int64_t c_cs);

// svd_jacobi.hip: X (or X') = U diag(S) V' for k <= 288, one-sided Jacobi
size_t svd_small_work_bytes(int64_t k);
int svd_small_f64(rla_ctx* ctx, int64_t k, const double* X, int64_t ldx, int transpose,
                  double* U, int64_t ldu, double* S, double* V, int64_t ldv, void* work,
                  size_t work_bytes, int* sweeps_out);

// srht_fused.hip: stages 0..13 of Float32 FWHTs of length 2^L, L >= 14
int fwht_chunks_f32(rla_ctx* ctx, const float* src, int64_t src_cs, int64_t m_valid, float* P,
                    int64_t ldp, int64_t ncols, int L, const uint64_t* signs, double scale);

}  // namespace rla
