// Context, error reporting, workspace and index conversion for
// librla_mi355x.so.
#include <string>

#include "philox.hpp"
#include "rla_internal.hpp"

namespace rla {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int use_device(rla_ctx* ctx) {
  RLA_HIP(hipSetDevice(ctx->device));
  return RLA_OK;
}

int workspace(rla_ctx* ctx, size_t bytes, void** out) {
  if (bytes <= ctx->ws_bytes && ctx->ws) {
    *out = ctx->ws;
    return RLA_OK;
  }
  size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes;
  want = (want + 255) & ~size_t(255);
  if (ctx->ws) {
    RLA_HIP(hipStreamSynchronize(ctx->stream));
    RLA_HIP(hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  void* p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    return fail(RLA_ENOMEM, "device workspace allocation of " + std::to_string(want) +
                                " bytes failed");
  }
  ctx->ws = p;
  ctx->ws_bytes = want;
  *out = p;
  return RLA_OK;
}

// out-of-range detection for rla_index_from_julia: flag gets the first bad
// position + 1 (atomicMin over positions keeps the report deterministic).
__global__ void index_from_julia_kernel(const int64_t* __restrict__ in, int64_t n,
                                        int64_t maxval, int32_t* __restrict__ out,
                                        unsigned long long* bad) {
  int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  int64_t v = in[i];
  if (v < 1 || v > maxval) {
    atomicMin(bad, static_cast<unsigned long long>(i));
    out[i] = 0;
  } else {
    out[i] = static_cast<int32_t>(v - 1);
  }
}

__global__ void set_u64_kernel(unsigned long long* p, unsigned long long v) { *p = v; }

}  // namespace rla

using namespace rla;

extern "C" {

const char* rla_version(void) { return "rla-mi355x 0.1.0 (gfx950)"; }

const char* rla_last_error(void) { return g_last_error.c_str(); }

int rla_ctx_create(int device, void* hip_stream, rla_ctx** out) {
  if (!out) return fail(RLA_EARG, "rla_ctx_create: out is NULL");
  int ndev = 0;
  RLA_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(RLA_EARG, "rla_ctx_create: device " + std::to_string(device) +
                              " out of range (" + std::to_string(ndev) + " devices)");
  RLA_HIP(hipSetDevice(device));
  rla_ctx* c = new rla_ctx();
  c->device = device;
  c->stream = static_cast<hipStream_t>(hip_stream);  // NULL: the null stream
  c->own_stream = false;
  *out = c;
  return RLA_OK;
}

int rla_ctx_destroy(rla_ctx* ctx) {
  if (!ctx) return RLA_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->ws) (void)hipFree(ctx->ws);
  destroy_blas(ctx);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RLA_OK;
}

int rla_ctx_set_stream(rla_ctx* ctx, void* hip_stream) {
  RLA_CHECK_CTX(ctx);
  if (ctx->own_stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    ctx->own_stream = false;
  }
  ctx->stream = static_cast<hipStream_t>(hip_stream);
  return RLA_OK;
}

int rla_ctx_sync(rla_ctx* ctx) {
  RLA_CHECK_CTX(ctx);
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  return RLA_OK;
}

int rla_ctx_workspace_bytes(rla_ctx* ctx, size_t* out) {
  RLA_CHECK_CTX(ctx);
  if (!out) return fail(RLA_EARG, "rla_ctx_workspace_bytes: out is NULL");
  *out = ctx->ws_bytes;
  return RLA_OK;
}

void rla_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  u32x4 c;
  for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
  u32x4 r = philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

int rla_index_from_julia(rla_ctx* ctx, const int64_t* idx1, int64_t n, int64_t maxval,
                         int32_t* idx0) {
  RLA_CHECK_CTX(ctx);
  if (n < 0) return fail(RLA_EARG, "rla_index_from_julia: negative length");
  if (maxval > INT32_MAX) return fail(RLA_ENOTSUP, "rla_index_from_julia: maxval exceeds int32");
  if (n == 0) return RLA_OK;
  if (!idx1 || !idx0) return fail(RLA_EARG, "rla_index_from_julia: NULL pointer");
  int rc = use_device(ctx);
  if (rc) return rc;
  void* ws = nullptr;
  rc = workspace(ctx, 256, &ws);
  if (rc) return rc;
  unsigned long long* bad = static_cast<unsigned long long*>(ws);
  set_u64_kernel<<<1, 1, 0, ctx->stream>>>(bad, ~0ull);
  int64_t blocks = ceil_div(n, 256);
  index_from_julia_kernel<<<dim3(static_cast<unsigned>(blocks)), 256, 0, ctx->stream>>>(
      idx1, n, maxval, idx0, bad);
  RLA_LAUNCH_CHECK("index_from_julia");
  unsigned long long h = 0;
  RLA_HIP(hipMemcpyAsync(&h, bad, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  RLA_HIP(hipStreamSynchronize(ctx->stream));
  if (h != ~0ull)
    return fail(RLA_EARG, "index at position " + std::to_string(h + 1) + " is outside 1:" +
                              std::to_string(maxval));
  return RLA_OK;
}

}  // extern "C"
