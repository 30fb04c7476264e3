// Context, error reporting, workspace and index conversion for
// librla_mi355x.so.
#include <mutex>
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

// ---- HBM admission account, one per device, process-wide -----------------
// Admitted request bytes and the library's own device allocations (context
// workspaces, batched least-squares plans) are charged to the same account;
// a charge that would take `used` past a non-zero budget fails with
// RLA_EADMIT and changes nothing.
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_hbm_mu;
size_t g_hbm_budget[kMaxDevices];
size_t g_hbm_used[kMaxDevices];
}  // namespace

static int hbm_device_ok(int device, const char* what) {
  if (device < 0 || device >= kMaxDevices)
    return fail(RLA_EARG, std::string(what) + ": device " + std::to_string(device) +
                              " out of range");
  return RLA_OK;
}

int hbm_resize(int device, size_t old_bytes, size_t new_bytes, const char* what) {
  int rc = hbm_device_ok(device, what);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(g_hbm_mu);
  size_t& used = g_hbm_used[device];
  const size_t base = used >= old_bytes ? used - old_bytes : 0;
  const size_t budget = g_hbm_budget[device];
  if (budget && new_bytes > old_bytes && (new_bytes > budget || base > budget - new_bytes))
    return fail(RLA_EADMIT, std::string(what) + ": " + std::to_string(new_bytes) +
                                " bytes exceed the HBM budget of device " +
                                std::to_string(device) + " (" + std::to_string(base) + " of " +
                                std::to_string(budget) + " bytes in use)");
  used = base + new_bytes;
  return RLA_OK;
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
  int rc = hbm_resize(ctx->device, ctx->ws_bytes, want, "device workspace");
  if (rc) return rc;  // the old workspace stays
  if (ctx->ws) {
    RLA_HIP(hipStreamSynchronize(ctx->stream));
    RLA_HIP(hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  void* p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    (void)hbm_resize(ctx->device, want, 0, "device workspace");
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
  (void)hbm_resize(ctx->device, ctx->ws_bytes, 0, "rla_ctx_destroy");
  destroy_blas(ctx);
  if (ctx->order_ev) (void)hipEventDestroy(ctx->order_ev);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RLA_OK;
}

int rla_ctx_set_stream(rla_ctx* ctx, void* hip_stream) {
  RLA_CHECK_CTX(ctx);
  hipStream_t next = static_cast<hipStream_t>(hip_stream);
  if (next == ctx->stream) return RLA_OK;
  int rc = use_device(ctx);
  if (rc) return rc;
  if (ctx->own_stream) {
    RLA_HIP(hipStreamSynchronize(ctx->stream));
    (void)hipStreamDestroy(ctx->stream);
    ctx->own_stream = false;
  } else {
    // Work already enqueued on the old stream may still use the context's
    // workspace and scratch: the new stream waits for it before anything
    // enqueued from now on.
    if (!ctx->order_ev) RLA_HIP(hipEventCreateWithFlags(&ctx->order_ev, hipEventDisableTiming));
    RLA_HIP(hipEventRecord(ctx->order_ev, ctx->stream));
    RLA_HIP(hipStreamWaitEvent(next, ctx->order_ev, 0));
  }
  ctx->stream = next;
  return RLA_OK;
}

int rla_hbm_set_budget(int device, size_t bytes) {
  int rc = hbm_device_ok(device, "rla_hbm_set_budget");
  if (rc) return rc;
  std::lock_guard<std::mutex> g(g_hbm_mu);
  g_hbm_budget[device] = bytes;
  return RLA_OK;
}

int rla_hbm_admit(int device, size_t bytes) {
  return hbm_resize(device, 0, bytes, "rla_hbm_admit");
}

int rla_hbm_release(int device, size_t bytes) {
  int rc = hbm_device_ok(device, "rla_hbm_release");
  if (rc) return rc;
  std::lock_guard<std::mutex> g(g_hbm_mu);
  if (bytes > g_hbm_used[device])
    return fail(RLA_EARG, "rla_hbm_release: " + std::to_string(bytes) +
                              " bytes is more than the " + std::to_string(g_hbm_used[device]) +
                              " bytes in use on device " + std::to_string(device));
  g_hbm_used[device] -= bytes;
  return RLA_OK;
}

int rla_hbm_query(int device, size_t* used, size_t* budget) {
  int rc = hbm_device_ok(device, "rla_hbm_query");
  if (rc) return rc;
  std::lock_guard<std::mutex> g(g_hbm_mu);
  if (used) *used = g_hbm_used[device];
  if (budget) *budget = g_hbm_budget[device];
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
