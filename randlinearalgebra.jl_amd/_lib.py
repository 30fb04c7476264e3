"""ctypes binding of librla_mi355x.so (the C ABI declared in include/rla.h).

The host mirror calls the library only through this module.  There is no
CPU fallback: if the shared library is missing, or no GPU is visible, calls
raise instead of computing anything on the host.
"""
import ctypes
import itertools
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librla_mi355x.so")

RLA_OK, RLA_EARG, RLA_EDIM, RLA_EDOMAIN, RLA_ENOMEM, RLA_EADMIT, RLA_EHIP, RLA_ENOTSUP = range(8)


class ArgumentError(ValueError):
    """Julia's ArgumentError (bad parameter or unsupported combination)."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


class DomainError(ValueError):
    """Julia's DomainError."""


class AdmissionError(RuntimeError):
    """Request exceeds the device-memory budget."""


class DeviceError(RuntimeError):
    """HIP runtime failure or device allocation failure."""


_ERR = {
    RLA_EARG: ArgumentError,
    RLA_EDIM: DimensionMismatch,
    RLA_EDOMAIN: DomainError,
    RLA_ENOMEM: DeviceError,
    RLA_EADMIT: AdmissionError,
    RLA_EHIP: DeviceError,
    RLA_ENOTSUP: ArgumentError,
}

i32, i64, u64, f64, vp = (ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double,
                          ctypes.c_void_p)

# name -> argtypes (restype int unless listed in _RESTYPE)
_SIGS = {
    "rla_version": [],
    "rla_last_error": [],
    "rla_ctx_create": [i32, vp, vp],
    "rla_ctx_destroy": [vp],
    "rla_ctx_set_stream": [vp, vp],
    "rla_ctx_sync": [vp],
    "rla_hbm_set_budget": [i32, ctypes.c_size_t],
    "rla_hbm_admit": [i32, ctypes.c_size_t],
    "rla_hbm_release": [i32, ctypes.c_size_t],
    "rla_hbm_query": [i32, vp, vp],
    "rla_ctx_workspace_bytes": [vp, vp],
    "rla_philox4x32_10": [vp, vp, vp],
    "rla_index_from_julia": [vp, vp, i64, i64, vp],
    "rla_sparse_sign_gen_f32": [vp, u64, u64, i64, i64, i64, vp, vp],
    "rla_sparse_sign_gen_f64": [vp, u64, u64, i64, i64, i64, vp, vp],
    "rla_host_sparse_sign_gen_f64": [u64, u64, i64, i64, i64, vp, vp],
    "rla_count_sketch_gen_f32": [vp, u64, u64, i64, i64, vp, vp],
    "rla_count_sketch_gen_f64": [vp, u64, u64, i64, i64, vp, vp],
    "rla_host_count_sketch_gen_f64": [u64, u64, i64, i64, vp, vp],
    "rla_srht_gen": [vp, u64, u64, i64, i64, vp, vp],
    "rla_host_srht_gen": [u64, u64, i64, i64, vp, vp],
    "rla_gaussian_gen_f32": [vp, u64, u64, i64, i64, f64, vp, i64],
    "rla_gaussian_gen_f64": [vp, u64, u64, i64, i64, f64, vp, i64],
    "rla_host_gaussian_gen_f64": [u64, u64, i64, i64, f64, vp, i64],
    "rla_sparse_transpose_f32": [vp, i64, i64, i64, vp, vp, vp, vp, vp, vp],
    "rla_sparse_transpose_f64": [vp, i64, i64, i64, vp, vp, vp, vp, vp, vp],
    "rla_sparse_apply_f32": [vp, i32, i64, i64, vp, vp, vp, vp, vp, vp, i64, vp, i64, i64, vp,
                             i64, i64, f64, f64],
    "rla_sparse_apply_f64": [vp, i32, i64, i64, vp, vp, vp, vp, vp, vp, i64, vp, i64, i64, vp,
                             i64, i64, f64, f64],
    "rla_sparse_ell_width": [vp, i64, vp, vp],
    "rla_sparse_bell_widths": [vp, i64, i64, i64, vp, vp, vp, vp],
    "rla_sparse_bell_build_f32": [vp, i64, i64, i64, vp, vp, vp, vp, vp, vp],
    "rla_sparse_bell_build_f64": [vp, i64, i64, i64, vp, vp, vp, vp, vp, vp],
    "rla_sparse_apply_bell_f32": [vp, i64, i64, i64, vp, vp, vp, i64, vp, i64, i64, vp, i64, i64,
                                  f64, f64],
    "rla_sparse_apply_bell_f64": [vp, i64, i64, i64, vp, vp, vp, i64, vp, i64, i64, vp, i64, i64,
                                  f64, f64],
    "rla_sparse_apply_bell_split_f32": [vp, i64, i64, i64, vp, vp, vp, i64, vp, i64, i64, vp, i64,
                                        i64, f64, f64],
    "rla_sparse_apply_bell_split_f64": [vp, i64, i64, i64, vp, vp, vp, i64, vp, i64, i64, vp, i64,
                                        i64, f64, f64],
    "rla_sparse_apply_bell_split_bytes": [i64, i64, i64, i64, vp],
    "rla_sparse_tile_shape": [vp, i64, i64, vp, vp, i32, vp, vp],
    "rla_sparse_tile_build_f32": [vp, i64, i64, vp, vp, vp, i32, i64, vp, vp],
    "rla_sparse_apply_tile_bytes": [i64, i64, i64, vp],
    "rla_sparse_apply_tile_f32": [vp, i64, i64, i32, i64, vp, f64, i64, vp, i64, i64, vp, i64, i64,
                                  f64, f64],
    "rla_fjlt_apply_tile_f32": [vp, i64, i64, i64, i64, i64, vp, vp, f64, vp, i64, i64, vp, i64,
                                i64, f64, f64],
    "rla_fjlt_apply_bell_f32": [vp, i64, i64, i64, i64, i64, vp, vp, vp, vp, f64, vp, i64, i64, vp,
                                i64, i64, f64, f64],
    "rla_fjlt_apply_bell_f64": [vp, i64, i64, i64, i64, i64, vp, vp, vp, vp, f64, vp, i64, i64, vp,
                                i64, i64, f64, f64],
    "rla_agmon_select_f32": [vp, i64, vp, i64, vp, i64, vp],
    "rla_fjlt_apply_ell_f32": [vp, i64, i64, i64, i64, vp, vp, vp, vp, f64, vp, i64, i64, vp, i64,
                               i64, f64, f64],
    "rla_fjlt_apply_ell_f64": [vp, i64, i64, i64, i64, vp, vp, vp, vp, f64, vp, i64, i64, vp, i64,
                               i64, f64, f64],
    "rla_agmon_select_f64": [vp, i64, vp, i64, vp, i64, vp],
    "rla_sparse_ell_sign_build_f32": [vp, i64, i32, vp, vp, vp, vp, vp],
    "rla_sparse_ell_sign_build_f64": [vp, i64, i32, vp, vp, vp, vp, vp],
    "rla_sparse_apply_ell_sign_f32": [vp, i64, i64, vp, vp, f64, i64, vp, i64, i64, vp, i64, i64,
                                      f64, f64],
    "rla_sparse_apply_ell_sign_f64": [vp, i64, i64, vp, vp, f64, i64, vp, i64, i64, vp, i64, i64,
                                      f64, f64],
    "rla_sparse_ell_build_f32": [vp, i64, i32, vp, vp, vp, vp, vp],
    "rla_sparse_ell_build_f64": [vp, i64, i32, vp, vp, vp, vp, vp],
    "rla_sparse_apply_ell_f32": [vp, i64, i64, vp, vp, vp, i64, vp, i64, i64, vp, i64, i64, f64,
                                 f64],
    "rla_sparse_apply_ell_f64": [vp, i64, i64, vp, vp, vp, i64, vp, i64, i64, vp, i64, i64, f64,
                                 f64],
    "rla_fjlt_gen_count": [vp, u64, u64, i64, i64, f64, vp, vp],
    "rla_fjlt_gen_fill_f32": [vp, u64, u64, i64, i64, f64, vp, vp, vp],
    "rla_fjlt_gen_fill_f64": [vp, u64, u64, i64, i64, f64, vp, vp, vp],
    "rla_host_fjlt_gen_f64": [u64, u64, i64, i64, f64, vp, vp, vp, i64, vp],
    "rla_fjlt_apply_f32": [vp, i32, i64, i64, i64, i64, vp, vp, vp, vp, vp, vp, vp, f64, vp, i64,
                           i64, vp, i64, i64, f64, f64],
    "rla_fjlt_apply_f64": [vp, i32, i64, i64, i64, i64, vp, vp, vp, vp, vp, vp, vp, f64, vp, i64,
                           i64, vp, i64, i64, f64, f64],
    "rla_fwht_f32": [vp, i64, i64, vp, f64, vp, i64],
    "rla_fwht_f64": [vp, i64, i64, vp, f64, vp, i64],
    "rla_srht_apply_f32": [vp, i32, i64, i64, i64, i64, vp, vp, f64, vp, vp, i64, i64, vp, i64,
                           i64, f64, f64],
    "rla_srht_apply_f64": [vp, i32, i64, i64, i64, i64, vp, vp, f64, vp, vp, i64, i64, vp, i64,
                           i64, f64, f64],
    "rla_srht_plan_bytes": [i32, i64, i64, vp],
    "rla_srht_plan": [vp, i32, i64, i64, vp, vp, vp],
    "rla_srht_apply_exact_f32": [vp, i32, i64, i64, i64, i64, vp, vp, f64, vp, vp, i64, i64, vp,
                                 i64, i64, f64, f64],
    "rla_gemm_f32": [vp, i32, i32, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64],
    "rla_gemm_f64": [vp, i32, i32, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64],
    "rla_gemm_f64_mfma": [vp, i32, i32, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64],
    "rla_gemm_strided_f32": [vp, i64, i64, i64, f64, vp, i64, i64, vp, i64, i64, f64, vp, i64,
                             i64],
    "rla_gemm_strided_f64": [vp, i64, i64, i64, f64, vp, i64, i64, vp, i64, i64, f64, vp, i64,
                             i64],
    "rla_gaussian_apply_f32": [vp, i32, i64, i64, vp, i64, i64, i64, i64, vp, i64, i64, f64, f64,
                               vp, i64, i64],
    "rla_gaussian_apply_f64": [vp, i32, i64, i64, vp, i64, i64, i64, i64, vp, i64, i64, f64, f64,
                               vp, i64, i64],
    "rla_gaussian_apply_noalloc_f32": [vp, i32, i64, i64, vp, i64, i64, i64, i64, vp, i64, i64,
                                       f64, f64, vp, i64, i64],
    "rla_to_bf16_f32": [vp, i64, i64, vp, i64, i64, vp, i64],
    "rla_to_bf16_f64": [vp, i64, i64, vp, i64, i64, vp, i64],
    "rla_gemm_bf16": [vp, i64, i64, i64, ctypes.c_float, vp, i64, vp, i64, ctypes.c_float, vp,
                      i64, i64],
    "rla_to_bf16_tiled_bytes": [i64, i64, vp],
    "rla_to_bf16_tiled_f32": [vp, i64, i64, vp, i64, i64, vp],
    "rla_to_bf16_tiled_f64": [vp, i64, i64, vp, i64, i64, vp],
    "rla_gemm_bf16_tiled": [vp, i64, i64, i64, ctypes.c_float, vp, vp, i64, ctypes.c_float, vp,
                            i64, i64],
    "rla_convert_f32_f64": [vp, i64, i64, vp, i64, vp, i64],
    "rla_convert_f64_f32": [vp, i64, i64, vp, i64, vp, i64],
    "rla_scale_cols_f64": [vp, i64, i64, vp, i64, vp, vp, i64],
    "rla_qr_workspace_bytes": [i64, i64, vp],
    "rla_qr_f64": [vp, i64, i64, vp, i64, vp, i64, vp, ctypes.c_size_t, vp],
    "rla_trsv_upper_f64": [vp, i32, i64, vp, i64, vp],
    "rla_chol_inv_f64": [vp, i64, vp, i64, vp, i64, vp, i64, vp],
    "rla_row_norms2_f32": [vp, i64, i64, vp, i64, i64, vp],
    "rla_row_norms2_f64": [vp, i64, i64, vp, i64, i64, vp],
    "rla_wsample": [vp, u64, u64, i64, vp, i64, ctypes.c_int, i64, vp],
    "rla_residual_norm_f32": [vp, i64, i64, vp, i64, i64, vp, vp, vp],
    "rla_kaczmarz_workspace_bytes": [i64, i64, i64, i64, vp],
    "rla_kaczmarz_lq_workspace_bytes": [i64, i64, vp],
    "rla_kaczmarz_lq_update_f32": [vp, i64, i64, vp, i64, i64, vp, i64, vp, ctypes.c_double, vp,
                                   ctypes.c_size_t],
    "rla_kaczmarz_lq_update_f64": [vp, i64, i64, vp, i64, i64, vp, i64, vp, ctypes.c_double, vp,
                                   ctypes.c_size_t],
    "rla_residual_norm_f64": [vp, i64, i64, vp, i64, i64, vp, vp, vp],
    "rla_kaczmarz_block_update_f32": [vp, i64, i64, vp, i64, i64, vp, i64, vp, ctypes.c_double,
                                      vp, ctypes.c_size_t],
    "rla_kaczmarz_solve_f32": [vp, i64, i64, vp, i64, vp, i64, vp, vp, i64, ctypes.c_double, vp,
                               ctypes.c_int, u64, u64, i64, i64, ctypes.c_double, vp, i64, vp,
                               vp, vp, vp, ctypes.c_size_t],
    "rla_rangefinder_workspace_bytes": [i32, i32, i64, i64, i64, vp],
    "rla_rangefinder": [vp, i32, i32, i64, i64, i64, i64, i32, vp, i64, vp, i64, vp, i64, vp,
                        i64, vp, ctypes.c_size_t],
    "rla_svd_small_workspace_bytes": [i64, vp],
    "rla_svd_small_f64": [vp, i64, vp, i64, i32, vp, i64, vp, vp, i64, vp, ctypes.c_size_t, vp],
    "rla_randsvd_workspace_bytes": [i32, i32, i64, i64, i64, vp],
    "rla_randsvd_finish": [vp, i32, i64, i64, i64, i32, vp, i64, vp, i64, vp, i64, vp, i64, vp,
                           vp, i64, vp, ctypes.c_size_t],
}
_RESTYPE = {"rla_version": ctypes.c_char_p, "rla_last_error": ctypes.c_char_p,
            "rla_philox4x32_10": None}

_lib = None
_lock = threading.Lock()


def exported_symbols():
    """Names of every entry point the host mirror binds (all declared in rla.h)."""
    return sorted(_SIGS)


def load():
    """Loads and binds the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                    "(make -C randlinearalgebra.jl_amd); there is no CPU fallback")
            lib = ctypes.CDLL(LIB_PATH)
            for name, args in _SIGS.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = _RESTYPE.get(name, ctypes.c_int)
            _lib = lib
    return _lib


def last_error():
    return load().rla_last_error().decode()


def check(rc, what=""):
    if rc != RLA_OK:
        msg = last_error()
        raise _ERR.get(rc, RuntimeError)(msg if msg else f"{what} failed with status {rc}")


class _Ctx:
    def __init__(self, device):
        self.device = device
        self.handle = ctypes.c_void_p()
        self.stream = None

    def bind(self):
        lib = load()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        if not self.handle:
            check(lib.rla_ctx_create(self.device, ctypes.c_void_p(stream),
                                     ctypes.byref(self.handle)), "rla_ctx_create")
        elif stream != self.stream:
            check(lib.rla_ctx_set_stream(self.handle, ctypes.c_void_p(stream)))
        self.stream = stream
        return self.handle


_ctxs = {}
_ctxs_lock = threading.Lock()


def _destroy(keys):
    lib = load()
    for key in keys:
        _ctx_use.pop(key, None)
        c = _ctxs.pop(key, None)
        if c is not None and c.handle:
            lib.rla_ctx_destroy(c.handle)  # waits for its stream, frees its workspace


def drop_thread_contexts():
    """Destroys the calling thread's contexts (a worker thread about to exit)."""
    me = threading.get_ident()
    with _ctxs_lock:
        _destroy([k for k in _ctxs if k[0] == me])


# contexts one thread keeps per device (one per stream it uses), least
# recently used destroyed first
MAX_STREAM_CONTEXTS = 8
_ctx_use = {}                 # key -> tick of its last use
_ticks = itertools.count(1)


def ctx(device=None):
    """The calling thread's context for `device` and torch's current stream.

    A context owns a workspace that its calls use asynchronously, so one
    context serves one stream: calls on two streams from one thread then run
    concurrently instead of being ordered through a shared workspace."""
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the rla kernels run on the device only")
    if device is None:
        device = torch.cuda.current_device()
    elif isinstance(device, torch.device):
        device = device.index if device.index is not None else torch.cuda.current_device()
    me = threading.get_ident()
    stream = torch.cuda.current_stream(device).cuda_stream
    key = (me, device, stream)
    c = _ctxs.get(key)
    if c is None:
        with _ctxs_lock:
            # contexts of threads that have ended hold workspaces nobody uses
            alive = {t.ident for t in threading.enumerate()}
            _destroy([k for k in _ctxs if k[0] not in alive])
            mine = sorted((k for k in _ctxs if k[0] == me and k[1] == device),
                          key=lambda k: _ctx_use.get(k, 0))
            _destroy(mine[:max(0, len(mine) + 1 - MAX_STREAM_CONTEXTS)])
            c = _Ctx(device)
            _ctxs[key] = c
    _ctx_use[key] = next(_ticks)
    return c.bind()


def call(name, *args):
    """Calls an entry point and raises the mapped exception on failure."""
    rc = getattr(load(), name)(*args)
    check(rc, name)


def call_rc(name, *args):
    """Calls an entry point and returns its status without raising."""
    return int(getattr(load(), name)(*args))


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def sfx(dtype):
    if dtype == torch.float32:
        return "f32"
    if dtype == torch.float64:
        return "f64"
    raise ArgumentError(f"element type {dtype} is not supported (Float32 / Float64)")
