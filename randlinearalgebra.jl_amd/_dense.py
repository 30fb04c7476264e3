"""Dense plumbing for the host mirror: strided views -> column-major GEMM
calls, BF16 copies and precision conversions, all through librla_mi355x."""
import ctypes

import torch

from . import _lib
from ._lib import ArgumentError


def _colmajor(v):
    """(pointer tensor, transposed?, leading dim) for a View usable by BLAS."""
    if v.rs == 1 and v.cs >= max(1, v.rows):
        return v.t, 0, v.cs
    if v.cs == 1 and v.rs >= max(1, v.cols):
        return v.t, 1, v.rs
    if v.cols == 1 and v.rs == 1:
        return v.t, 0, max(1, v.rows)
    return None


def gemm(C, A, B, alpha=1.0, beta=0.0):
    """C = alpha*A*B + beta*C on strided Views (Float32 / Float64), on the
    library's FP32 / FP64 MFMA GEMM (any strides)."""
    m, n, k = C.rows, C.cols, A.cols
    if m == 0 or n == 0:
        return
    _lib.call(f"rla_gemm_strided_{_lib.sfx(C.dtype)}", _lib.ctx(C.t.device), m, n, k,
              float(alpha), _lib.ptr(A.t), A.rs, A.cs, _lib.ptr(B.t), B.rs, B.cs, float(beta),
              _lib.ptr(C.t), C.rs, C.cs)


class SourceKey:
    """Identity of a caller's tensor, for caches of data derived from it.

    The key keeps a reference to the tensor, so its memory cannot be freed and
    handed to another allocation while the key lives: an equal data pointer
    then means the same memory, and the version counter, which every view of
    the storage shares, catches in-place changes.  Derived data must be keyed
    on the caller's tensor, never on a temporary made from it (a temporary's
    counter starts at zero and its address is reused by the next one)."""

    __slots__ = ("ref", "sig")

    def __init__(self, t):
        self.ref = t
        self.sig = self._sig(t)

    @staticmethod
    def _sig(t):
        return (t.data_ptr(), t._version, tuple(t.shape), tuple(t.stride()), t.dtype, t.device)

    def matches(self, t):
        return self._sig(t) == self.sig


def key_matches(key, t):
    return key is not None and key.matches(t)


def ld8(n):
    return (n + 7) & ~7


def bf16_rows(v):
    """BF16 copy of the (rows x cols) View with each row contiguous
    (row i at out[i, :cols]); rows padded to a multiple of 8 entries."""
    out = torch.empty(v.rows, ld8(v.cols), dtype=torch.bfloat16, device=v.t.device)
    _lib.call(f"rla_to_bf16_{_lib.sfx(v.dtype)}", _lib.ctx(v.t.device), v.rows, v.cols,
              _lib.ptr(v.t), v.rs, v.cs, _lib.ptr(out), ld8(v.cols))
    return out


def bf16_tiled(v):
    """Tiled BF16 copy of the (rows x cols) View: the layout in which the BF16
    GEMM streams its A operand (rla_to_bf16_tiled: 128 x 64 tiles, one
    contiguous kilobyte per wavefront load, padding zero)."""
    nbytes = query_bytes("rla_to_bf16_tiled_bytes", v.rows, v.cols)
    out = torch.empty(max(1, nbytes // 2), dtype=torch.bfloat16, device=v.t.device)
    _lib.call(f"rla_to_bf16_tiled_{_lib.sfx(v.dtype)}", _lib.ctx(v.t.device), v.rows, v.cols,
              _lib.ptr(v.t), v.rs, v.cs, _lib.ptr(out))
    return out


def gemm_bf16_tiled(M, N, K, A_tiled, B_cols, C, alpha=1.0, beta=0.0):
    """gemm_bf16 with A (M x K) as a tiled copy from bf16_tiled (N <= 288)."""
    _lib.call("rla_gemm_bf16_tiled", _lib.ctx(C.t.device), M, N, K, float(alpha),
              _lib.ptr(A_tiled), _lib.ptr(B_cols), B_cols.shape[1], float(beta), _lib.ptr(C.t),
              C.rs, C.cs)


def gemm_bf16(M, N, K, A_rows, B_cols, C, alpha=1.0, beta=0.0):
    """C (Float32 View, M x N) = alpha * A_rows[:M, :K] @ B_cols[:N, :K]' + beta*C,
    with both operands BF16 and K-contiguous (from bf16_rows)."""
    _lib.call("rla_gemm_bf16", _lib.ctx(C.t.device), M, N, K, float(alpha), _lib.ptr(A_rows),
              A_rows.shape[1], _lib.ptr(B_cols), B_cols.shape[1], float(beta), _lib.ptr(C.t),
              C.rs, C.cs)


def convert(src, dtype):
    """Column-major copy of a column-major device matrix in another precision.
    With dtype equal to the source's, a source that is already column-major
    and dense is returned itself (no copy)."""
    from .compressors import as_view
    v = as_view(src)
    if v.rs != 1:
        raise ArgumentError("convert needs a column-major source")
    if src.dtype == dtype:
        return src.clone() if src.dim() == 1 else src.T.contiguous().T
    shape = (v.cols, v.rows)
    out = torch.empty(shape, dtype=dtype, device=src.device).T
    name = "rla_convert_f32_f64" if dtype == torch.float64 else "rla_convert_f64_f32"
    _lib.call(name, _lib.ctx(src.device), v.rows, v.cols, _lib.ptr(src), max(1, v.cs),
              _lib.ptr(out), max(1, v.rows))
    return out if src.dim() == 2 else out[:, 0]


def workspace(nbytes, device):
    return torch.empty(max(1, (int(nbytes) + 7) // 8), dtype=torch.float64, device=device)


def query_bytes(name, *args):
    out = ctypes.c_size_t(0)
    _lib.call(name, *args, ctypes.byref(out))
    return out.value
