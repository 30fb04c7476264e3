"""Compressor surface of RandLinearAlgebra.jl on the MI355X kernels.

Mirrors src/Compressors.jl: a `Compressor` holds user parameters, a
`CompressorRecipe` holds the realised operator for a particular matrix (here:
device tensors), `complete_compressor` builds a recipe, `update_compressor`
redraws it in place, and `mul(C, S, A, alpha, beta)` / `mul(C, A, S, alpha,
beta)` are the reference's 5-argument `mul!` (3-argument forms default
alpha = 1.0, beta = 0.0, Compressors.jl:389-410).  `adjoint(S)` (or `S.T`)
gives the `CompressorAdjoint` of Compressors.jl:94-104, and `S @ A` / `A @ S`
are the allocating `*` of Compressors.jl:416-449.

Matrices are torch tensors on the GPU, indexed like the Julia matrices they
stand for (A[i, j] is row i, column j); any strides are accepted, so Julia
views, transposes and column-major storage all map directly.  All arithmetic
runs in librla_mi355x.so; nothing here computes on the host.
"""
import ctypes
import math

import torch

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, DomainError

# ---------------------------------------------------------------------------
# cardinality (Compressors.jl:24-45)


class Cardinality:
    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self).__name__)

    def __repr__(self):
        return f"{type(self).__name__}()"


class Left(Cardinality):
    pass


class Right(Cardinality):
    pass


class Undef(Cardinality):
    pass


# ---------------------------------------------------------------------------
# global draw state (the reference draws from Julia's task-local RNG; here a
# recipe is a pure function of (seed, epoch), and every complete/update call
# takes the next epoch)

_rng = {"seed": 0x5EED_0000_1234, "epoch": 0}


def seed(s):
    """Counterpart of Random.seed!(s) for recipe generation."""
    _rng["seed"] = int(s) & 0xFFFFFFFFFFFFFFFF
    _rng["epoch"] = 0


def _draw(count=1):
    """(seed, first epoch) for `count` consecutive draws."""
    e = _rng["epoch"]
    _rng["epoch"] = e + count
    return _rng["seed"], e


# ---------------------------------------------------------------------------
# strided views


class View:
    """A (rows x cols) strided window on a device tensor's storage."""

    __slots__ = ("t", "rows", "cols", "rs", "cs")

    def __init__(self, t, rows, cols, rs, cs):
        self.t, self.rows, self.cols, self.rs, self.cs = t, rows, cols, rs, cs

    @property
    def T(self):
        return View(self.t, self.cols, self.rows, self.cs, self.rs)

    @property
    def dtype(self):
        return self.t.dtype

    def size(self, d=None):
        s = (self.rows, self.cols)
        return s if d is None else s[d - 1]


def as_view(x):
    if isinstance(x, View):
        return x
    if not isinstance(x, torch.Tensor):
        raise ArgumentError(f"expected a torch tensor, got {type(x).__name__}")
    if x.dim() == 1:
        n = x.shape[0]
        return View(x, n, 1, x.stride(0) if n > 1 else 1, max(n, 1))
    if x.dim() == 2:
        return View(x, x.shape[0], x.shape[1], x.stride(0), x.stride(1))
    raise ArgumentError("only vectors and matrices are supported")


def _size2(x):
    v = as_view(x)
    return v.rows, v.cols


def _check_device(*vs):
    for v in vs:
        if not v.t.is_cuda:
            raise ArgumentError("arrays must live on the GPU (torch cuda tensors)")


# ---------------------------------------------------------------------------
# abstract types


class Compressor:
    """Parameters of a compression technique (Compressors.jl:7)."""

    cardinality = Left()


class CompressorRecipe:
    """A realised compression operator (Compressors.jl:16); subclasses set
    n_rows and n_cols."""

    n_rows = 0
    n_cols = 0

    def size(self, dim=None):
        if dim is None:
            return (self.n_rows, self.n_cols)
        if dim < 1 or dim > 2:
            raise DomainError("`dim` must be 1 or 2.")
        return self.n_rows if dim == 1 else self.n_cols

    @property
    def T(self):
        return CompressorAdjoint(self)

    def __matmul__(self, A):
        return times(self, A)

    def __rmatmul__(self, A):
        return times(A, self)

    # implemented by concrete recipes
    def _left(self, C, A, alpha, beta):
        raise ArgumentError(
            f"No method `mul!` defined for ({type(C).__name__}, {type(self).__name__}, "
            f"{type(A).__name__}, Float64, Float64).")

    def _right(self, C, A, alpha, beta):
        raise ArgumentError(
            f"No method `mul!` defined for ({type(C).__name__}, {type(A).__name__}, "
            f"{type(self).__name__}, Float64, Float64).")


class CompressorAdjoint:
    """Adjoint of a recipe (Compressors.jl:94-104)."""

    def __init__(self, parent):
        self.parent = parent

    def size(self, dim=None):
        p = self.parent
        if dim is None:
            return (p.n_cols, p.n_rows)
        if dim < 1 or dim > 2:
            raise DomainError("`dim` must be 1 or 2.")
        return p.n_cols if dim == 1 else p.n_rows

    @property
    def T(self):
        return self.parent

    def __matmul__(self, A):
        return times(self, A)

    def __rmatmul__(self, A):
        return times(A, self)


def adjoint(S):
    return S.T


# ---------------------------------------------------------------------------
# dimension checks (src/helpers/mul_dim_checks.jl:41-159), same messages


def left_mul_dimcheck(C, S, A):
    if isinstance(S, CompressorAdjoint):
        return right_mul_dimcheck(as_view(C).T, as_view(A).T, S.parent)
    s_rows, s_cols = S.size()
    a_rows, a_cols = _size2(A)
    c_rows, c_cols = _size2(C)
    if a_rows != s_cols:
        raise DimensionMismatch(f"Matrix A has {a_rows} rows while S has {s_cols} columns.")
    if a_cols != c_cols:
        raise DimensionMismatch(f"Matrix A has {a_cols} columns while C has {c_cols} columns.")
    if c_rows != s_rows:
        raise DimensionMismatch(f"Matrix C has {c_rows} rows while S has {s_rows} rows.")


def right_mul_dimcheck(C, A, S):
    if isinstance(S, CompressorAdjoint):
        return left_mul_dimcheck(as_view(C).T, S.parent, as_view(A).T)
    s_rows, s_cols = S.size()
    a_rows, a_cols = _size2(A)
    c_rows, c_cols = _size2(C)
    if a_cols != s_rows:
        raise DimensionMismatch(f"Matrix A has {a_cols} columns while S has {s_rows} rows.")
    if c_cols != s_cols:
        raise DimensionMismatch(f"Matrix C has {c_cols} columns while S has {s_cols} columns.")
    if c_rows != a_rows:
        raise DimensionMismatch(f"Matrix C has {c_rows} rows while A has {a_rows} rows.")


# ---------------------------------------------------------------------------
# generic entry points


def complete_compressor(compressor, *args):
    """complete_compressor(compressor, A) / (compressor, A, b) / (compressor, x, A, b)
    (Compressors.jl:126-206: the longer forms forward to the matrix form unless
    the technique needs them)."""
    if not isinstance(compressor, Compressor):
        raise ArgumentError(f"No method `complete_compressor` defined for {type(compressor)}.")
    return compressor._complete(*args)


def update_compressor(S, *args):
    """update_compressor!(S[, A[, b]] / [x, A, b]) (Compressors.jl:225-305)."""
    if isinstance(S, CompressorAdjoint):
        S = S.parent
    S._update(*args)


update_compressor_ = update_compressor


_mul_hooks = []


def _register_mul_hook(h):
    """Lets other operator families (approximators) join `mul`."""
    _mul_hooks.append(h)


def mul(C, X, Y, alpha=1.0, beta=0.0):
    """mul!(C, S, A, alpha, beta) or mul!(C, A, S, alpha, beta)."""
    alpha, beta = float(alpha), float(beta)
    for h in _mul_hooks:
        if h(C, X, Y, alpha, beta):
            return None
    if isinstance(X, CompressorAdjoint):
        # alpha*S'*A + beta*C -> C  ==  alpha*A'*S + beta*C' -> C'
        return mul(as_view(C).T, as_view(Y).T, X.parent, alpha, beta)
    if isinstance(Y, CompressorAdjoint):
        return mul(as_view(C).T, Y.parent, as_view(X).T, alpha, beta)
    if isinstance(X, CompressorRecipe):
        left_mul_dimcheck(C, X, Y)
        Cv, Av = as_view(C), as_view(Y)
        _check_device(Cv, Av)
        _check_dtypes(Cv, Av)
        X._left(Cv, Av, alpha, beta)
        return None
    if isinstance(Y, CompressorRecipe):
        right_mul_dimcheck(C, X, Y)
        Cv, Av = as_view(C), as_view(X)
        _check_device(Cv, Av)
        _check_dtypes(Cv, Av)
        Y._right(Cv, Av, alpha, beta)
        return None
    raise ArgumentError("mul needs a CompressorRecipe or CompressorAdjoint operand")


mul_ = mul


def times(X, Y):
    """The allocating product `S * A` / `A * S` (Compressors.jl:416-449); also
    for approximator recipes (Approximators.jl:420-453)."""
    if not isinstance(X, torch.Tensor):
        A = Y
        s_rows = X.size(1)
        C = (torch.zeros(s_rows, dtype=A.dtype, device=A.device) if A.dim() == 1
             else _colmajor_zeros(s_rows, A.shape[1], A))
        mul(C, X, A)
        return C
    A, S = X, Y
    C = _colmajor_zeros(A.shape[0], S.size(2), A)
    mul(C, A, S)
    return C


def _colmajor_zeros(r, c, like):
    return torch.zeros(c, r, dtype=like.dtype, device=like.device).T


def _check_dtypes(Cv, Av):
    if Cv.dtype != Av.dtype:
        raise ArgumentError(f"C and A must share an element type ({Cv.dtype} vs {Av.dtype})")
    _lib.sfx(Av.dtype)


def _dtype(type_):
    if type_ in (torch.float64, float, "Float64", None):
        return torch.float64
    if type_ in (torch.float32, "Float32"):
        return torch.float32
    raise ArgumentError(f"unsupported element type {type_}")


def _device_of(*xs):
    for x in xs:
        if isinstance(x, torch.Tensor) and x.is_cuda:
            return x.device
        if isinstance(x, View) and x.t.is_cuda:
            return x.t.device
    return torch.device("cuda", torch.cuda.current_device())


# ---------------------------------------------------------------------------
# sparse operators: SparseSign, CountSketch


# partial sums the block-split sparse gather may hold (larger: block by block)
_BELL_SPLIT_MAX_BYTES = 64 << 20
# output rows up to which the block-split gather carries each row's sums over
# a run of blocks (2 per thread of a 512-thread workgroup)
_BELL_CARRY_ROWS = 1024
# Float32 X at least this tall and wide (and M of one magnitude with at most
# 1024 rows) takes the row-group tile gather: each workgroup's sums (256 KiB
# for 1024 rows and 64 columns) then stay well below the X it reads
_TILE_MIN_ROWS = 1 << 19
_TILE_MIN_COLS = 16
_TILE_MAX_ROWS = 1024
# FJLT's K (values carried) takes the tile gather up to this many rows (a
# workgroup per 1024 of them)
_FJLT_TILE_MAX_ROWS = 4096


class _SparseOp:
    """An nrows x ncols CSC matrix M on the device and its CSR twin.
    `_uniform`: all nonzeros share one magnitude (the sign-packed gather is
    tried first; its builder verifies it).  `exact`: tall X (columns above
    64 KiB) take the block-by-block gather whose sums are added in
    SparseArrays' entry order (identical bits); otherwise the block-split
    gather, whose per-block FP64 partial sums are within rounding of it."""

    _uniform = False
    exact = False

    def __init__(self, nrows, ncols, colptr, rowidx, val):
        self.nrows, self.ncols = nrows, ncols
        self.colptr, self.rowidx, self.val = colptr, rowidx, val
        self.nnz = rowidx.numel()
        self._vals = {}
        self._ell = {}
        self.rebuild_csr()

    def rebuild_csr(self):
        dev = self.rowidx.device
        self.rowptr = torch.empty(self.nrows + 1, dtype=torch.int32, device=dev)
        self.colidx = torch.empty(self.nnz, dtype=torch.int32, device=dev)
        self.val_t = torch.empty(self.nnz, dtype=self.val.dtype, device=dev)
        _lib.call(f"rla_sparse_transpose_{_lib.sfx(self.val.dtype)}", _lib.ctx(dev), self.nrows,
                  self.ncols, self.nnz, _lib.ptr(self.colptr), _lib.ptr(self.rowidx),
                  _lib.ptr(self.val), _lib.ptr(self.rowptr), _lib.ptr(self.colidx),
                  _lib.ptr(self.val_t))
        self._vals = {}
        self._ell = {}

    def _vals_for(self, dtype):
        if dtype == self.val.dtype:
            return self.val, self.val_t
        if dtype not in self._vals:
            self._vals[dtype] = (self.val.to(dtype), self.val_t.to(dtype))
        return self._vals[dtype]

    def _ell_for(self, trans_m, dtype):
        """Row-interleaved gather arrays of op(M) in `dtype` (built once per
        operator and cached), or None when rows average under 4 entries (the
        per-row reads are then few) or padding every row to the longest would
        more than double the entries."""
        key = (int(trans_m), dtype)
        if key in self._ell:
            return self._ell[key]
        rows = self.ncols if trans_m else self.nrows
        if self.nnz < 4 * rows:
            self._ell[key] = None
            return None
        dev = self.rowidx.device
        ctx = _lib.ctx(dev)
        ptr, idx = (self.colptr, self.rowidx) if trans_m else (self.rowptr, self.colidx)
        val, val_t = self._vals_for(dtype)
        v = val if trans_m else val_t
        w = ctypes.c_int32(0)
        _lib.call("rla_sparse_ell_width", ctx, rows, _lib.ptr(ptr), ctypes.byref(w))
        width = w.value
        out = None
        if rows > 0 and 0 < width and width * rows <= 2 * self.nnz + 64 * rows:
            in_rows = self.nrows if trans_m else self.ncols
            if self._uniform and in_rows * torch.finfo(dtype).bits // 8 <= 160 * 1024:
                # one magnitude: signs packed into the index words, whole X
                # columns staged in LDS
                es = torch.empty(width * rows, dtype=torch.int32, device=dev)
                mag = ctypes.c_double(0.0)
                rc = _lib.call_rc(f"rla_sparse_ell_sign_build_{_lib.sfx(dtype)}", ctx, rows,
                                  width, _lib.ptr(ptr), _lib.ptr(idx), _lib.ptr(v), _lib.ptr(es),
                                  ctypes.byref(mag))
                if rc == 0:
                    out = ("sign", ptr, es, mag.value)
            if out is None:
                ei = torch.empty(width * rows, dtype=torch.int32, device=dev)
                ev = torch.empty(width * rows, dtype=dtype, device=dev)
                _lib.call(f"rla_sparse_ell_build_{_lib.sfx(dtype)}", ctx, rows, width,
                          _lib.ptr(ptr), _lib.ptr(idx), _lib.ptr(v), _lib.ptr(ei), _lib.ptr(ev))
                out = ("ell", ptr, ei, ev)
        self._ell[key] = out
        return out

    def _bell_for(self, trans_m, dtype, block_bytes=65536):
        """Block-interleaved gather arrays of op(M) (X blocks of `block_bytes`)
        in `dtype`, built once per operator and block size and cached: (rb,
        widths, bidx, bval), or None when X fits one block, rows average under
        4 entries, or the copy would hold more than four times the entries
        plus one slot per row and block."""
        key = ("bell", int(trans_m), dtype, block_bytes)
        if key in self._ell:
            return self._ell[key]
        rows, in_rows = (self.ncols, self.nrows) if trans_m else (self.nrows, self.ncols)
        rb = block_bytes // (torch.finfo(dtype).bits // 8)
        out = None
        if rows > 0 and in_rows > rb and self.nnz >= 4 * rows:
            dev = self.rowidx.device
            ctx = _lib.ctx(dev)
            ptr, idx = (self.colptr, self.rowidx) if trans_m else (self.rowptr, self.colidx)
            nb = (in_rows + rb - 1) // rb
            widths = torch.empty(nb, dtype=torch.int32, device=dev)
            total = ctypes.c_int64(0)
            rc = _lib.call_rc("rla_sparse_bell_widths", ctx, rows, in_rows, rb, _lib.ptr(ptr),
                              _lib.ptr(idx), _lib.ptr(widths), ctypes.byref(total))
            if rc == 0 and total.value * rows <= 4 * self.nnz + nb * rows:
                val, val_t = self._vals_for(dtype)
                v = val if trans_m else val_t
                bidx = torch.empty(total.value * rows, dtype=torch.int32, device=dev)
                bval = torch.empty(total.value * rows, dtype=dtype, device=dev)
                _lib.call(f"rla_sparse_bell_build_{_lib.sfx(dtype)}", ctx, rows, in_rows, rb,
                          _lib.ptr(ptr), _lib.ptr(idx), _lib.ptr(v), _lib.ptr(widths),
                          _lib.ptr(bidx), _lib.ptr(bval))
                out = (rb, widths, bidx, bval)
        self._ell[key] = out
        return out

    def _tile_for(self, trans_m, values=False):
        """Row-group tile copy of op(M) for Float32 X (built once per operator
        and cached): (words per record, records, magnitude), or None when M has
        too many rows, nonzeros of several magnitudes (without values), or a
        record would not fit the LDS."""
        key = ("tile", int(trans_m), bool(values))
        if key in self._ell:
            return self._ell[key]
        rows, in_rows = (self.ncols, self.nrows) if trans_m else (self.nrows, self.ncols)
        out = None
        limit = _FJLT_TILE_MAX_ROWS if values else _TILE_MAX_ROWS
        if (values or self._uniform) and 0 < rows <= limit:
            dev = self.rowidx.device
            ctx = _lib.ctx(dev)
            ptr, idx = (self.colptr, self.rowidx) if trans_m else (self.rowptr, self.colidx)
            words, nrec = ctypes.c_int64(0), ctypes.c_int64(0)
            rc = _lib.call_rc("rla_sparse_tile_shape", ctx, rows, in_rows, _lib.ptr(ptr),
                              _lib.ptr(idx), int(values), ctypes.byref(words),
                              ctypes.byref(nrec))
            if rc == 0:
                val, val_t = self._vals_for(torch.float32)
                v = val if trans_m else val_t
                tiles = torch.empty(nrec.value * words.value, dtype=torch.int32, device=dev)
                mag = ctypes.c_double(0.0)
                rc = _lib.call_rc("rla_sparse_tile_build_f32", ctx, rows, in_rows, _lib.ptr(ptr),
                                  _lib.ptr(idx), _lib.ptr(v), int(values), words.value,
                                  _lib.ptr(tiles), ctypes.byref(mag))
                if rc == 0:
                    out = (words.value, tiles, mag.value)
        self._ell[key] = out
        return out

    def apply(self, trans_m, X, C, alpha, beta):
        """C = alpha*op(M)*X + beta*C on strided views (op = M or M')."""
        dt = X.dtype
        if (dt == torch.float32 and X.rs == 1 and not self.exact and X.rows >= _TILE_MIN_ROWS
                and X.cols >= _TILE_MIN_COLS and C.rows <= _TILE_MAX_ROWS):
            tile = self._tile_for(trans_m)
            if tile is not None:
                words, tiles, mag = tile
                _lib.call("rla_sparse_apply_tile_f32", _lib.ctx(X.t.device), C.rows, X.rows, 0,
                          words, _lib.ptr(tiles), mag, X.cols, _lib.ptr(X.t), X.rs, X.cs,
                          _lib.ptr(C.t), C.rs, C.cs, alpha, beta)
                return
        if X.rs == 1 and X.rows > 0 and X.cols > 0:
            col_bytes = X.rows * (torch.finfo(dt).bits // 8)
            if col_bytes > 64 * 1024 and not (self._uniform and col_bytes <= 160 * 1024):
                # X taller than one LDS block (and no whole-column sign path):
                # the block-interleaved copy keeps every metadata load coalesced
                # the block-split apply shares each entry read among up to 8
                # columns staged side by side: 16 KiB blocks let 8 fit the LDS
                from . import _dense
                small = not self.exact and X.cols >= 2 and C.rows <= _BELL_CARRY_ROWS
                bell = self._bell_for(trans_m, dt, 16384 if small else 65536)
                if bell is not None:
                    rb, widths, bidx, bval = bell
                    nblk = (X.rows + rb - 1) // rb
                    split = (not self.exact and nblk >= 2
                             and _dense.query_bytes("rla_sparse_apply_bell_split_bytes", C.rows,
                                                    X.rows, rb, X.cols) <= _BELL_SPLIT_MAX_BYTES)
                    name = "rla_sparse_apply_bell_split_" if split else "rla_sparse_apply_bell_"
                    _lib.call(f"{name}{_lib.sfx(dt)}", _lib.ctx(X.t.device),
                              C.rows, X.rows, rb, _lib.ptr(widths), _lib.ptr(bidx),
                              _lib.ptr(bval), X.cols, _lib.ptr(X.t), X.rs, X.cs, _lib.ptr(C.t),
                              C.rs, C.cs, alpha, beta)
                    return
            ell = self._ell_for(trans_m, dt)
            if ell is not None and ell[0] == "sign":
                _, ptr, es, mag = ell
                _lib.call(f"rla_sparse_apply_ell_sign_{_lib.sfx(dt)}", _lib.ctx(X.t.device),
                          C.rows, X.rows, _lib.ptr(ptr), _lib.ptr(es), mag, X.cols,
                          _lib.ptr(X.t), X.rs, X.cs, _lib.ptr(C.t), C.rs, C.cs, alpha, beta)
                return
            if ell is not None:
                _, ptr, ei, ev = ell
                _lib.call(f"rla_sparse_apply_ell_{_lib.sfx(dt)}", _lib.ctx(X.t.device),
                          C.rows, X.rows, _lib.ptr(ptr), _lib.ptr(ei), _lib.ptr(ev), X.cols,
                          _lib.ptr(X.t), X.rs, X.cs, _lib.ptr(C.t), C.rs, C.cs, alpha, beta)
                return
        val, val_t = self._vals_for(dt)
        _lib.call(f"rla_sparse_apply_{_lib.sfx(dt)}", _lib.ctx(X.t.device), int(trans_m),
                  self.nrows, self.ncols, _lib.ptr(self.colptr), _lib.ptr(self.rowidx),
                  _lib.ptr(val), _lib.ptr(self.rowptr), _lib.ptr(self.colidx), _lib.ptr(val_t),
                  X.cols, _lib.ptr(X.t), X.rs, X.cs, _lib.ptr(C.t), C.rs, C.cs, alpha, beta)

    def dense(self):
        """M as a dense host float64 array (for inspection and tests)."""
        import numpy as np
        out = np.zeros((self.nrows, self.ncols))
        cp = self.colptr.cpu().numpy()
        ri = self.rowidx.cpu().numpy()
        v = self.val.double().cpu().numpy()
        for j in range(self.ncols):
            out[ri[cp[j]:cp[j + 1]], j] = v[cp[j]:cp[j + 1]]
        return out


class _SparseRecipe(CompressorRecipe):
    """Shared mul! plumbing: Left recipes hold S = M (s x m); Right recipes
    hold S = M' with M (s x n) — the reference's `op` / `adjoint(op)`
    (sparse_sign.jl:229-295, count_sketch.jl:113-147)."""

    def _left(self, C, A, alpha, beta):
        # S*A: op(M) = M for Left, M' for Right
        self.op.apply(0 if self.cardinality == Left() else 1, A, C, alpha, beta)

    def _right(self, C, A, alpha, beta):
        # A*S = (S'*A')'
        self.op.apply(1 if self.cardinality == Left() else 0, A.T, C.T, alpha, beta)


class SparseSign(Compressor):
    """SparseSign(cardinality=Left(), compression_dim=2, nnz=min(8, compression_dim),
    type=Float64) — src/Compressors/sparse_sign.jl:65-127.  exact=True keeps
    SparseArrays' summation order (identical bits) for tall operands as well;
    by default X with columns above 64 KiB is applied block-split, within
    rounding of it."""

    def __init__(self, cardinality=None, compression_dim=2, nnz=None, type=torch.float64,
                 exact=False):
        cardinality = Left() if cardinality is None else cardinality
        nnz = min(8, compression_dim) if nnz is None else nnz
        if compression_dim <= 0:
            raise ArgumentError("Field `compression_dim` must be positive.")
        if nnz > compression_dim:
            raise ArgumentError(
                f"Number of non-zero indices, {nnz}, must be less than or equal to "
                f"compression dimension, {compression_dim}.")
        if nnz <= 0:
            raise ArgumentError("Field `nnz` must be positive.")
        object.__setattr__(self, "cardinality", cardinality)
        object.__setattr__(self, "compression_dim", int(compression_dim))
        object.__setattr__(self, "nnz", int(nnz))
        object.__setattr__(self, "type", _dtype(type))
        object.__setattr__(self, "exact", bool(exact))

    def __setattr__(self, name, val):
        if name == "compression_dim":
            if val <= 0:
                raise ArgumentError("Field `compression_dim` must be positive.")
            if self.nnz > val:
                raise ArgumentError(
                    f"New `compression_dim`, {val}, must be greater than or equal to current "
                    f"`nnz`, {self.nnz}.")
        elif name == "nnz":
            if val <= 0:
                raise ArgumentError("Field `nnz` must be positive.")
            if val > self.compression_dim:
                raise ArgumentError(
                    f"New `nnz`, {val}, must be less than or equal to current "
                    f"`compression_dim`, {self.compression_dim}.")
        object.__setattr__(self, name, val)

    def _complete(self, *args):
        A = args[0] if len(args) <= 2 else args[1]
        rec = SparseSignRecipe(self.cardinality, self.compression_dim, A, self.nnz, self.type)
        rec.op.exact = self.exact
        return rec


class SparseSignRecipe(_SparseRecipe):
    """sparse_sign.jl:219-295.  Left: n_rows = s, n_cols = size(A, 1);
    Right: n_rows = size(A, 2), n_cols = s.  `op` is M (nnz sorted rows per
    column, values +-1/sqrt(nnz))."""

    def __init__(self, cardinality, compression_dim, A, nnz, type_):
        if cardinality == Undef():
            raise ArgumentError("SparseSign needs cardinality Left() or Right().")
        a_rows, a_cols = _size2(A)
        self.cardinality = cardinality
        self.nnz = nnz
        if cardinality == Left():
            self.n_rows, self.n_cols = compression_dim, a_rows
            initial = a_rows
        else:
            self.n_rows, self.n_cols = a_cols, compression_dim
            initial = a_cols
        self.type = type_
        sc = float(torch.tensor(1.0 / math.sqrt(nnz), dtype=type_))
        self.scale = [-sc, sc]
        dev = _device_of(A)
        colptr = torch.arange(0, initial * nnz + 1, nnz, dtype=torch.int32, device=dev)
        rowidx = torch.empty(initial * nnz, dtype=torch.int32, device=dev)
        val = torch.empty(initial * nnz, dtype=type_, device=dev)
        self._gen(rowidx, val, compression_dim, initial)
        self.op = _SparseOp(compression_dim, initial, colptr, rowidx, val)
        self.op._uniform = True  # values +-1/sqrt(nnz)

    def _gen(self, rowidx, val, s, initial):
        sd, ep = _draw()
        self.seed, self.epoch = sd, ep
        _lib.call(f"rla_sparse_sign_gen_{_lib.sfx(self.type)}", _lib.ctx(rowidx.device), sd, ep,
                  initial, s, self.nnz, _lib.ptr(rowidx), _lib.ptr(val))

    def _update(self, *args):
        op = self.op
        self._gen(op.rowidx, op.val, op.nrows, op.ncols)
        op.rebuild_csr()


class CountSketch(Compressor):
    """CountSketch(cardinality=Left(), compression_dim=2, type=Float64) —
    src/Compressors/count_sketch.jl:60-81.  exact: as for SparseSign."""

    def __init__(self, cardinality=None, compression_dim=2, type=torch.float64, exact=False):
        cardinality = Left() if cardinality is None else cardinality
        if compression_dim <= 0:
            raise ArgumentError("Field 'compression_dim' must be positive.")
        if cardinality == Undef():
            raise ArgumentError("`cardinality` must be specified as `Left()` or `Right()`."
                                "`Undef()` is not allowed in `CountSketch` structure.")
        self.cardinality = cardinality
        self.compression_dim = int(compression_dim)
        self.type = _dtype(type)
        self.exact = bool(exact)

    def _complete(self, *args):
        A = args[0] if len(args) <= 2 else args[1]
        rec = CountSketchRecipe(self.cardinality, self.compression_dim, A, self.type)
        rec.op.exact = self.exact
        return rec


class CountSketchRecipe(_SparseRecipe):
    """count_sketch.jl:105-147: one +-1 per column of M (s x initial)."""

    def __init__(self, cardinality, compression_dim, A, type_):
        a_rows, a_cols = _size2(A)
        self.cardinality = cardinality
        self.compression_dim = compression_dim
        if cardinality == Left():
            self.n_rows, self.n_cols = compression_dim, a_rows
            initial = a_rows
        else:
            self.n_rows, self.n_cols = a_cols, compression_dim
            initial = a_cols
        self.type = type_
        dev = _device_of(A)
        colptr = torch.arange(0, initial + 1, dtype=torch.int32, device=dev)
        rowidx = torch.empty(initial, dtype=torch.int32, device=dev)
        val = torch.empty(initial, dtype=type_, device=dev)
        self._gen(rowidx, val, compression_dim, initial)
        self.op = _SparseOp(compression_dim, initial, colptr, rowidx, val)
        self.op._uniform = True  # values +-1

    @property
    def mat(self):
        return self.op

    def _gen(self, rowidx, val, s, initial):
        sd, ep = _draw()
        self.seed, self.epoch = sd, ep
        _lib.call(f"rla_count_sketch_gen_{_lib.sfx(self.type)}", _lib.ctx(rowidx.device), sd, ep,
                  initial, s, _lib.ptr(rowidx), _lib.ptr(val))

    def _update(self, *args):
        op = self.op
        self._gen(op.rowidx, op.val, op.nrows, op.ncols)
        op.rebuild_csr()


# ---------------------------------------------------------------------------
# SRHT


def padded_size(n):
    """Int64(2^(ceil(log(2, n)))) exactly as srht.jl:147 evaluates it
    (log(2, n) = log(n)/log(2) in Float64, which overshoots at n = 2^29,
    2^31, 2^39 and then pads to the next power)."""
    return int(2.0 ** math.ceil(math.log(float(n)) / math.log(2.0)))


class SRHT(Compressor):
    """SRHT(cardinality=Left(), compression_dim=2, block_size=10, type=Float64)
    — src/Compressors/srht.jl:60-91.  An extra `exact` keyword (default False)
    keeps Float32 applies in the reference's operation order, bit for bit;
    otherwise Float32 forward applies of 2^14..2^24 rows may take the pruned
    kernel, within Float32 rounding of the reference (1e-5 relative)."""

    def __init__(self, cardinality=None, compression_dim=2, block_size=10, type=torch.float64,
                 exact=False):
        cardinality = Left() if cardinality is None else cardinality
        if compression_dim <= 0:
            raise ArgumentError("Field `compression_dim` must be positive.")
        if block_size <= 0:
            raise ArgumentError("Field `block_size` must be positive.")
        if cardinality == Undef():
            raise ArgumentError("Cardinality must be of type `Left` or `Right`.")
        self.cardinality = cardinality
        self.compression_dim = int(compression_dim)
        self.block_size = int(block_size)
        self.type = _dtype(type)
        self.exact = bool(exact)

    def _complete(self, *args):
        A = args[0] if len(args) <= 2 else args[1]
        rec = SRHTRecipe(self.compression_dim, self.block_size, self.cardinality, A, self.type)
        rec.exact = getattr(self, "exact", False)
        return rec


class SRHTRecipe(CompressorRecipe):
    """srht.jl:125-191.  Left: S = (scale*H*D*pad)[op, :] is s x m; Right: S is
    n x s, the transpose of the same construction on size(A, 2).
    `op` holds 0-based sampled rows, `signs` the BitVector chunks, `scale` the
    Float64 type(sqrt(s)/sqrt(n_pad)).  block_size is kept for parity of the
    recipe fields only: the kernels do not block columns."""

    def __init__(self, compression_dim, block_size, cardinality, A, type_):
        a_rows, a_cols = _size2(A)
        self.cardinality = cardinality
        self.block_size = block_size
        if cardinality == Left():
            self.n_rows, self.n_cols = compression_dim, a_rows
            self.dim = a_rows
        else:
            self.n_rows, self.n_cols = a_cols, compression_dim
            self.dim = a_cols
        self.n_pad = padded_size(self.dim)
        if compression_dim > self.n_pad:
            raise ArgumentError(
                f"Cannot draw a sample of size {compression_dim} without replacement from "
                f"1:{self.n_pad}.")
        self.s = compression_dim
        self.type = type_
        self.scale = float(torch.tensor(math.sqrt(compression_dim) / math.sqrt(self.n_pad),
                                        dtype=type_))
        dev = _device_of(A)
        self.op = torch.empty(compression_dim, dtype=torch.int32, device=dev)
        self.signs = torch.empty((self.n_pad + 63) // 64, dtype=torch.int64, device=dev)
        self._gen()

    def _gen(self):
        sd, ep = _draw()
        self.seed, self.epoch = sd, ep
        _lib.call("rla_srht_gen", _lib.ctx(self.op.device), sd, ep, self.n_pad, self.s,
                  _lib.ptr(self.op), _lib.ptr(self.signs))
        self._plans = {}

    def _update(self, *args):
        self._gen()

    def plan(self, dtype):
        """Device plan of the fused forward kernel for element type `dtype`
        (None when the shape has no fused kernel); rebuilt after updates."""
        if dtype not in self._plans:
            elem = 4 if dtype == torch.float32 else 8
            nb = ctypes.c_size_t(0)
            _lib.call("rla_srht_plan_bytes", elem, self.n_pad, self.s, ctypes.byref(nb))
            pl = None
            if nb.value:
                pl = torch.empty((nb.value + 3) // 4, dtype=torch.int32, device=self.op.device)
                _lib.call("rla_srht_plan", _lib.ctx(self.op.device), elem, self.n_pad, self.s,
                          _lib.ptr(self.op), _lib.ptr(self.signs), _lib.ptr(pl))
            self._plans[dtype] = pl
        return self._plans[dtype]

    def sign_vector(self):
        """The signs as a host bool array of length n_pad (BitVector order)."""
        import numpy as np
        ch = self.signs.cpu().numpy().view(np.uint64)
        bits = ((ch[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(bool)
        return bits.reshape(-1)[: self.n_pad]

    def _apply(self, adjoint, X, Y, alpha, beta):
        dt = X.dtype
        plan = self.plan(dt) if not adjoint else None
        name = f"rla_srht_apply_{_lib.sfx(dt)}"
        if getattr(self, "exact", False) and dt == torch.float32:
            name = "rla_srht_apply_exact_f32"
        _lib.call(name, _lib.ctx(X.t.device), int(adjoint), self.dim,
                  self.n_pad, self.s, X.cols, _lib.ptr(self.op), _lib.ptr(self.signs),
                  self.scale, _lib.ptr(plan), _lib.ptr(X.t), X.rs, X.cs, _lib.ptr(Y.t), Y.rs, Y.cs,
                  alpha, beta)

    def _left(self, C, A, alpha, beta):
        # Left S*A: T*A (srht.jl:220); Right S*A: T'*A (srht.jl:432)
        self._apply(0 if self.cardinality == Left() else 1, A, C, alpha, beta)

    def _right(self, C, A, alpha, beta):
        # Left A*S = (T'*A')' (srht.jl:289); Right A*S = (T*A')' (srht.jl:366)
        self._apply(1 if self.cardinality == Left() else 0, A.T, C.T, alpha, beta)


# ---------------------------------------------------------------------------
# FJLT


class FJLT(Compressor):
    """FJLT(cardinality=Left(), compression_dim=2, block_size=10, sparsity=0.0,
    type=Float64) — src/Compressors/fjlt.jl:65-107."""

    def __init__(self, cardinality=None, compression_dim=2, block_size=10, sparsity=0.0,
                 type=torch.float64):
        cardinality = Left() if cardinality is None else cardinality
        if compression_dim <= 0:
            raise ArgumentError("Field `compression_dim` must be positive.")
        if sparsity < 0 or sparsity > 1:
            raise ArgumentError("Field `sparsity` must be between 0 and 1.")
        if block_size <= 0:
            raise ArgumentError("Field `block_size` must be positive.")
        if cardinality == Undef():
            raise ArgumentError("Cardinality must be of type `Left` or `Right`.")
        self.cardinality = cardinality
        self.compression_dim = int(compression_dim)
        self.block_size = int(block_size)
        self.sparsity = float(sparsity)
        self.type = _dtype(type)

    def _complete(self, *args):
        A = args[0] if len(args) <= 2 else args[1]
        return FJLTRecipe(self.compression_dim, self.block_size, self.cardinality,
                          self.sparsity, A, self.type)


class FJLTRecipe(CompressorRecipe):
    """fjlt.jl:147-251.  Left: S = scale*K*H*D*pad is s x m with K =
    sprandn(s, n_pad, q); Right: S = D*H*K*scale is n x s with K =
    sprandn(n_pad, s, q).  Sparsity 0.0 selects q = 0.25*log(dim)^2/dim
    (fjlt.jl:181, 223), scale = type(1/(sqrt(n_pad)*sqrt(q)*sqrt(s))).
    `op` holds the s x n_pad factor M (M = K for Left, M = K' for Right; both
    are sprandn laws), `signs` the BitVector chunks.  block_size is kept for
    parity of the recipe fields: the kernels chunk columns by workspace size."""

    def __init__(self, compression_dim, block_size, cardinality, sparsity, A, type_):
        a_rows, a_cols = _size2(A)
        self.cardinality = cardinality
        self.block_size = block_size
        if cardinality == Left():
            self.n_rows, self.n_cols = compression_dim, a_rows
            self.dim = a_rows
        else:
            self.n_rows, self.n_cols = a_cols, compression_dim
            self.dim = a_cols
        self.n_pad = padded_size(self.dim)
        self.s = compression_dim
        self.type = type_
        self.sparsity = (0.25 * math.log(self.dim) ** 2 / self.dim) if sparsity == 0.0 \
            else float(sparsity)
        q = self.sparsity
        inv = 1.0 / (math.sqrt(self.n_pad) * math.sqrt(q) * math.sqrt(compression_dim)) \
            if q > 0 else math.inf
        self.scale = float(torch.tensor(inv, dtype=type_))
        self._dev = _device_of(A)
        self.signs = torch.empty((self.n_pad + 63) // 64, dtype=torch.int64, device=self._dev)
        self.op = None
        self._gen()

    def _gen(self):
        sd, ep = _draw()
        self.seed, self.epoch = sd, ep
        dev = self._dev
        ctx = _lib.ctx(dev)
        colptr = torch.empty(self.n_pad + 1, dtype=torch.int32, device=dev)
        nnz = ctypes.c_int64(0)
        _lib.call("rla_fjlt_gen_count", ctx, sd, ep, self.s, self.n_pad, self.sparsity,
                  _lib.ptr(colptr), ctypes.byref(nnz))
        rowidx = torch.empty(max(1, nnz.value), dtype=torch.int32, device=dev)[: nnz.value]
        val = torch.empty(max(1, nnz.value), dtype=self.type, device=dev)[: nnz.value]
        _lib.call(f"rla_fjlt_gen_fill_{_lib.sfx(self.type)}", ctx, sd, ep, self.s, self.n_pad,
                  self.sparsity, _lib.ptr(colptr), _lib.ptr(rowidx), _lib.ptr(val))
        _lib.call("rla_srht_gen", ctx, sd, ep, self.n_pad, self.s, None, _lib.ptr(self.signs))
        self.op = _SparseOp(self.s, self.n_pad, colptr, rowidx, val)

    def _update(self, *args):
        # update_compressor! (fjlt.jl:239-247): a new K with the same sparsity, new signs
        self._gen()

    def sign_vector(self):
        """The signs as a host bool array of length n_pad (BitVector order)."""
        import numpy as np
        ch = self.signs.cpu().numpy().view(np.uint64)
        bits = ((ch[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(bool)
        return bits.reshape(-1)[: self.n_pad]

    def K(self):
        """K as the reference's recipe holds it, a dense host float64 array
        (s x n_pad for Left, n_pad x s for Right)."""
        M = self.op.dense()
        return M if self.cardinality == Left() else M.T

    def _apply(self, adjoint, X, Y, alpha, beta):
        dt = X.dtype
        op = self.op
        if (dt == torch.float32 and not adjoint and not op.exact and X.cols >= _TILE_MIN_COLS
                and self.s <= _FJLT_TILE_MAX_ROWS):
            tile = op._tile_for(0, values=True)
            if tile is not None:
                # forward: the gather reads K's entries, with their values, once
                # for 64 columns of the transform
                words, tiles, _ = tile
                _lib.call("rla_fjlt_apply_tile_f32", _lib.ctx(X.t.device), self.dim,
                          self.n_pad, self.s, X.cols, words, _lib.ptr(tiles),
                          _lib.ptr(self.signs), self.scale, _lib.ptr(X.t), X.rs, X.cs,
                          _lib.ptr(Y.t), Y.rs, Y.cs, alpha, beta)
                return
        bell = op._bell_for(0, dt) if not adjoint else None
        if bell is not None:
            # forward: the gather reads K's rows from the block-interleaved copy
            rb, widths, bidx, bval = bell
            _lib.call(f"rla_fjlt_apply_bell_{_lib.sfx(dt)}", _lib.ctx(X.t.device), self.dim,
                      self.n_pad, self.s, X.cols, rb, _lib.ptr(widths), _lib.ptr(bidx),
                      _lib.ptr(bval), _lib.ptr(self.signs), self.scale, _lib.ptr(X.t), X.rs,
                      X.cs, _lib.ptr(Y.t), Y.rs, Y.cs, alpha, beta)
            return
        ell = op._ell_for(0, dt) if not adjoint else None
        if ell is not None and ell[0] == "ell":
            # forward: the gather reads K's rows from the row-interleaved copy
            _, ptr, ei, ev = ell
            _lib.call(f"rla_fjlt_apply_ell_{_lib.sfx(dt)}", _lib.ctx(X.t.device), self.dim,
                      self.n_pad, self.s, X.cols, _lib.ptr(ptr), _lib.ptr(ei), _lib.ptr(ev),
                      _lib.ptr(self.signs), self.scale, _lib.ptr(X.t), X.rs, X.cs, _lib.ptr(Y.t),
                      Y.rs, Y.cs, alpha, beta)
            return
        val, val_t = op._vals_for(dt)
        _lib.call(f"rla_fjlt_apply_{_lib.sfx(dt)}", _lib.ctx(X.t.device), int(adjoint), self.dim,
                  self.n_pad, self.s, X.cols, _lib.ptr(op.colptr), _lib.ptr(op.rowidx),
                  _lib.ptr(val), _lib.ptr(op.rowptr), _lib.ptr(op.colidx), _lib.ptr(val_t),
                  _lib.ptr(self.signs), self.scale, _lib.ptr(X.t), X.rs, X.cs, _lib.ptr(Y.t),
                  Y.rs, Y.cs, alpha, beta)

    def _left(self, C, A, alpha, beta):
        # Left S*A: T*A (fjlt.jl:257-319); Right S*A: T'*A (fjlt.jl:470-535)
        self._apply(0 if self.cardinality == Left() else 1, A, C, alpha, beta)

    def _right(self, C, A, alpha, beta):
        # Left A*S = (T'*A')' (fjlt.jl:321-393); Right A*S = (T*A')' (fjlt.jl:395-468)
        self._apply(1 if self.cardinality == Left() else 0, A.T, C.T, alpha, beta)


def fwht(x, signs=None, scaling=1.0):
    """fwht!(x, signs; scaling) / fwht!(x; scaling) on each column of a device
    matrix (or a vector) — src/Compressors/helpers/fwht.jl:21-118.  `signs` is
    a bool tensor of length n or a BitVector-chunk int64 tensor."""
    v = as_view(x)
    n = v.rows
    if n < 2 or (n & (n - 1)) != 0:
        raise DimensionMismatch("Size of vector must be power of 2.")
    if v.rs != 1:
        raise ArgumentError("fwht needs unit row stride")
    chunks = None
    if signs is not None:
        if signs.dtype == torch.bool:
            if signs.numel() != n:
                raise DimensionMismatch("Vector `x` and vector `signs` must be same length.")
            chunks = pack_bits(signs).to(x.device)
        else:
            chunks = signs
    _lib.call(f"rla_fwht_{_lib.sfx(x.dtype)}", _lib.ctx(x.device), n, v.cols, _lib.ptr(chunks),
              float(scaling), _lib.ptr(v.t), v.cs)


def pack_bits(bits):
    """bool vector -> Julia BitVector chunks (int64 tensor, bit i of chunk i>>6)."""
    import numpy as np
    b = bits.detach().cpu().numpy().astype(np.uint64)
    n = b.shape[0]
    pad = np.zeros(((n + 63) // 64) * 64, dtype=np.uint64)
    pad[:n] = b
    w = (pad.reshape(-1, 64) << np.arange(64, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
    return torch.from_numpy(w.view(np.int64).copy())


# ---------------------------------------------------------------------------
# Gaussian


GAUSSIAN_PRECISIONS = ("plain", "bf16")


class Gaussian(Compressor):
    """Gaussian(cardinality=Left(), compression_dim=2, type=Float64) —
    src/Compressors/gaussian.jl:49-61.  An extra `precision` keyword picks the
    arithmetic of mul!: "plain" (the data's own type on the FP32 / FP64
    matrix cores, as the reference's BLAS gemm) or "bf16" (operands rounded to
    BF16, FP32 accumulation; Float32 data only)."""

    def __init__(self, cardinality=None, compression_dim=2, type=torch.float64,
                 precision="plain"):
        cardinality = Left() if cardinality is None else cardinality
        if compression_dim <= 0:
            raise ArgumentError("Field 'compression_dim' must be positive.")
        if precision not in GAUSSIAN_PRECISIONS:
            raise ArgumentError(f"precision must be one of {GAUSSIAN_PRECISIONS}")
        self.cardinality = cardinality
        self.compression_dim = int(compression_dim)
        self.type = _dtype(type)
        self.precision = precision

    def _complete(self, *args):
        A = args[0] if len(args) <= 2 else args[1]
        return GaussianRecipe(self.cardinality, self.compression_dim, self.type, A,
                              getattr(self, "precision", "plain"))


class GaussianRecipe(CompressorRecipe):
    """gaussian.jl:87-124: op = scale .* randn(type, n_rows, n_cols) with
    scale = type(1/sqrt(compression_dim)); Left: s x size(A, 1), Right:
    size(A, 2) x s.  `op` is the dense operator S itself (column-major)."""

    def __init__(self, cardinality, compression_dim, type_, A, precision="plain"):
        a_rows, a_cols = _size2(A)
        self.cardinality = cardinality
        self.compression_dim = compression_dim
        self.precision = precision
        self._bf16_op = None  # (key of op, BF16 copy of op, its orientation)
        if cardinality == Right():
            self.n_rows, self.n_cols = a_cols, compression_dim
        else:
            self.n_rows, self.n_cols = compression_dim, a_rows
        self.type = type_
        self.scale = float(torch.tensor(1.0 / math.sqrt(compression_dim), dtype=type_))
        dev = _device_of(A)
        self.op = torch.empty(self.n_cols, self.n_rows, dtype=type_, device=dev).T
        self._gen()

    def _gen(self):
        sd, ep = _draw()
        self.seed, self.epoch = sd, ep
        self._bf16_op = None  # op is rewritten in place below
        _lib.call(f"rla_gaussian_gen_{_lib.sfx(self.type)}", _lib.ctx(self.op.device), sd, ep,
                  self.n_rows, self.n_cols, self.scale, _lib.ptr(self.op), max(1, self.n_rows))

    def _update(self, *args):
        self._gen()

    def _op_as(self, dtype):
        return self.op if dtype == self.op.dtype else self.op.to(dtype)

    def bf16_operand(self, side):
        """BF16 copy of op read K-contiguous by the BF16 GEMM: rows contiguous
        for S*A (side 0), columns for A*S (side 1).  Kept until op changes
        (in place by _gen, which drops it, or by the caller)."""
        from . import _dense
        c = self._bf16_op
        if c is None or c[2] != side or not _dense.key_matches(c[0], self.op):
            S = as_view(self.op)
            ob = _dense.bf16_rows(S if side == 0 else S.T)
            self._bf16_op = c = (_dense.SourceKey(self.op), ob, side)
        return c[1]

    def _apply(self, side, C, A, alpha, beta):
        """side 0: C = alpha*S*A + beta*C; side 1: C = alpha*A*S + beta*C
        (gaussian.jl:145-170), on the library's MFMA kernels."""
        if self.precision == "bf16":
            return self._apply_bf16(side, C, A, alpha, beta)
        S = as_view(self._op_as(A.dtype))
        _lib.call(f"rla_gaussian_apply_{_lib.sfx(A.dtype)}", _lib.ctx(C.t.device), side, S.rows,
                  S.cols, _lib.ptr(S.t), S.rs, S.cs, A.rows, A.cols, _lib.ptr(A.t), A.rs, A.cs,
                  alpha, beta, _lib.ptr(C.t), C.rs, C.cs)

    def _apply_bf16(self, side, C, A, alpha, beta):
        from . import _dense
        if C.dtype != torch.float32 or A.dtype != torch.float32:
            raise ArgumentError("the bf16 Gaussian apply takes Float32 data")
        ob = self.bf16_operand(side)
        if side == 0:   # C (s x n) = S (s x m) * A (m x n)
            Ab = _dense.bf16_rows(A.T)
            _dense.gemm_bf16(C.rows, C.cols, A.rows, ob, Ab, C, alpha, beta)
        else:           # C (m x s) = A (m x n) * S (n x s)
            Ab = _dense.bf16_rows(A)
            _dense.gemm_bf16(C.rows, C.cols, A.cols, Ab, ob, C, alpha, beta)

    def _left(self, C, A, alpha, beta):
        self._apply(0, C, A, alpha, beta)

    def _right(self, C, A, alpha, beta):
        self._apply(1, C, A, alpha, beta)
