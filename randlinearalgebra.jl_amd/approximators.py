"""Approximator surface of RandLinearAlgebra.jl on the MI355X kernels:
RangeFinder and RandSVD (src/Approximators.jl,
src/Approximators/RangeApproximators/{rangefinder,randsvd}.jl).

`rapproximate(approx, A)` = `complete_approximator` + `rapproximate_`
(Approximators.jl:222-226).  The power / orthogonalised iterations, every QR
(FP64 CholeskyQR2) and the RandSVD finish run in librla_mi355x.  An extra
`precision` field selects the product precision: "plain" (A's own type, the
reference's arithmetic) or "bf16" (BF16 MFMA products with FP32 accumulation
on resident BF16 copies of A; QR and the small SVD stay FP64).
"""
import warnings

import torch

from . import _dense, _lib
from ._lib import ArgumentError, DomainError
from .compressors import (GaussianRecipe, Left, Right, SparseSign, _register_mul_hook, as_view,
                          complete_compressor, left_mul_dimcheck, mul, right_mul_dimcheck)

# host precision -> the rla_rangefinder code ("bf16": products on the BF16
# MFMA reading tiled BF16 copies of A)
PRECISIONS = {"plain": 0, "bf16": 2}


class Approximator:
    """Parameters of a low-rank approximation technique (Approximators.jl:10)."""


class ApproximatorRecipe:
    """A completed approximation (Approximators.jl:20); n_rows, n_cols."""

    def size(self, dim=None):
        if dim is None:
            return (self.n_rows, self.n_cols)
        if dim < 1 or dim > 2:
            raise DomainError("`dim` must be 1 or 2.")
        return self.n_rows if dim == 1 else self.n_cols

    @property
    def T(self):
        return ApproximatorAdjoint(self)


class ApproximatorAdjoint:
    """Approximators.jl:122-132."""

    def __init__(self, parent):
        self.parent = parent

    def size(self, dim=None):
        p = self.parent
        if dim is None:
            return (p.n_cols, p.n_rows)
        if dim < 1 or dim > 2:
            raise DomainError("`dim` must be 1 or 2.")
        return p.n_cols if dim == 1 else p.n_rows


def complete_approximator(approx, A):
    if not isinstance(approx, Approximator):
        raise ArgumentError(f"No method `complete_approximator` defined for {type(approx)}.")
    return approx._complete(A)


def rapproximate_(recipe, A):
    """rapproximate!(recipe, A)."""
    recipe._run(A)


def rapproximate(approx, A):
    """rapproximate(approx, A) = complete_approximator + rapproximate!."""
    recipe = complete_approximator(approx, A)
    rapproximate_(recipe, A)
    return recipe


# ---------------------------------------------------------------------------


class _Bf16Copies:
    """The two resident BF16 copies of A used by the bf16 products, in the
    GEMM's tiled layout: of A (for A*X) and of A' (for A'*X).  Rebuilt only
    when A's storage or version changes."""

    def __init__(self):
        self.key = None
        self.rm = self.cm = None

    def get(self, A):
        """A is the caller's matrix (any strides), not a copy of it."""
        if not _dense.key_matches(self.key, A):
            v = as_view(A)
            self.rm = _dense.bf16_tiled(v)      # A, m x n
            self.cm = _dense.bf16_tiled(v.T)    # A', n x m
            self.key = _dense.SourceKey(A)
        return self.rm, self.cm


class _RowCopy:
    """The plain path's resident copy of A in A's type with rows contiguous:
    A'*X then reads it as a column-major A' (a plain product) instead of a
    transposed A.  A matrix that is already row-major is its own copy.
    Rebuilt only when A's storage or version changes."""

    def __init__(self):
        self.key = None
        self.rows = None

    def get(self, A):
        if A.dtype not in (torch.float32, torch.float64):
            return None
        v = as_view(A)
        if v.cs == 1 and v.rs >= v.cols:   # rows already contiguous
            return A
        if not _dense.key_matches(self.key, A):
            self.rows = A.contiguous()
            self.key = _dense.SourceKey(A)
        return self.rows


def _range(recipe, A):
    """Q (m x k, FP64, column-major) for recipe's compressor, power_its and
    orthogonalize settings (power_its.jl:15-80)."""
    A_ = A
    v = as_view(A_)
    if v.rs != 1:  # the vendor GEMMs of the plain path take column-major A
        A_ = A.T.contiguous().T
        v = as_view(A_)
    m, n = v.rows, v.cols
    S = recipe.compressor
    k = S.size(2)
    dev = A.device
    prec = recipe.precision
    Y = torch.empty(k, m, dtype=torch.float64, device=dev).T
    if prec == "bf16":
        rm, cm = recipe._bf16.get(A)
        if isinstance(S, GaussianRecipe):
            om = S.bf16_operand(1)    # columns of the sketch, K-contiguous
            Y32 = torch.empty(k, m, dtype=torch.float32, device=dev).T
            _dense.gemm_bf16_tiled(m, k, n, rm, om, as_view(Y32))
            Y = _dense.convert(Y32, torch.float64)
        else:
            Yt = torch.empty(k, m, dtype=A.dtype, device=dev).T
            mul(Yt, A_, S)
            Y = _dense.convert(Yt, torch.float64)
    else:
        # the row copy makes A'*X a plain product, as A*X is on A itself
        rm, cm = recipe._rows.get(A), None
        Yt = torch.empty(k, m, dtype=A.dtype, device=dev).T
        mul(Yt, A_, S)
        Y = _dense.convert(Yt, torch.float64)
    a_elem = 4 if A.dtype == torch.float32 else 8
    p = PRECISIONS[prec]
    nbytes = _dense.query_bytes("rla_rangefinder_workspace_bytes", p, a_elem, m, n, k)
    work = _dense.workspace(nbytes, dev)
    # tiled BF16 copies carry no leading dimension
    ld_rm = 0 if (rm is None or prec == "bf16") else as_view(rm).rs
    _lib.call("rla_rangefinder", _lib.ctx(dev), p, int(bool(recipe.orthogonalize)), m, n, k,
              recipe.power_its, a_elem, _lib.ptr(A_), v.cs, _lib.ptr(rm), ld_rm, _lib.ptr(cm),
              0, _lib.ptr(Y), m, _lib.ptr(work),
              work.numel() * 8)
    return Y, A_, rm, cm


# ---------------------------------------------------------------------------
# RangeFinder


class RangeFinder(Approximator):
    """RangeFinder(compressor=SparseSign(cardinality=Right()), power_its=0,
    orthogonalize=False[, precision="plain"]) — rangefinder.jl:71-89."""

    def __init__(self, compressor=None, power_its=0, orthogonalize=False, precision="plain"):
        if power_its < 0:
            raise ArgumentError("Field `power_its` must be non-negative.")
        if precision not in PRECISIONS:
            raise ArgumentError(f"precision must be one of {sorted(PRECISIONS)}")
        self.compressor = SparseSign(cardinality=Right()) if compressor is None else compressor
        self.power_its = int(power_its)
        self.orthogonalize = bool(orthogonalize)
        self.precision = precision

    def _complete(self, A):
        if self.compressor.cardinality == Left():
            warnings.warn("Compressor with cardinality `Left` being applied from `Right`.")
        comp = complete_compressor(self.compressor, A)
        return RangeFinderRecipe(A.shape[0], comp.size(2), comp, self.power_its,
                                 self.orthogonalize, self.precision)


class RangeFinderRecipe(ApproximatorRecipe):
    """rangefinder.jl:104-132: n_rows = size(A, 1), n_cols = size(S, 2);
    `range` holds Q after rapproximate!."""

    def __init__(self, n_rows, n_cols, compressor, power_its, orthogonalize, precision):
        self.n_rows, self.n_cols = n_rows, n_cols
        self.compressor = compressor
        self.power_its = power_its
        self.orthogonalize = orthogonalize
        self.precision = precision
        self.range = None
        self._bf16 = _Bf16Copies()
        self._rows = _RowCopy()

    def _run(self, A):
        Q, _, _, _ = _range(self, A)
        self.range = Q if A.dtype == torch.float64 else _dense.convert(Q, A.dtype)

    def _left(self, C, A, alpha, beta):
        _dense.gemm(C, as_view(self.range), A, alpha, beta)

    def _right(self, C, A, alpha, beta):
        _dense.gemm(C, A, as_view(self.range), alpha, beta)

    def _left_adj(self, C, A, alpha, beta):
        _dense.gemm(C, as_view(self.range).T, A, alpha, beta)

    def _right_adj(self, C, A, alpha, beta):
        _dense.gemm(C, A, as_view(self.range).T, alpha, beta)


# ---------------------------------------------------------------------------
# RandSVD


class RandSVD(Approximator):
    """RandSVD(compressor=SparseSign(cardinality=Right()), power_its=0,
    orthogonalize=False, block_size=0[, precision="plain"]) —
    randsvd.jl:87-111."""

    def __init__(self, compressor=None, power_its=0, orthogonalize=False, block_size=0,
                 precision="plain"):
        if power_its < 0:
            raise ArgumentError("Field `power_its` must be non-negative.")
        if block_size < 0:
            raise ArgumentError("Field `block_size` must be non-negative.")
        if precision not in PRECISIONS:
            raise ArgumentError(f"precision must be one of {sorted(PRECISIONS)}")
        self.compressor = SparseSign(cardinality=Right()) if compressor is None else compressor
        self.power_its = int(power_its)
        self.orthogonalize = bool(orthogonalize)
        self.block_size = int(block_size)
        self.precision = precision

    def _complete(self, A):
        if self.compressor.cardinality == Left():
            warnings.warn("Compressor with cardinality `Left` being applied from `Right`.")
        comp = complete_compressor(self.compressor, A)
        return RandSVDRecipe(A.shape[0], A.shape[1], comp, self.power_its, self.orthogonalize,
                             self.block_size or A.shape[1], self.precision)


class RandSVDRecipe(ApproximatorRecipe):
    """randsvd.jl:136-189: U (m x k), S (k), V (n x k) in Float64 (the
    reference allocates Q'A as Matrix{Float64})."""

    def __init__(self, n_rows, n_cols, compressor, power_its, orthogonalize, block_size,
                 precision):
        self.n_rows, self.n_cols = n_rows, n_cols
        self.compressor = compressor
        self.power_its = power_its
        self.orthogonalize = orthogonalize
        self.block_size = block_size
        self.precision = precision
        self.U = self.S = self.V = None
        self._bf16 = _Bf16Copies()
        self._rows = _RowCopy()

    def _run(self, A):
        Q, A_, rm, cm = _range(self, A)
        m, n = A_.shape
        k = Q.shape[1]
        if k > n:
            raise ArgumentError("RandSVD needs compression_dim <= size(A, 2)")
        dev = A.device
        a_elem = 4 if A.dtype == torch.float32 else 8
        p = PRECISIONS[self.precision]
        U = torch.empty(k, m, dtype=torch.float64, device=dev).T
        V = torch.empty(k, n, dtype=torch.float64, device=dev).T
        S = torch.empty(k, dtype=torch.float64, device=dev)
        nbytes = _dense.query_bytes("rla_randsvd_workspace_bytes", p, a_elem, m, n, k)
        work = _dense.workspace(nbytes, dev)
        # A'Q reads the BF16 columns copy, or the plain path's row copy
        at = cm if self.precision == "bf16" else rm
        ld_at = 0 if (at is None or self.precision == "bf16") else as_view(at).rs
        _lib.call("rla_randsvd_finish", _lib.ctx(dev), p, m, n, k, a_elem, _lib.ptr(A_),
                  as_view(A_).cs, _lib.ptr(at), ld_at,
                  _lib.ptr(Q), m, _lib.ptr(U), m, _lib.ptr(S), _lib.ptr(V), n, _lib.ptr(work),
                  work.numel() * 8)
        self.U, self.S, self.V = U, S, V
        US = torch.empty(k, m, dtype=torch.float64, device=dev).T
        _lib.call("rla_scale_cols_f64", _lib.ctx(dev), m, k, _lib.ptr(U), m, _lib.ptr(S),
                  _lib.ptr(US), m)
        self._US = US

    # C = alpha*U*S*V'*A + beta*C and C = alpha*A*U*S*V' + beta*C
    # (randsvd.jl:198-293; the S scaling is folded into U once)
    def _left(self, C, A, alpha, beta):
        buf = torch.empty(A.cols, self.V.shape[1], dtype=C.dtype, device=C.t.device).T
        _dense.gemm(as_view(buf), as_view(self.V.to(C.dtype)).T, A, alpha, 0.0)
        _dense.gemm(C, as_view(self._US.to(C.dtype)), as_view(buf), 1.0, beta)

    def _right(self, C, A, alpha, beta):
        buf = torch.empty(self.V.shape[1], A.rows, dtype=C.dtype, device=C.t.device).T
        _dense.gemm(as_view(buf), A, as_view(self._US.to(C.dtype)), alpha, 0.0)
        _dense.gemm(C, as_view(buf), as_view(self.V.to(C.dtype)).T, 1.0, beta)

    def _left_adj(self, C, A, alpha, beta):
        buf = torch.empty(A.cols, self.V.shape[1], dtype=C.dtype, device=C.t.device).T
        _dense.gemm(as_view(buf), as_view(self._US.to(C.dtype)).T, A, alpha, 0.0)
        _dense.gemm(C, as_view(self.V.to(C.dtype)), as_view(buf), 1.0, beta)

    def _right_adj(self, C, A, alpha, beta):
        buf = torch.empty(self.V.shape[1], A.rows, dtype=C.dtype, device=C.t.device).T
        _dense.gemm(as_view(buf), A, as_view(self.V.to(C.dtype)), alpha, 0.0)
        _dense.gemm(C, as_view(buf), as_view(self._US.to(C.dtype)).T, 1.0, beta)


def approx_mul(C, X, Y, alpha=1.0, beta=0.0):
    """mul! with an approximator recipe or its adjoint (Approximators.jl:334-453)."""
    Cv = as_view(C)
    if isinstance(X, ApproximatorAdjoint):
        left_mul_dimcheck(C, X, Y)
        X.parent._left_adj(Cv, as_view(Y), alpha, beta)
    elif isinstance(Y, ApproximatorAdjoint):
        right_mul_dimcheck(C, X, Y)
        Y.parent._right_adj(Cv, as_view(X), alpha, beta)
    elif isinstance(X, ApproximatorRecipe):
        left_mul_dimcheck(C, X, Y)
        X._left(Cv, as_view(Y), alpha, beta)
    else:
        right_mul_dimcheck(C, X, Y)
        Y._right(Cv, as_view(X), alpha, beta)


def _mul_hook(C, X, Y, alpha, beta):
    if isinstance(X, (ApproximatorRecipe, ApproximatorAdjoint)) or \
            isinstance(Y, (ApproximatorRecipe, ApproximatorAdjoint)):
        approx_mul(C, X, Y, alpha, beta)
        return True
    return False


_register_mul_hook(_mul_hook)
