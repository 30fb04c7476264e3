"""Solver surface of RandLinearAlgebra.jl on the MI355X kernels: the Sampling
compressor with its Uniform / L2Norm distributions, block Kaczmarz with the LQ
sub-solver, FullResidual and BasicLogger.

Mirrors src/Solvers.jl (`complete_solver`, `rsolve!` returning `(x, recipe)`
for a solver and nothing for a recipe), src/Solvers/kaczmarz.jl,
src/Compressors/sampling.jl, src/Compressors/Distributions/uniform.jl and
strohmer_vershynin.jl, src/Solvers/Loggers/basic_logger.jl and
src/Solvers/ErrorMethods/full_residual.jl.

A Kaczmarz recipe whose compressor is a Left Sampling recipe of at most 64
rows on Float32 data runs the whole solve in librla_mi355x.so
(rla_kaczmarz_solve_f32: row sets drawn on the device, fused gather/update
kernels, FP64 residual); other compressors run the reference loop here with
each step on the device (update_compressor, mul, rla_kaczmarz_block_update_f32,
rla_residual_norm_f32).
"""
import ctypes

import numpy as np
import torch

from . import _dense, _lib
from ._lib import ArgumentError, DimensionMismatch
from .compressors import (Compressor, CompressorRecipe, Left, Right, SparseSign, Undef,
                          _SparseOp, _SparseRecipe, _draw, _size2, as_view, complete_compressor,
                          mul, update_compressor)

# ---------------------------------------------------------------------------
# distributions (src/Compressors/Distributions/)


class Distribution:
    cardinality = Undef()
    replace = False


class Uniform(Distribution):
    """Uniform(cardinality=Undef(), replace=False) — uniform.jl:7-22."""

    def __init__(self, cardinality=None, replace=False):
        self.cardinality = Undef() if cardinality is None else cardinality
        self.replace = bool(replace)


class L2Norm(Distribution):
    """L2Norm(cardinality=Undef(), replace=False) — strohmer_vershynin.jl:20-37:
    rows (Left) or columns (Right) drawn with probability proportional to
    their squared Euclidean norm."""

    def __init__(self, cardinality=None, replace=False):
        self.cardinality = Undef() if cardinality is None else cardinality
        self.replace = bool(replace)


class DistributionRecipe:
    """Weights over the state space 0:n-1 (device FP64, None = uniform)."""

    def __init__(self, dist, A):
        self.cardinality = dist.cardinality
        self.replace = dist.replace
        self.kind = type(dist)
        self._key = None
        self.weights = None
        self.n = 0
        self._compute(A)

    def _compute(self, A):
        name = "complete_distribution" if self._key is None else "update_distribution!"
        if self.cardinality == Undef():
            raise ArgumentError(
                f"`{self.kind.__name__}` cardinality must be specified as `Left()` or "
                f"`Right()`.`Undef()` is not allowed in `{name}`.")
        a_rows, a_cols = _size2(A)
        self.n = a_rows if self.cardinality == Left() else a_cols
        if self.kind is Uniform:
            self.weights = None
            return
        v = as_view(A)
        if _dense.key_matches(self._key, v.t):
            return  # A unchanged since the weights were computed
        w = torch.empty(self.n, dtype=torch.float64, device=v.t.device)
        rs, cs = (v.rs, v.cs) if self.cardinality == Left() else (v.cs, v.rs)
        cols = v.cols if self.cardinality == Left() else v.rows
        _lib.call(f"rla_row_norms2_{_lib.sfx(v.dtype)}", _lib.ctx(v.t.device), self.n, cols,
                  _lib.ptr(v.t), rs, cs, _lib.ptr(w))
        self.weights = w
        self._key = _dense.SourceKey(v.t)

    def positive(self):
        if self.weights is None:
            return self.n
        return int((self.weights > 0).sum())


class Agmon(Distribution):
    """Agmon(cardinality=Undef(), replace=False, beta=1) — maximal-residual
    selection among beta random candidates, agmon.jl:1-70."""

    def __init__(self, cardinality=None, replace=False, beta=1):
        if beta < 1:
            raise ArgumentError(f"`Agmon` beta must be >= 1, got beta={beta}")
        self.cardinality = Undef() if cardinality is None else cardinality
        self.replace = bool(replace)
        self.beta = int(beta)


class AgmonRecipe:
    """agmon.jl:72-133: cardinality, replace, beta, the size of the state
    space (indices 0..d-1), the candidate buffer (device int32, beta entries,
    allocated on first use), references to A, b, x and the optional stored
    residual r (A x - b, maintained by the caller) or None."""

    def __init__(self, dist, x, A, b):
        self.cardinality = dist.cardinality
        self.replace = dist.replace
        self.beta = dist.beta
        self.sample_buffer = None
        self.r = None
        self._set(x, A, b, "complete_distribution")

    def _set(self, x, A, b, name):
        """The checks of complete_distribution / update_distribution!
        (agmon.jl:156-205, 235-301), with their messages."""
        n_rows, n_cols = _size2(A)
        if self.cardinality == Left():
            d, x_len, b_len = n_rows, n_cols, n_rows
        elif self.cardinality == Right():
            d, x_len, b_len = n_cols, n_cols, n_rows
        else:
            raise ArgumentError("`Agmon` cardinality must be specified as `Left()` or `Right()`."
                                f"`Undef()` is not allowed in `{name}`.")
        if b.shape[0] != b_len:
            raise DimensionMismatch(
                f"Vector b has length {b.shape[0]}, expected {b_len} to match rows of A")
        if x.shape[0] != x_len:
            raise DimensionMismatch(
                f"Vector x has length {x.shape[0]}, expected {x_len} to match columns of A")
        if self.beta > d:
            dim_name = "rows" if self.cardinality == Left() else "columns"
            raise ArgumentError(
                f"`Agmon` beta must be <= number of {dim_name} ({d}), got beta={self.beta}")
        self.state_space = d
        self.A, self.b, self.x = A, b, x


def complete_distribution(dist, *args):
    """complete_distribution(dist, A) for Uniform / L2Norm; (dist, x, A, b)
    for Agmon (agmon.jl:156-205)."""
    if isinstance(dist, Agmon):
        if len(args) != 3:
            raise ArgumentError("`Agmon` is completed with complete_distribution(distribution, "
                                "x, A, b).")
        return AgmonRecipe(dist, *args)
    A = args[0] if len(args) == 1 else args[1]
    return DistributionRecipe(dist, A)


def update_distribution(recipe, *args, r=None):
    """update_distribution!(recipe, A) or (recipe, x, A, b): L2Norm recomputes
    its weights from A (strohmer_vershynin.jl:115-134); they are cached while
    A is unchanged.  Agmon (agmon.jl:235-301) stores x, A, b and the optional
    residual r = A x - b kept by the caller (None: computed when sampling)."""
    if isinstance(recipe, AgmonRecipe):
        x, A, b = args
        recipe._set(x, A, b, "update_distribution!")
        recipe.r = r
        return
    A = args[0] if len(args) == 1 else args[1]
    recipe._compute(A)


def sample_distribution(indices, recipe):
    """sample_distribution!(indices, ::AgmonRecipe) — agmon.jl:354-416:
    indices (device int32, 0-based) receive the len(indices) candidates with
    the largest |residual|, descending, ties to the smaller index.  Candidates:
    one uniform draw when beta = 1, all indices when beta >= d, else beta
    uniform draws (with or without replacement; at most 1024).  Residuals:
    rows |A x - b|, columns |A'(A x - b)|, with the stored r when given."""
    if not isinstance(recipe, AgmonRecipe):
        raise ArgumentError(f"No method `sample_distribution!` defined for {type(recipe)}.")
    k = indices.shape[0]
    if k > recipe.beta:
        raise ArgumentError(f"`Agmon` cannot output {k} indices when beta={recipe.beta}.")
    d = recipe.state_space
    dev = indices.device
    ctx = _lib.ctx(dev)
    sd, ep = _draw()
    if recipe.beta == 1:
        _lib.call("rla_wsample", ctx, sd, ep, d, None, 1, 0, 1, _lib.ptr(indices))
        return
    cand, n_cand = None, d
    if recipe.beta < d:
        if recipe.beta > 1024:
            raise ArgumentError("the device sampler draws at most 1024 `Agmon` candidates "
                                "(beta < d)")
        if recipe.sample_buffer is None or recipe.sample_buffer.shape[0] != recipe.beta:
            recipe.sample_buffer = torch.empty(recipe.beta, dtype=torch.int32, device=dev)
        cand, n_cand = recipe.sample_buffer, recipe.beta
        _lib.call("rla_wsample", ctx, sd, ep, d, None, recipe.beta, 1 if recipe.replace else 0,
                  1, _lib.ptr(cand))
    Av = as_view(recipe.A)
    if recipe.r is not None:
        res = recipe.r
    else:
        res = recipe.b.clone()  # r = A x - b
        _dense.gemm(_col(res), Av, _col(recipe.x), 1.0, -1.0)
    if recipe.cardinality == Left():
        v = res
    else:
        v = torch.empty(Av.cols, dtype=res.dtype, device=dev)
        _dense.gemm(_col(v), Av.T, _col(res), 1.0, 0.0)
    _lib.call(f"rla_agmon_select_{_lib.sfx(v.dtype)}", ctx, d, _lib.ptr(cand), n_cand,
              _lib.ptr(v.contiguous()), k, _lib.ptr(indices))


# ---------------------------------------------------------------------------
# Sampling compressor (src/Compressors/sampling.jl)


class Sampling(Compressor):
    """Sampling(cardinality=Left(), compression_dim=2, distribution=Uniform()) —
    sampling.jl:39-74."""

    def __init__(self, cardinality=None, compression_dim=2, distribution=None):
        cardinality = Left() if cardinality is None else cardinality
        if compression_dim <= 0:
            raise ArgumentError("Field `compression_dim` must be positive.")
        if cardinality == Undef():
            raise ArgumentError("`cardinality` must be specified as `Left()` or `Right()`."
                                "`Undef()` is not allowed in `Sampling` structure.")
        self.cardinality = cardinality
        self.compression_dim = int(compression_dim)
        self.distribution = Uniform() if distribution is None else distribution

    def _complete(self, *args):
        A = args[0] if len(args) <= 2 else args[1]
        return SamplingRecipe(self, A)


class SamplingRecipe(_SparseRecipe):
    """sampling.jl:76-123.  `idx` is the sorted index set (0-based, device
    int32).  As an operator, Left: S (s x m) selects rows, S*A = A[idx, :];
    Right: S (n x s), A*S = A[:, idx]; other products follow from the
    selection matrix M (s x n_items, M[r, idx[r]] = 1)."""

    def __init__(self, sampling, A):
        a_rows, a_cols = _size2(A)
        self.cardinality = sampling.cardinality
        self.compression_dim = sampling.compression_dim
        if self.cardinality == Left():
            self.n_rows, self.n_cols = sampling.compression_dim, a_rows
        else:
            self.n_rows, self.n_cols = a_cols, sampling.compression_dim
        sampling.distribution.cardinality = sampling.cardinality
        self.distribution_recipe = complete_distribution(sampling.distribution, A)
        self.dtype = as_view(A).dtype
        self.device = as_view(A).t.device
        self.idx = torch.empty(self.compression_dim, dtype=torch.int32, device=self.device)
        self._sample()

    def _sample(self):
        d = self.distribution_recipe
        k = self.compression_dim
        if not d.replace and d.positive() < k:
            raise ArgumentError(
                f"cannot draw {k} distinct indices from {d.positive()} with positive weight")
        sd, ep = _draw()
        self.seed, self.epoch = sd, ep
        _lib.call("rla_wsample", _lib.ctx(self.device), sd, ep, d.n,
                  _lib.ptr(d.weights) if d.weights is not None else None, k,
                  1 if d.replace else 0, 1, _lib.ptr(self.idx))
        self._build_op()

    def _build_op(self):
        # M (k x n_items) with M[r, idx[r]] = 1: CSR of M is (0:k, idx); its CSC
        # is the transpose of M' = CSC (0:k, idx)
        k, n = self.compression_dim, self.distribution_recipe.n
        dev, dt = self.device, self.dtype
        rowptr = torch.arange(0, k + 1, dtype=torch.int32, device=dev)
        ones = torch.ones(k, dtype=dt, device=dev)
        colptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
        rowidx = torch.empty(k, dtype=torch.int32, device=dev)
        val = torch.empty(k, dtype=dt, device=dev)
        _lib.call(f"rla_sparse_transpose_{_lib.sfx(dt)}", _lib.ctx(dev), n, k, k,
                  _lib.ptr(rowptr), _lib.ptr(self.idx), _lib.ptr(ones), _lib.ptr(colptr),
                  _lib.ptr(rowidx), _lib.ptr(val))
        op = _SparseOp.__new__(_SparseOp)
        op.nrows, op.ncols, op.nnz = k, n, k
        op.colptr, op.rowidx, op.val = colptr, rowidx, val
        op.rowptr, op.colidx, op.val_t = rowptr, self.idx, ones
        op._vals = {}
        op._ell = {}
        self.op = op

    def _update(self, *args):
        # update_compressor!(S, x, A, b): update_distribution! then a new draw
        if args:
            update_distribution(self.distribution_recipe, *args)
        self._sample()


# ---------------------------------------------------------------------------
# loggers, errors, sub-solvers


def threshold_stop(log):
    """basic_logger.jl:130-134."""
    return log.error < log.threshold_info


class BasicLogger:
    """BasicLogger(max_it=0, collection_rate=1, threshold=0.0,
    stopping_criterion=threshold_stop) — basic_logger.jl:15-45."""

    def __init__(self, max_it=0, collection_rate=1, threshold=0.0,
                 stopping_criterion=threshold_stop):
        if max_it < 0:
            raise ArgumentError("Field `max_it` must be positive or 0.")
        if collection_rate < 1:
            raise ArgumentError("Field `colection_rate` must be positive.")
        if collection_rate > max_it and max_it > 0:
            raise ArgumentError("Field `colection_rate` must be smaller than `max_it`.")
        self.max_it = int(max_it)
        self.collection_rate = int(collection_rate)
        self.threshold_info = threshold
        self.stopping_criterion = stopping_criterion


class BasicLoggerRecipe:
    """basic_logger.jl:47-125."""

    def __init__(self, log):
        self.max_it = log.max_it
        self.error = 0.0
        self.threshold_info = log.threshold_info
        self.iteration = 1
        self.record_location = 1
        self.collection_rate = log.collection_rate
        self.converged = False
        self.stopping_criterion = log.stopping_criterion
        self.hist = np.zeros(int(np.ceil(log.max_it / log.collection_rate)) + 1)

    def update(self, error, iteration):
        """update_logger! (basic_logger.jl:93-109)."""
        self.iteration = iteration
        self.error = float(error)
        self.converged = (self.stopping_criterion(self) if iteration < self.max_it else True)
        if self.converged:
            self.hist[self.record_location - 1] = error
        elif iteration % self.collection_rate == 0:
            self.hist[self.record_location - 1] = error
            self.record_location += 1

    def reset(self):
        """reset_logger! (basic_logger.jl:111-118)."""
        self.error = 0.0
        self.iteration = 1
        self.record_location = 1
        self.converged = False
        self.hist[:] = 0.0


def complete_logger(log):
    return BasicLoggerRecipe(log)


class FullResidual:
    """FullResidual() — full_residual.jl:1-20: error = ||b - A x||_2, here
    accumulated in FP64.  `every` (an extension; 1 is the reference's
    behaviour) evaluates it only at iterations 1, 1+every, ... and at max_it;
    the logger sees those evaluations only."""

    def __init__(self, every=1):
        if every < 1:
            raise ArgumentError("Field `every` must be positive.")
        self.every = int(every)


class FullResidualRecipe:
    def __init__(self, err, b):
        self.every = err.every
        self.residual = None  # the norm is reduced on the device


class LSGradient:
    """LSGradient() — ls_gradient.jl:1-11: the error is the norm of the
    least-squares gradient -A'(b - A x)."""


class LSGradientRecipe:
    """ls_gradient.jl:13-36: holds the gradient (Float32, length n)."""

    def __init__(self, gradient):
        self.gradient = gradient


def _norm2(v):
    """||v||_2 of a contiguous Float32 / Float64 device vector, accumulated in
    FP64 by the residual kernel (||v - 0 x|| with a matrix of no columns)."""
    out = ctypes.c_double(0.0)
    _lib.call(f"rla_residual_norm_{_lib.sfx(v.dtype)}", _lib.ctx(v.device), v.shape[0], 0,
              _lib.ptr(v), 1, 1, _lib.ptr(v), _lib.ptr(v), ctypes.byref(out))
    return out.value


def compute_error(err, solver, A, b):
    """compute_error: FullResidual (full_residual.jl:32-41) in FP64;
    LSGradient (ls_gradient.jl:39-51): gradient = -A' r with the solver's
    residual r (Float32 product), norm accumulated in FP64."""
    if isinstance(err, LSGradientRecipe):
        _dense.gemm(_col(err.gradient), as_view(A).T, _col(solver.residual_vec), -1.0, 0.0)
        return _norm2(err.gradient)
    Av = as_view(A)
    out = ctypes.c_double(0.0)
    _lib.call(f"rla_residual_norm_{_lib.sfx(Av.dtype)}", _lib.ctx(Av.t.device), Av.rows,
              Av.cols, _lib.ptr(Av.t), Av.rs, Av.cs, _lib.ptr(solver.solution_vec), _lib.ptr(b),
              ctypes.byref(out))
    return out.value


class LQSolver:
    """LQSolver() — lq.jl:1-12: the minimum-norm solution of the block system."""


class LQSolverRecipe:
    def __init__(self, A=None):
        self.A = A


# ---------------------------------------------------------------------------
# Kaczmarz (src/Solvers/kaczmarz.jl)


class Solver:
    pass


class Kaczmarz(Solver):
    """Kaczmarz(compressor=SparseSign(cardinality=Left()), log=BasicLogger(),
    error=FullResidual(), sub_solver=LQSolver(), alpha=1.0) — kaczmarz.jl:68-101."""

    def __init__(self, compressor=None, log=None, error=None, sub_solver=None, alpha=1.0):
        self.compressor = SparseSign(cardinality=Left()) if compressor is None else compressor
        if self.compressor.cardinality != Left():
            import warnings
            warnings.warn("Compressor has cardinality `Right` but kaczmarzcompresses  from the"
                          "  `Left`.")
        self.log = BasicLogger() if log is None else log
        self.error = FullResidual() if error is None else error
        self.sub_solver = LQSolver() if sub_solver is None else sub_solver
        self.alpha = float(alpha)


class KaczmarzRecipe:
    """kaczmarz.jl:103-160.  Holds the compressor/logger/error/sub-solver
    recipes, alpha, the compressed block buffers and, for the device loop,
    the row-major copy of A."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def _check_float(*ts):
    """Float32 or Float64 data, one type for A, b and x."""
    dt = ts[0].dtype
    for t in ts:
        if t.dtype not in (torch.float32, torch.float64) or t.dtype != dt:
            raise ArgumentError("the GPU Kaczmarz solver takes Float32 or Float64 data of one "
                                f"type (got {[u.dtype for u in ts]})")


def complete_solver(ingredients, x, A, b):
    """complete_solver(::Kaczmarz, x, A, b) — kaczmarz.jl:162-246; for an IHS
    solver, ihs.jl:122-201; for a ColumnProjection solver,
    col_projection.jl:175-247."""
    if isinstance(ingredients, IHS):
        return _complete_ihs(ingredients, x, A, b)
    if isinstance(ingredients, ColumnProjection):
        return _complete_colproj(ingredients, x, A, b)
    if not isinstance(ingredients, Kaczmarz):
        raise ArgumentError(f"No method `complete_solver` defined for {type(ingredients)}.")
    compressor = complete_compressor(ingredients.compressor, x, A, b)
    logger = complete_logger(ingredients.log)
    if not isinstance(ingredients.error, FullResidual):
        raise ArgumentError(
            f"ErrorRecipe {type(ingredients.error).__name__} does not contain the field "
            f"'residual' and is not valid for a kaczmarz solver.")
    error = FullResidualRecipe(ingredients.error, b)
    if not isinstance(ingredients.sub_solver, LQSolver):
        raise ArgumentError("the GPU Kaczmarz solver projects with LQSolver() only")
    _check_float(as_view(A).t, b, x)
    sample_size = compressor.n_rows
    n = _size2(A)[1]
    dev = as_view(A).t.device
    dt = as_view(A).dtype
    # compressed block buffers (kaczmarz.jl:207-224), column-major
    compressed_mat = torch.zeros(n, sample_size, dtype=dt, device=dev).T
    compressed_vec = torch.zeros(sample_size, dtype=dt, device=dev)
    return KaczmarzRecipe(compressor=compressor, log=logger, error=error,
                          sub_solver=LQSolverRecipe(compressed_mat), alpha=ingredients.alpha,
                          compressed_mat=compressed_mat, compressed_vec=compressed_vec,
                          solution_vec=x, A_rm=None, _rm_key=None)


def _row_major(recipe, A):
    """A with each row contiguous (the block gathers read rows): A itself, or
    a copy kept while A is unchanged."""
    v = as_view(A)
    if v.cs == 1 and v.rs >= v.cols:
        return v.t, v.rs
    if not _dense.key_matches(recipe._rm_key, v.t):
        recipe.A_rm = torch.empty(v.rows, v.cols, dtype=v.dtype, device=v.t.device)
        recipe.A_rm.copy_(A)
        recipe._rm_key = _dense.SourceKey(v.t)
    return recipe.A_rm, v.cols


def _fast_path(recipe):
    S = recipe.compressor
    return (isinstance(S, SamplingRecipe) and S.cardinality == Left()
            and S.compression_dim <= 64 and recipe.compressed_mat.dtype == torch.float32
            and isinstance(recipe.log, BasicLoggerRecipe)
            and recipe.log.stopping_criterion is threshold_stop
            and np.isscalar(recipe.log.threshold_info))


def rsolve(solver, x, A, b):
    """rsolve!(solver, x, A, b) -> (x, recipe) for a Kaczmarz, IHS or
    ColumnProjection solver (Solvers.jl:145-149), or rsolve!(recipe, x, A, b)
    -> None (kaczmarz.jl:313-368, ihs.jl:214-254, col_projection.jl:322-372):
    x is updated in place."""
    if isinstance(solver, Solver):
        recipe = complete_solver(solver, x, A, b)
        _rsolve(recipe, x, A, b)
        return x, recipe
    _rsolve(solver, x, A, b)
    return None


rsolve_ = rsolve


def _rsolve(recipe, x, A, b):
    if isinstance(recipe, IHSRecipe):
        return _rsolve_ihs(recipe, x, A, b)
    if isinstance(recipe, ColumnProjectionRecipe):
        return _rsolve_colproj(recipe, x, A, b)
    log = recipe.log
    log.reset()
    a_rows, a_cols = _size2(A)
    if x.shape[0] != a_cols:
        raise DimensionMismatch(
            f"Dimension of `x`, {x.shape[0]} is different from number of columns in `A`,"
            f"{a_cols}.")
    if b.shape[0] != a_rows:
        raise DimensionMismatch(
            f"Dimension of `b`, {b.shape[0]} is different from number of rows in `A`,"
            f"{a_rows}.")
    _check_float(as_view(A).t, b, x)
    recipe.solution_vec = x
    if log.max_it == 0:
        return
    if _fast_path(recipe):
        return _rsolve_device(recipe, x, A, b)
    return _rsolve_loop(recipe, x, A, b)


def _rsolve_device(recipe, x, A, b):
    S, log = recipe.compressor, recipe.log
    Av = as_view(A)
    if Av.rs != 1:
        raise ArgumentError("the device Kaczmarz loop reads A column-major (unit row stride)")
    A_rm, ld_rm = _row_major(recipe, A)
    d = S.distribution_recipe
    update_distribution(d, x, A, b)
    if not d.replace and d.positive() < S.compression_dim:
        raise ArgumentError(f"cannot draw {S.compression_dim} distinct indices from "
                            f"{d.positive()} with positive weight")
    m, n = Av.rows, Av.cols
    s, every = S.compression_dim, recipe.error.every
    nb = _dense.query_bytes("rla_kaczmarz_workspace_bytes", m, n, s, every)
    work = torch.empty(max(1, (nb + 7) // 8), dtype=torch.float64, device=Av.t.device)
    cap = log.max_it // every + 2
    hist = np.zeros(cap)
    n_hist, iters, conv = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int(0)
    # one epoch per iteration's row set; a recipe may carry a fixed (seed,
    # epoch) so its solve does not depend on other draws
    sd, ep = recipe.rng if getattr(recipe, "rng", None) else _draw(log.max_it)
    _lib.call("rla_kaczmarz_solve_f32", _lib.ctx(Av.t.device), m, n, _lib.ptr(Av.t), Av.cs,
              _lib.ptr(A_rm), ld_rm, _lib.ptr(b), _lib.ptr(x), s, recipe.alpha,
              _lib.ptr(d.weights) if d.weights is not None else None, 1 if d.replace else 0,
              sd, ep, log.max_it, every, float(log.threshold_info),
              hist.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(n_hist),
              ctypes.byref(iters), ctypes.byref(conv), _lib.ptr(work), work.numel() * 8)
    # replay the evaluated errors through the logger (same bookkeeping as
    # update_logger! at each evaluated iteration)
    k = int(n_hist.value)
    its = [1 + j * every for j in range(k)]
    if k and its[-1] != int(iters.value):
        its[-1] = int(iters.value)
    for j in range(k):
        log.update(hist[j], its[j])
    log.iteration = int(iters.value)
    log.converged = bool(conv.value)


def _rsolve_loop(recipe, x, A, b):
    """The host-driven loop for the blocks the one-launch device step does not
    take: other compressors, Float64 data, more than 64 rows.  Blocks of at
    most 64 Float32 rows use the device step (FP64 Gram, LDL'); the others
    the QR-based LQ projection (rla_kaczmarz_lq_update_*)."""
    S, log = recipe.compressor, recipe.log
    every = recipe.error.every
    n = _size2(A)[1]
    f32 = x.dtype == torch.float32
    nb = _dense.query_bytes("rla_kaczmarz_workspace_bytes", 64, n, 64, 1)
    work = torch.empty(max(1, (nb + 7) // 8), dtype=torch.float64, device=x.device)
    lq_work = None
    for i in range(1, log.max_it + 1):
        if (i - 1) % every == 0 or i == log.max_it:
            err = compute_error(recipe.error, recipe, A, b)
            log.update(err, i)
            if log.converged:
                return
        update_compressor(S, x, A, b)
        rows_s = S.size(1)
        Mt = recipe.compressed_mat[:rows_s]
        bt = recipe.compressed_vec[:rows_s]
        mul(Mt, S, A)
        mul(bt, S, b)
        Mv = as_view(Mt)
        if f32 and rows_s <= 64:
            _lib.call("rla_kaczmarz_block_update_f32", _lib.ctx(x.device), rows_s, n,
                      _lib.ptr(Mv.t), Mv.rs, Mv.cs, _lib.ptr(bt), bt.stride(0), _lib.ptr(x),
                      recipe.alpha, _lib.ptr(work), work.numel() * 8)
            continue
        if rows_s > n:
            raise ArgumentError("a Kaczmarz block of more rows than columns is not supported")
        if lq_work is None or lq_work[0] != rows_s:
            nbl = _dense.query_bytes("rla_kaczmarz_lq_workspace_bytes", rows_s, n)
            lq_work = (rows_s, _dense.workspace(nbl, x.device))
        w = lq_work[1]
        _lib.call(f"rla_kaczmarz_lq_update_{_lib.sfx(x.dtype)}", _lib.ctx(x.device), rows_s, n,
                  _lib.ptr(Mv.t), Mv.rs, Mv.cs, _lib.ptr(bt), bt.stride(0), _lib.ptr(x),
                  recipe.alpha, _lib.ptr(w), w.numel() * 8)


# ---------------------------------------------------------------------------
# IHS (src/Solvers/ihs.jl)


class IHS(Solver):
    """IHS(compressor=SparseSign(cardinality=Left()), log=BasicLogger(),
    error=FullResidual(), alpha=1.0) — iterative Hessian sketch for
    overdetermined least squares, ihs.jl:1-76."""

    def __init__(self, compressor=None, log=None, error=None, alpha=1.0):
        import warnings
        self.compressor = SparseSign(cardinality=Left()) if compressor is None else compressor
        if self.compressor.cardinality != Left():
            warnings.warn("Compressor has cardinality `Right` but IHS compresses from the `Left`.")
        if alpha < 0:
            warnings.warn("Negative step size could lead to divergent iterates.")
        self.log = BasicLogger() if log is None else log
        self.error = FullResidual() if error is None else error
        self.alpha = float(alpha)


class IHSRecipe:
    """ihs.jl:78-121: the compressor / logger / error recipes, alpha, the
    compressed matrix buffer (s x n, Float32, column-major), the residual
    b - A x, the gradient A'(b - A x), and on the device side the FP64 copy of
    the compressed matrix, its triangular factor R and the QR workspace."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def _complete_ihs(ingredients, x, A, b):
    """complete_solver(::IHS, x, A, b) — ihs.jl:122-201 (argument checks and
    their messages as there)."""
    compressor = complete_compressor(ingredients.compressor, x, A, b)
    logger = complete_logger(ingredients.log)
    if not isinstance(ingredients.error, FullResidual):
        raise ArgumentError(
            f"ErrorRecipe {type(ingredients.error).__name__} does not contain the field "
            f"'residual' and is not valid for an IHS solver.")
    error = FullResidualRecipe(ingredients.error, b)
    sample_size = compressor.n_rows
    rows_a, cols_a = _size2(A)
    if cols_a > sample_size:
        raise ArgumentError(
            "Compression dimension not larger than column dimension this will lead to singular "
            "QR decompositions, which cannot be inverted.")
    if rows_a < sample_size:
        raise ArgumentError("Compression dimension larger row dimension.")
    if cols_a >= rows_a:
        raise ArgumentError("Matrix must have more rows than columns.")
    for t in (as_view(A).t, b, x):
        if t.dtype != torch.float32:
            raise ArgumentError(f"the GPU IHS solver runs on Float32 data (got {t.dtype})")
    dev = as_view(A).t.device
    F32, F64 = torch.float32, torch.float64
    qr_bytes = _dense.query_bytes("rla_qr_workspace_bytes", sample_size, cols_a)
    return IHSRecipe(
        log=logger, compressor=compressor, error=error, alpha=ingredients.alpha,
        compressed_mat=torch.zeros(cols_a, sample_size, dtype=F32, device=dev).T,
        residual_vec=torch.zeros(rows_a, dtype=F32, device=dev),
        gradient_vec=torch.zeros(cols_a, dtype=F32, device=dev),
        solution_vec=x,
        R=torch.zeros(cols_a, cols_a, dtype=F64, device=dev).T,
        qr_work=_dense.workspace(qr_bytes, dev), qr_bytes=qr_bytes)


def _col(v):
    """(n x 1) View of a contiguous device vector (leading dimension n)."""
    return as_view(v.reshape(-1))


def _rsolve_ihs(recipe, x, A, b):
    """rsolve!(::IHSRecipe, x, A, b) — ihs.jl:214-254.  Per iteration:
    g = A'r; the error and logger; a new sketch S and SA = S*A (the sparse
    and SRHT/Gaussian kernels); R with SA = QR (FP64 CholeskyQR2 of SA,
    Householder if it breaks down); u = R^-1 R'^-1 g (FP64); x += alpha u and
    r -= A u (as the reference: the residual update does not scale u by
    alpha).  FP32 vectors and products, as the reference on Float32 data."""
    log = recipe.log
    log.reset()
    a_rows, a_cols = _size2(A)
    if x.shape[0] != a_cols:
        raise DimensionMismatch(
            f"Dimension of `x`, {x.shape[0]} is different from number of columns in `A`,"
            f"{a_cols}.")
    if b.shape[0] != a_rows:
        raise DimensionMismatch(
            f"Dimension of `b`, {b.shape[0]} is different from number of rows in `A`,"
            f"{a_rows}.")
    recipe.solution_vec = x
    Av = as_view(A)
    dev = Av.t.device
    ctx = _lib.ctx(dev)
    r, g = recipe.residual_vec, recipe.gradient_vec
    one = torch.ones(1, 1, dtype=torch.float32, device=dev)
    # r = b - A x  (ihs.jl:218-221)
    r.copy_(b)
    _dense.gemm(_col(r), Av, _col(x), -1.0, 1.0)
    S = recipe.compressor
    for i in range(1, log.max_it + 1):
        _dense.gemm(_col(g), Av.T, _col(r), 1.0, 0.0)  # g = A'r
        err = compute_error(recipe.error, recipe, A, b)
        log.update(err, i)
        if log.converged:
            return
        update_compressor(S, x, A, b)
        rows_s = S.size(1)
        SA = recipe.compressed_mat[:rows_s]
        mul(SA, S, A)
        SA64 = _dense.convert(SA, torch.float64)  # s x n, column-major
        hh = ctypes.c_int(0)
        _lib.call("rla_qr_f64", ctx, rows_s, a_cols, _lib.ptr(SA64), max(1, rows_s),
                  _lib.ptr(recipe.R), a_cols, _lib.ptr(recipe.qr_work), recipe.qr_bytes,
                  ctypes.byref(hh))
        u = _dense.convert(g.view(-1, 1), torch.float64)[:, 0].contiguous()
        _lib.call("rla_trsv_upper_f64", ctx, 1, a_cols, _lib.ptr(recipe.R), a_cols, _lib.ptr(u))
        _lib.call("rla_trsv_upper_f64", ctx, 0, a_cols, _lib.ptr(recipe.R), a_cols, _lib.ptr(u))
        u32 = _dense.convert(u.view(-1, 1), torch.float32)
        _dense.gemm(_col(x), as_view(u32), as_view(one), recipe.alpha, 1.0)  # x += alpha u
        _dense.gemm(_col(r), Av, as_view(u32), -1.0, 1.0)                    # r -= A u


# ---------------------------------------------------------------------------
# ColumnProjection (src/Solvers/col_projection.jl)


class ColumnProjection(Solver):
    """ColumnProjection(alpha=1.0, compressor=SparseSign(cardinality=Right()),
    error=LSGradient(), log=BasicLogger(), sub_solver=QRSolver()) — block
    coordinate descent for least squares, col_projection.jl:1-121."""

    def __init__(self, alpha=1.0, compressor=None, error=None, log=None, sub_solver=None):
        import warnings

        from .batched import QRSolver
        self.compressor = SparseSign(cardinality=Right()) if compressor is None else compressor
        if self.compressor.cardinality != Right():
            warnings.warn("Compressor has cardinality `Left` but ColumnProjection compresses "
                          "from the `Right`. This may cause an inefficiency.")
        self.alpha = float(alpha)
        self.error = LSGradient() if error is None else error
        self.log = BasicLogger() if log is None else log
        self.sub_solver = QRSolver() if sub_solver is None else sub_solver


class ColumnProjectionRecipe:
    """col_projection.jl:123-173: the compressor / logger / error recipes,
    alpha, the compressed matrix buffer A*S (m x k, Float32, column-major),
    the solution, the update v, the residual b - A x, and on the device side
    the triangular factor R of A*S (FP64) and the QR workspace."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def _complete_colproj(ingredients, x, A, b):
    """complete_solver(::ColumnProjection, x, A, b) — col_projection.jl:175-247
    (the error-recipe check and its message as there)."""
    from .batched import QRSolver
    compressor = complete_compressor(ingredients.compressor, A, b)
    logger = complete_logger(ingredients.log)
    if not isinstance(ingredients.error, LSGradient):
        raise ArgumentError(
            f"ErrorRecipe {type(ingredients.error).__name__} does not contain the field "
            f"'gradient' and is not valid for a `ColumnProjection` solver.")
    if not isinstance(ingredients.sub_solver, QRSolver):
        raise ArgumentError("the GPU ColumnProjection solver projects with QRSolver() only")
    for t in (as_view(A).t, b, x):
        if t.dtype != torch.float32:
            raise ArgumentError(
                f"the GPU ColumnProjection solver runs on Float32 data (got {t.dtype})")
    rows_a, cols_a = _size2(A)
    sample_size = compressor.n_cols
    if sample_size > rows_a:
        raise ArgumentError("Compression dimension larger than the row dimension of A.")
    dev = as_view(A).t.device
    F32, F64 = torch.float32, torch.float64
    qr_bytes = _dense.query_bytes("rla_qr_workspace_bytes", rows_a, sample_size)
    return ColumnProjectionRecipe(
        compressor=compressor, log=logger,
        error=LSGradientRecipe(torch.zeros(cols_a, dtype=F32, device=dev)),
        sub_solver=ingredients.sub_solver, alpha=ingredients.alpha,
        compressed_mat=torch.zeros(sample_size, rows_a, dtype=F32, device=dev).T,
        solution_vec=x, update_vec=torch.zeros(sample_size, dtype=F32, device=dev),
        residual_vec=torch.zeros(rows_a, dtype=F32, device=dev),
        R=torch.zeros(sample_size, sample_size, dtype=F64, device=dev).T,
        qr_work=_dense.workspace(qr_bytes, dev), qr_bytes=qr_bytes)


def _rsolve_colproj(recipe, x, A, b):
    """rsolve!(::ColumnProjectionRecipe, x, A, b) — col_projection.jl:322-372.
    Per iteration: the LSGradient error and logger (iteration i - 1); a new
    sketch S (n x k) and AS = A*S (the compressor kernels); v minimises
    ||AS v - r||: R from FP64 CholeskyQR2 of AS (Householder if it breaks
    down), v = R^-1 R'^-1 (AS)'r (FP64 solves); x += alpha S v (the
    compressor applied to v) and r -= alpha AS v.  One column (k = 1) is the
    same formula, v = a'r / a'a (colproj_update!, :265-275).  The error after
    the last iteration is logged at max_it (:368-371)."""
    log = recipe.log
    log.reset()
    a_rows, a_cols = _size2(A)
    if x.shape[0] != a_cols:
        raise DimensionMismatch(
            f"Dimension of `x`, {x.shape[0]} is different from number of columns in `A`,"
            f"{a_cols}.")
    if b.shape[0] != a_rows:
        raise DimensionMismatch(
            f"Dimension of `b`, {b.shape[0]} is different from number of rows in `A`,"
            f"{a_rows}.")
    recipe.solution_vec = x
    Av = as_view(A)
    ctx = _lib.ctx(Av.t.device)
    r = recipe.residual_vec
    # r = b - A x  (col_projection.jl:329-333)
    r.copy_(b)
    _dense.gemm(_col(r), Av, _col(x), -1.0, 1.0)
    S = recipe.compressor
    for i in range(1, log.max_it + 1):
        err = compute_error(recipe.error, recipe, A, b)
        log.update(err, i - 1)
        if log.converged:
            return
        update_compressor(S, x, A, b)
        cols_s = S.size(2)
        AS = recipe.compressed_mat[:, :cols_s]
        mul(AS, A, S)
        AS64 = _dense.convert(AS, torch.float64)  # m x k, column-major
        hh = ctypes.c_int(0)
        _lib.call("rla_qr_f64", ctx, a_rows, cols_s, _lib.ptr(AS64), max(1, a_rows),
                  _lib.ptr(recipe.R), cols_s, _lib.ptr(recipe.qr_work), recipe.qr_bytes,
                  ctypes.byref(hh))
        v = recipe.update_vec[:cols_s]
        _dense.gemm(_col(v), as_view(AS).T, _col(r), 1.0, 0.0)  # v = (AS)'r
        v64 = _dense.convert(v.view(-1, 1), torch.float64)[:, 0].contiguous()
        _lib.call("rla_trsv_upper_f64", ctx, 1, cols_s, _lib.ptr(recipe.R), cols_s,
                  _lib.ptr(v64))
        _lib.call("rla_trsv_upper_f64", ctx, 0, cols_s, _lib.ptr(recipe.R), cols_s,
                  _lib.ptr(v64))
        v.copy_(_dense.convert(v64.view(-1, 1), torch.float32)[:, 0])
        mul(x, S, v, recipe.alpha, 1.0)                                  # x += alpha S v
        _dense.gemm(_col(r), as_view(AS), _col(v), -recipe.alpha, 1.0)  # r -= alpha AS v
    err = compute_error(recipe.error, recipe, A, b)
    log.update(err, log.max_it)
