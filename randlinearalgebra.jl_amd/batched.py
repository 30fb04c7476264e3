"""Batched sketch-and-solve least squares and the QR sub-solver.

`SketchSolveBatch` runs many problems x_p = argmin ||S_p A_p x - S_p b_p||
(S_p a CountSketch, count_sketch.jl) as one replayable hipGraph
(rla_ls_plan_*): the reference's per-problem
`mul!(SA, S, A); mul!(Sb, S, b); ldiv!(x, QRSolverRecipe, Sb)`
(src/Solvers/SubSolvers/qr.jl:36-43) for a whole batch at once.

`QRSolver` / `complete_sub_solver` / `ldiv` mirror qr.jl for one (already
sketched) matrix of any size, Float32 or Float64: an FP64 QR of A (CholeskyQR2
with the Householder fallback), Q'b and a triangular solve.
"""
import ctypes
import json
import math
import os
import threading

import torch

from . import _dense, _lib
from ._lib import ArgumentError
from .compressors import CountSketch, Left, as_view, complete_compressor


class _Problem(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int64), ("n", ctypes.c_int64), ("lda", ctypes.c_int64),
                ("A", ctypes.c_void_p), ("b", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("s_rowptr", ctypes.c_void_p), ("s_colidx", ctypes.c_void_p),
                ("s_val", ctypes.c_void_p), ("s", ctypes.c_int64)]


def _addr(t):
    return t.data_ptr()


def _plan_fns():
    lib = _lib.load()
    if not getattr(lib, "_ls_sigs", False):
        vp = ctypes.c_void_p
        lib.rla_ls_plan_create.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(_Problem),
                                           ctypes.c_int, ctypes.POINTER(vp)]
        lib.rla_ls_plan_execute.argtypes = [vp]
        lib.rla_ls_plan_info.argtypes = [vp, vp]
        lib.rla_ls_plan_config.argtypes = [vp, vp, vp, vp]
        lib.rla_ls_plan_sketch.argtypes = [vp, ctypes.c_int64, vp]
        lib.rla_ls_plan_destroy.argtypes = [vp]
        for f in ("create", "execute", "info", "config", "sketch", "destroy"):
            getattr(lib, f"rla_ls_plan_{f}").restype = ctypes.c_int
        lib._ls_sigs = True
    return lib


class _Plan:
    """Owns one rla_ls_plan and keeps every tensor it points at alive."""

    def __init__(self, device, probs, keep, autotune):
        self.lib = _plan_fns()
        self.device = device
        self.keep = keep
        arr = (_Problem * len(probs))(*probs)
        h = ctypes.c_void_p()
        _lib.check(self.lib.rla_ls_plan_create(_lib.ctx(device), len(probs), arr, int(autotune),
                                               ctypes.byref(h)), "rla_ls_plan_create")
        self.h = h
        self.n = len(probs)

    def execute(self):
        _lib.check(self.lib.rla_ls_plan_execute(self.h), "rla_ls_plan_execute")

    def info(self):
        out = (ctypes.c_int32 * self.n)()
        _lib.check(self.lib.rla_ls_plan_info(self.h, out), "rla_ls_plan_info")
        return list(out)

    def config(self):
        cb, rows = ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_float * 3)()
        _lib.check(self.lib.rla_ls_plan_config(self.h, ctypes.byref(cb), ctypes.byref(rows), ms),
                   "rla_ls_plan_config")
        return {"cols_per_tile": cb.value, "rows_per_chunk": rows.value,
                "tuned_ms": [float(v) for v in ms]}

    def sketch(self, p, s, n):
        out = torch.empty(n + 1, s, dtype=torch.float32, device=self.device).T
        _lib.check(self.lib.rla_ls_plan_sketch(self.h, p, _lib.ptr(out)), "rla_ls_plan_sketch")
        return out

    def close(self):
        if self.h:
            self.lib.rla_ls_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Tile shapes chosen by the autotuner, by batch bucket: a committed JSON file
# beside this module (read), or the file named by RLA_TILE_CACHE (read, and
# written when a new bucket is tuned).  A hit skips the plan-time timing.
_TILE_CACHE_DEFAULT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ls_tiles.json")
_tile_lock = threading.Lock()


def _tile_cache_path():
    return os.environ.get("RLA_TILE_CACHE", _TILE_CACHE_DEFAULT)


def _load_tiles(path):
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def tile_bucket(device, shapes):
    """The autotuner's key for a batch: the GPU's architecture and the batch's
    (m, n) bucket, i.e. the problem count, the median log2(m) and the median
    n in steps of 32 (rounded), so batches of the same make share a tile."""
    arch = torch.cuda.get_device_properties(device).gcnArchName.split(":")[0]
    ms = sorted(int(round(math.log2(max(m, 1)))) for m, _, _ in shapes)
    ns = sorted((n + 16) // 32 for _, n, _ in shapes)
    cnt = int(round(math.log2(len(shapes))))
    return f"{arch}|count2^{cnt}|m2^{ms[len(ms) // 2]}|n{32 * ns[len(ns) // 2]}"


def _check_problem(A, b, x):
    Av = as_view(A)
    if Av.dtype != torch.float32 or b.dtype != torch.float32 or x.dtype != torch.float32:
        raise ArgumentError("batched least squares runs on Float32 A, b and x")
    if Av.rs != 1:
        raise ArgumentError("A must be column-major (unit row stride)")
    if b.dim() != 1 or x.dim() != 1 or b.stride(0) != 1 or x.stride(0) != 1:
        raise ArgumentError("b and x must be contiguous vectors")
    if b.shape[0] != Av.rows or x.shape[0] != Av.cols:
        raise ArgumentError("b must have size(A, 1) entries and x size(A, 2)")
    return Av


class SketchSolveBatch:
    """x_p = argmin ||S_p A_p x - S_p b_p|| for every p, S_p =
    CountSketch(compression_dim = s_p) drawn by complete_compressor (default
    s_p = 4 n_p).  A_p: m_p x n_p Float32 column-major; b_p, x_p vectors.
    `run()` replays the plan (asynchronous); `info()` gives per-problem
    breakdown flags."""

    def __init__(self, As, bs, xs, compression_dims=None, autotune=1):
        if not (len(As) == len(bs) == len(xs)) or not As:
            raise ArgumentError("As, bs and xs must be non-empty and of equal length")
        probs, keep = [], [As, bs, xs]
        self.recipes = []
        self.shapes = []
        for p, (A, b, x) in enumerate(zip(As, bs, xs)):
            Av = _check_problem(A, b, x)
            s = 4 * Av.cols if compression_dims is None else int(compression_dims[p])
            S = complete_compressor(CountSketch(cardinality=Left(), compression_dim=s,
                                                type=torch.float32), A)
            self.recipes.append(S)
            op = S.op  # s x m; CSR rows = buckets
            lda = Av.cs if Av.cols > 1 else max(Av.cs, Av.rows)  # a single column has no stride
            probs.append(_Problem(Av.rows, Av.cols, lda, _addr(Av.t), _addr(b), _addr(x),
                                  _addr(op.rowptr), _addr(op.colidx), _addr(op.val_t), s))
            self.shapes.append((Av.rows, Av.cols, s))
        keep.append(self.recipes)
        dev = As[0].device
        self.tile_key, self.tile_cached = None, False
        if autotune == 1:
            self.tile_key = tile_bucket(dev, self.shapes)
            path = _tile_cache_path()
            hit = _load_tiles(path).get(self.tile_key)
            if hit is None and path != _TILE_CACHE_DEFAULT:
                hit = _load_tiles(_TILE_CACHE_DEFAULT).get(self.tile_key)
            if hit is not None and hit.get("shape") in (0, 1, 2):
                autotune = 2 + int(hit["shape"])   # the cached shape, no timing
                self.tile_cached = True
        self.plan = _Plan(dev, probs, keep, autotune)
        if self.tile_key is not None and not self.tile_cached:
            cfg = self.plan.config()
            ms = cfg["tuned_ms"]
            best = min(range(3), key=lambda c: ms[c] if ms[c] >= 0 else float("inf"))
            self._store_tile(self.tile_key, {"shape": best, "cols_per_tile": cfg["cols_per_tile"],
                                             "rows_per_chunk": cfg["rows_per_chunk"],
                                             "tuned_ms": ms})

    @staticmethod
    def _store_tile(key, entry):
        path = _tile_cache_path()
        with _tile_lock:
            d = _load_tiles(path)
            d[key] = entry
            try:
                tmp = path + ".tmp"
                with open(tmp, "w") as f:
                    json.dump(d, f, indent=1, sort_keys=True)
                os.replace(tmp, path)
            except OSError:
                pass  # a read-only install keeps tuning at plan time

    def run(self):
        self.plan.execute()

    def info(self):
        return self.plan.info()

    def config(self):
        return self.plan.config()

    def sketch(self, p):
        m, n, s = self.shapes[p]
        return self.plan.sketch(p, s, n)

    def close(self):
        self.plan.close()


# ---------------------------------------------------------------------------
# QR sub-solver (src/Solvers/SubSolvers/qr.jl)


class QRSolver:
    """QRSolver() — qr.jl:1-13."""


class QRSolverRecipe:
    """qr.jl:15-34: holds the matrix whose least-squares problems ldiv solves."""

    def __init__(self, A):
        self.A = A
        self._qr = None     # (Q, R) in FP64, for the A of _key
        self._key = None


def complete_sub_solver(solver, A, b=None):
    if not isinstance(solver, QRSolver):
        raise ArgumentError(f"No method `complete_sub_solver` defined for {type(solver)}.")
    return QRSolverRecipe(A)


def update_sub_solver(recipe, A):
    recipe.A = A


def _check_single(A, b, x):
    Av = as_view(A)
    if Av.dtype not in (torch.float32, torch.float64):
        raise ArgumentError("QRSolver runs on Float32 or Float64 data")
    if b.dtype != Av.dtype or x.dtype != Av.dtype:
        raise ArgumentError("A, b and x must have the same element type")
    if Av.rs != 1:
        raise ArgumentError("A must be column-major (unit row stride)")
    if b.dim() != 1 or x.dim() != 1:
        raise ArgumentError("b and x must be vectors")
    if b.shape[0] != Av.rows or x.shape[0] != Av.cols:
        raise ArgumentError("b must have size(A, 1) entries and x size(A, 2)")
    if Av.rows < Av.cols:
        raise ArgumentError("QRSolver needs size(A, 1) >= size(A, 2)")
    return Av


def _qr_factor(recipe, Av):
    """Q (m x n, orthonormal columns) and R (n x n upper) of the recipe's A in
    FP64: CholeskyQR2 on the matrix cores, Householder when the Cholesky factor
    breaks down (rla_qr_f64).  Cached while A is the same, unchanged tensor;
    A itself is not overwritten (qr.jl:42's qr! is)."""
    if recipe._key is not None and _dense.key_matches(recipe._key, recipe.A):
        return recipe._qr
    dev = Av.t.device
    m, n = Av.rows, Av.cols
    if recipe.A.dtype == torch.float64:
        Q = torch.empty((n, m), dtype=torch.float64, device=dev).T
        Q.copy_(recipe.A)  # the QR overwrites its input
    else:
        Q = _dense.convert(recipe.A, torch.float64)
    R = torch.empty((n, n), dtype=torch.float64, device=dev).T
    nbytes = _dense.query_bytes("rla_qr_workspace_bytes", m, n)
    work = _dense.workspace(nbytes, dev)
    hh = ctypes.c_int(0)
    _lib.call("rla_qr_f64", _lib.ctx(dev), m, n, _lib.ptr(Q), max(1, m), _lib.ptr(R), max(1, n),
              _lib.ptr(work), nbytes, ctypes.byref(hh))
    recipe._qr = (Q, R)
    recipe._key = _dense.SourceKey(recipe.A)
    return recipe._qr


def ldiv(x, recipe, b):
    """ldiv!(x, ::QRSolverRecipe, b) (qr.jl:36-43): x = argmin ||A x - b|| for
    Float32 or Float64 A (m >= n, column-major, any size), with A = QR in FP64
    (_qr_factor), c = Q'b on the FP64 matrix cores and x = R \\ c.  As after
    Julia's qr!, an exactly zero diagonal entry of R (an exactly rank-deficient
    A) is an error (the triangular solve's SingularException); a nearly
    singular R gives the large solution the reference gives."""
    Av = _check_single(recipe.A, b, x)
    n = Av.cols
    if n == 0:
        return
    Q, R = _qr_factor(recipe, Av)
    dev = Av.t.device
    ctx = _lib.ctx(dev)
    zero = torch.nonzero(torch.diagonal(R) == 0)
    if zero.numel():
        raise _lib.DomainError(f"SingularException({int(zero[0, 0]) + 1}): the least-squares "
                               "matrix is rank deficient")
    b64 = _dense.convert(b.contiguous(), torch.float64)
    direct = x.dtype == torch.float64 and x.stride(0) == 1
    c = x if direct else torch.empty(n, dtype=torch.float64, device=dev)
    _dense.gemm(as_view(c), as_view(Q).T, as_view(b64))  # c = Q'b
    _lib.call("rla_trsv_upper_f64", ctx, 0, n, _lib.ptr(R), n, _lib.ptr(c))
    if direct:
        return
    if x.dtype == torch.float64:
        x.copy_(c)
        return
    out = x if x.stride(0) == 1 else torch.empty(n, dtype=torch.float32, device=dev)
    _lib.call("rla_convert_f64_f32", ctx, n, 1, _lib.ptr(c), n, _lib.ptr(out), n)
    if out is not x:
        x.copy_(out)
