"""The C ABI loads and exports everything include/rla.h declares; host twins
of the recipe generators match the Python restatement bit for bit; host-side
validation mirrors the reference's errors.  No GPU needed."""
import ctypes
import math
import os
import re

import numpy as np
import pytest
import torch

from oracle import restate as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "rla.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rla_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(rla):
    lib = rla._lib.load()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bound_symbols_are_declared(rla):
    assert set(rla._lib.exported_symbols()) <= set(_declared())


def test_version_string(rla):
    assert b"gfx950" in rla._lib.load().rla_version()


def test_host_sparse_sign_gen_matches_restatement(rla):
    lib = rla._lib.load()
    n, s, nnz = 300, 50, 7
    idx = np.zeros(n * nnz, np.int32)
    val = np.zeros(n * nnz)
    assert lib.rla_host_sparse_sign_gen_f64(77, 3, n, s, nnz, idx.ctypes.data,
                                            val.ctypes.data) == 0
    ridx, rval = R.sparse_sign_gen(77, 3, n, s, nnz)
    assert np.array_equal(idx, ridx) and np.array_equal(val, rval)


def test_host_count_sketch_gen_matches_restatement(rla):
    lib = rla._lib.load()
    n, s = 1000, 37
    idx = np.zeros(n, np.int32)
    val = np.zeros(n)
    assert lib.rla_host_count_sketch_gen_f64(5, 9, n, s, idx.ctypes.data, val.ctypes.data) == 0
    ridx, rval = R.count_sketch_gen(5, 9, n, s)
    assert np.array_equal(idx, ridx) and np.array_equal(val, rval)


@pytest.mark.parametrize("n_pad,s", [(2, 1), (64, 64), (1000 and 1024, 300), (8, 5)])
def test_host_srht_gen_matches_restatement(rla, n_pad, s):
    lib = rla._lib.load()
    op = np.zeros(s, np.int32)
    ch = np.zeros((n_pad + 63) // 64, np.uint64)
    assert lib.rla_host_srht_gen(12, 4, n_pad, s, op.ctypes.data, ch.ctypes.data) == 0
    rop, rsg = R.srht_gen(12, 4, n_pad, s)
    assert np.array_equal(op, rop)
    bits = ((ch[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(bool).ravel()
    assert np.array_equal(bits[:n_pad], rsg) and not bits[n_pad:].any()


def test_host_gaussian_gen_is_standard_normal(rla):
    lib = rla._lib.load()
    r, c = 400, 250
    out = np.zeros((c, r))  # column-major r x c
    assert lib.rla_host_gaussian_gen_f64(3, 0, r, c, 1.0, out.ctypes.data, r) == 0
    z = out.ravel()
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    # Box–Muller pair 0 from the restated stream
    st = R.Stream(3, 0, R.DOM_GAUSS, 0)
    u1, u2 = st.open01(), st.open01()
    rr = np.sqrt(-2 * np.log(u1))
    assert abs(out[0, 0] - rr * np.cos(2 * np.pi * u2)) < 1e-14
    assert abs(out[0, 1] - rr * np.sin(2 * np.pi * u2)) < 1e-14


def test_host_gen_rejects_bad_arguments(rla):
    lib = rla._lib.load()
    idx = np.zeros(16, np.int32)
    val = np.zeros(16)
    assert lib.rla_host_sparse_sign_gen_f64(1, 0, 2, 4, 8, idx.ctypes.data,
                                            val.ctypes.data) == rla._lib.RLA_EARG
    assert b"must be less than or equal" in lib.rla_last_error()
    assert lib.rla_host_srht_gen(1, 0, 12, 3, idx.ctypes.data, None) == rla._lib.RLA_EDIM


def test_technique_validation_mirrors_reference(rla):
    with pytest.raises(rla.ArgumentError, match="compression_dim` must be positive"):
        rla.SparseSign(compression_dim=0)
    with pytest.raises(rla.ArgumentError, match="Number of non-zero indices, 9"):
        rla.SparseSign(compression_dim=4, nnz=9)
    with pytest.raises(rla.ArgumentError, match="`nnz` must be positive"):
        rla.SparseSign(compression_dim=4, nnz=0)
    s = rla.SparseSign(compression_dim=10, nnz=5)
    with pytest.raises(rla.ArgumentError, match="New `compression_dim`, 3"):
        s.compression_dim = 3
    with pytest.raises(rla.ArgumentError, match="New `nnz`, 11"):
        s.nnz = 11
    assert rla.SparseSign().nnz == 2 and rla.SparseSign(compression_dim=20).nnz == 8
    with pytest.raises(rla.ArgumentError):
        rla.CountSketch(cardinality=rla.Undef())
    with pytest.raises(rla.ArgumentError, match="block_size"):
        rla.SRHT(block_size=0)
    with pytest.raises(rla.ArgumentError, match="Left` or `Right"):
        rla.SRHT(cardinality=rla.Undef())


class _FakeRecipe:
    pass


def test_dimension_checks_mirror_reference_messages(rla):
    class Rec(rla.CompressorRecipe):
        n_rows, n_cols = 4, 10
    S = Rec()
    A = torch.zeros(9, 3)
    C = torch.zeros(4, 3)
    with pytest.raises(rla.DimensionMismatch, match="Matrix A has 9 rows while S has 10 columns."):
        rla.left_mul_dimcheck(C, S, A)
    with pytest.raises(rla.DimensionMismatch,
                       match="Matrix A has 3 columns while C has 2 columns."):
        rla.left_mul_dimcheck(torch.zeros(4, 2), S, torch.zeros(10, 3))
    with pytest.raises(rla.DimensionMismatch, match="Matrix C has 5 rows while S has 4 rows."):
        rla.left_mul_dimcheck(torch.zeros(5, 3), S, torch.zeros(10, 3))
    with pytest.raises(rla.DimensionMismatch, match="Matrix A has 3 columns while S has 4 rows."):
        rla.right_mul_dimcheck(torch.zeros(2, 10), torch.zeros(2, 3), S)
    with pytest.raises(rla.DimensionMismatch,
                       match="Matrix C has 9 columns while S has 10 columns."):
        rla.right_mul_dimcheck(torch.zeros(2, 9), torch.zeros(2, 4), S)
    with pytest.raises(rla.DimensionMismatch, match="Matrix C has 3 rows while A has 2 rows."):
        rla.right_mul_dimcheck(torch.zeros(3, 10), torch.zeros(2, 4), S)
    # adjoint forwarding: S' is 10 x 4
    rla.left_mul_dimcheck(torch.zeros(10, 2), S.T, torch.zeros(4, 2))
    with pytest.raises(rla.DomainError):
        S.size(3)
    assert S.T.size() == (10, 4)


def test_mul_without_recipe_is_argument_error(rla):
    with pytest.raises(rla.ArgumentError):
        rla.mul(torch.zeros(2, 2), torch.zeros(2, 2), torch.zeros(2, 2))


def _host_fjlt(lib, seed, epoch, s, n_pad, q):
    cap = s * n_pad
    cp = np.zeros(n_pad + 1, np.int32)
    ri = np.zeros(cap, np.int32)
    va = np.zeros(cap)
    nnz = ctypes.c_int64(0)
    assert lib.rla_host_fjlt_gen_f64(seed, epoch, s, n_pad, q, cp.ctypes.data, ri.ctypes.data,
                                     va.ctypes.data, cap, ctypes.byref(nnz)) == 0
    return cp, ri[: nnz.value], va[: nnz.value]


@pytest.mark.parametrize("s,n_pad,q", [(12, 32, 0.3), (7, 8, 1.0), (40, 16, 0.05), (3, 4, 0.0)])
def test_host_fjlt_gen_matches_restatement(rla, s, n_pad, q):
    cp, ri, va = _host_fjlt(rla._lib.load(), 21, 8, s, n_pad, q)
    rcp, rri, rva = R.fjlt_gen(21, 8, s, n_pad, q)
    assert np.array_equal(cp, rcp) and np.array_equal(ri, rri)
    assert np.allclose(va, rva, rtol=1e-15, atol=1e-15)


def test_host_fjlt_gen_law(rla):
    # entries independently nonzero with probability q, values N(0, 1)
    s, n_pad, q = 256, 4096, 0.01
    cp, ri, va = _host_fjlt(rla._lib.load(), 3, 1, s, n_pad, q)
    N = s * n_pad
    nnz = cp[-1]
    assert abs(nnz - N * q) < 5 * math.sqrt(N * q * (1 - q))
    assert abs(va.mean()) < 5 / math.sqrt(nnz) and abs(va.var() - 1) < 5 * math.sqrt(2 / nnz)
    rows = np.bincount(ri, minlength=s)          # each ~ Binomial(n_pad, q)
    z = (rows - n_pad * q) / math.sqrt(n_pad * q * (1 - q))
    assert np.abs(z).max() < 5 and abs(z.mean()) < 5 / math.sqrt(s)
    cols = np.diff(cp)                           # each ~ Binomial(s, q)
    zc = (cols.mean() - s * q) / math.sqrt(s * q * (1 - q) / n_pad)
    assert abs(zc) < 5


def test_fjlt_constructor_errors(rla):
    with pytest.raises(rla.ArgumentError, match="compression_dim"):
        rla.FJLT(compression_dim=0)
    with pytest.raises(rla.ArgumentError, match="sparsity"):
        rla.FJLT(sparsity=1.5)
    with pytest.raises(rla.ArgumentError, match="block_size"):
        rla.FJLT(block_size=0)
    with pytest.raises(rla.ArgumentError, match="Cardinality"):
        rla.FJLT(cardinality=rla.Undef())
    f = rla.FJLT()
    assert (f.compression_dim, f.block_size, f.sparsity, f.type) == (2, 10, 0.0, torch.float64)


def test_ihs_constructor_warnings_mirror_reference(rla):
    # ihs.jl:42-57: a Right compressor and a negative step size only warn
    import warnings
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        s = rla.IHS(compressor=rla.SparseSign(cardinality=rla.Right()), alpha=-1.0)
    msgs = [str(x.message) for x in w]
    assert any("cardinality `Right`" in m for m in msgs)
    assert any("Negative step size" in m for m in msgs)
    assert s.alpha == -1.0 and isinstance(s.log, rla.BasicLogger)


def test_column_projection_constructor_mirrors_reference(rla):
    # col_projection.jl:90-121: defaults, and a Left compressor only warns
    import warnings
    s = rla.ColumnProjection()
    assert s.alpha == 1.0 and s.compressor.cardinality == rla.Right()
    assert isinstance(s.error, rla.LSGradient) and isinstance(s.log, rla.BasicLogger)
    assert isinstance(s.sub_solver, rla.QRSolver)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        rla.ColumnProjection(compressor=rla.SparseSign(cardinality=rla.Left()), alpha=1.5)
    assert any("cardinality `Left`" in str(x.message) for x in w)


def test_agmon_checks_mirror_reference(rla):
    # agmon.jl:58-66, 156-205, 235-301 (shape checks only: no device work)
    with pytest.raises(rla.ArgumentError, match="beta must be >= 1, got beta=0"):
        rla.Agmon(beta=0)
    A, b, x = torch.zeros(10, 3), torch.zeros(10), torch.zeros(3)
    with pytest.raises(rla.ArgumentError, match="`Undef\\(\\)` is not allowed in `complete_distribution`"):
        rla.complete_distribution(rla.Agmon(), x, A, b)
    with pytest.raises(rla.ArgumentError, match=r"beta must be <= number of rows \(10\), got beta=11"):
        rla.complete_distribution(rla.Agmon(cardinality=rla.Left(), beta=11), x, A, b)
    with pytest.raises(rla.ArgumentError, match=r"number of columns \(3\), got beta=4"):
        rla.complete_distribution(rla.Agmon(cardinality=rla.Right(), beta=4), x, A, b)
    with pytest.raises(rla.DimensionMismatch, match="Vector b has length 9, expected 10"):
        rla.complete_distribution(rla.Agmon(cardinality=rla.Left()), x, A, torch.zeros(9))
    with pytest.raises(rla.DimensionMismatch, match="Vector x has length 4, expected 3"):
        rla.complete_distribution(rla.Agmon(cardinality=rla.Right()), torch.zeros(4), A, b)
    mr = rla.complete_distribution(rla.Agmon(cardinality=rla.Left(), beta=5), x, A, b)
    assert mr.state_space == 10 and mr.r is None
    with pytest.raises(rla.ArgumentError, match=r"number of rows \(4\), got beta=5"):
        rla.update_distribution(mr, x, torch.zeros(4, 3), torch.zeros(4))
    r = torch.ones(10)
    rla.update_distribution(mr, x, A, b, r=r)
    assert mr.r is r
    rla.update_distribution(mr, x, A, b)
    assert mr.r is None
    with pytest.raises(rla.ArgumentError, match="cannot output 6 indices when beta=5"):
        rla.sample_distribution(torch.zeros(6, dtype=torch.int32), mr)


def test_hbm_admission_account(rla):
    # the per-device HBM account behind RLA_EADMIT: host bookkeeping, no device
    lib = rla._lib.load()
    dev = 63  # an account no other test touches
    used, budget = ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.rla_hbm_query(dev, ctypes.byref(used), ctypes.byref(budget)) == 0
    assert used.value == 0 and budget.value == 0
    # unlimited by default
    assert lib.rla_hbm_admit(dev, 1 << 40) == 0
    assert lib.rla_hbm_release(dev, 1 << 40) == 0
    assert lib.rla_hbm_set_budget(dev, 1000) == 0
    assert lib.rla_hbm_admit(dev, 600) == 0
    assert lib.rla_hbm_admit(dev, 400) == 0          # exactly at the budget
    assert lib.rla_hbm_admit(dev, 1) == rla._lib.RLA_EADMIT
    assert b"exceed the HBM budget" in lib.rla_last_error()
    with pytest.raises(rla.AdmissionError):
        rla._lib.check(lib.rla_hbm_admit(dev, 5))
    lib.rla_hbm_query(dev, ctypes.byref(used), ctypes.byref(budget))
    assert used.value == 1000 and budget.value == 1000  # a refused charge changes nothing
    assert lib.rla_hbm_release(dev, 400) == 0
    assert lib.rla_hbm_admit(dev, 2 ** 64 - 1) == rla._lib.RLA_EADMIT  # no wrap-around
    assert lib.rla_hbm_admit(dev, 400) == 0
    assert lib.rla_hbm_release(dev, 1001) == rla._lib.RLA_EARG  # more than in use
    assert lib.rla_hbm_release(dev, 1000) == 0
    assert lib.rla_hbm_set_budget(dev, 0) == 0
    assert lib.rla_hbm_admit(-1, 1) == rla._lib.RLA_EARG
    assert lib.rla_hbm_set_budget(64, 1) == rla._lib.RLA_EARG


def test_sparse_tile_workspace_query(rla):
    # the row-group tile apply's workspace: FP32 sums of 64 columns for every
    # (column group, row group of 16 * RW rows, range of 384-row blocks), with
    # RW the smallest of 4..64 covering the rows in 16 wavefronts and about
    # 256 workgroups in all; host arithmetic, no device
    lib = rla._lib.load()
    out = ctypes.c_size_t(0)

    def q(rows, in_rows, n_rhs):
        assert lib.rla_sparse_apply_tile_bytes(rows, in_rows, n_rhs, ctypes.byref(out)) == 0
        return out.value

    assert q(1024, 1000000, 64) == 256 * 1024 * 64 * 4          # 256 ranges, one group
    assert q(2048, 65536, 512) == 8 * 2 * 16 * 1024 * 64 * 4    # 8 x 2 groups, 16 ranges
    assert q(1, 5000, 3) == 14 * 64 * 64 * 4                    # RW = 4; 14 blocks, 14 ranges
    assert q(300, 40001, 16) == 105 * 512 * 64 * 4              # RW = 32; 105 blocks
    assert q(0, 1000, 64) == 0 and q(1024, 1000, 0) == 0
    assert lib.rla_sparse_apply_tile_bytes(-1, 10, 1, ctypes.byref(out)) == rla._lib.RLA_EARG
    assert lib.rla_sparse_apply_tile_bytes(4, 0, 1, ctypes.byref(out)) == rla._lib.RLA_EARG
