"""Pins the oracle (oracle/restate.py, oracle/rla_oracle.c) against closed
forms, invariants and known-answer vectors, independently of the GPU."""
import ctypes
import math

import numpy as np
import pytest
import scipy.linalg

from oracle import restate as R


# Philox4x32-10 known-answer vectors published with Random123 (kat_vectors).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_kat_python(ctr, key, want):
    assert tuple(R.philox4x32_10(list(ctr), list(key))) == want


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_kat_library(rla, ctr, key, want):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    out = (ctypes.c_uint32 * 4)()
    rla._lib.load().rla_philox4x32_10(c, k, out)
    assert tuple(out) == want


def _signs(n, seed):
    return np.random.default_rng(seed).random(n) < 0.5


@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 256, 1024])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_fwht_matches_dense_hadamard(n, dtype, tol):
    # fwht.jl tests compare against Hadamard.hadamard(n) * x (Sylvester order)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, 3)).astype(dtype)
    sg = _signs(n, n + 1)
    sc = 0.37
    got = R.fwht(x, sg, sc, dtype)
    H = scipy.linalg.hadamard(n).astype(np.float64)
    want = H @ (np.where(sg, sc, -sc)[:, None] * x.astype(np.float64))
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want))
    # sign flipping really matters: the unsigned transform differs
    if n >= 4:
        assert not np.allclose(R.fwht(x, None, sc, dtype), got)


def test_fwht_self_inverse():
    # fwht.jl test: H(Hx)/n == x  (scaling = 1/n on the second application)
    n = 4096
    x = np.random.default_rng(3).standard_normal((n, 2))
    y = R.fwht(R.fwht(x, None, 1.0), None, 1.0 / n)
    assert np.max(np.abs(y - x)) < 1e-12


def test_fwht_stage_order_matters_in_float32():
    # a different stage order would change float32 rounding; the restatement
    # follows h = 1, 2, 4, ...  Check against a scalar transcription of
    # fwht.jl:21-70 on one column.
    n = 256
    x = np.random.default_rng(5).standard_normal(n).astype(np.float32) * 1e3
    sg = _signs(n, 6)
    sc = float(np.float32(math.sqrt(16) / math.sqrt(n)))
    ref = x.copy()
    for i in range(0, n, 2):
        z = float(ref[i]) * (sc if sg[i] else -sc)
        y = float(ref[i + 1]) * (sc if sg[i + 1] else -sc)
        ref[i], ref[i + 1] = np.float32(z + y), np.float32(z - y)
    h = 2
    while h < n:
        for i in range(0, n, 2 * h):
            for j in range(i, i + h):
                z, y = ref[j], ref[j + h]
                ref[j], ref[j + h] = np.float32(z + y), np.float32(z - y)
        h *= 2
    assert np.array_equal(R.fwht(x, sg, sc, np.float32), ref)


def _pack(sg):
    n = sg.shape[0]
    pad = np.zeros(((n + 63) // 64) * 64, dtype=np.uint64)
    pad[:n] = sg
    return (pad.reshape(-1, 64) << np.arange(64, dtype=np.uint64)).sum(1, dtype=np.uint64)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fwht_numpy_equals_c_restatement(coracle, dtype):
    n, k = 2048, 5
    x = np.random.default_rng(7).standard_normal((n, k)).astype(dtype)
    sg = _signs(n, 8)
    ch = _pack(sg)
    want = R.fwht(x, sg, 0.25, dtype)
    xf = np.asfortranarray(x.copy())
    fn = coracle.oracle_fwht_f32 if dtype == np.float32 else coracle.oracle_fwht_f64
    fn(xf.ctypes.data, n, k, n, ch.ctypes.data, 0.25, 2)
    assert np.array_equal(xf, want)


@pytest.mark.parametrize("m,s", [(8, 4), (5, 3), (1000, 64), (4096, 4096)])
def test_srht_left_matches_reference_test_formula(m, s):
    # test/Compressors/srht.jl:291-296: (H * (signs .* pad(A)) * alpha * scale)[op, :] + beta*C
    rng = np.random.default_rng(m)
    A = rng.standard_normal((m, 4))
    n_pad = R.padded_size(m)
    op = rng.choice(n_pad, s, replace=False)
    sg = _signs(n_pad, m + 2)
    scale = math.sqrt(s) / math.sqrt(n_pad)
    C0 = rng.standard_normal((s, 4))
    got = R.srht_left_mul(A, op, sg, scale, np.float64, C0, 1.5, 0.5)
    P = np.zeros((n_pad, 4))
    P[:m] = A
    H = scipy.linalg.hadamard(n_pad).astype(float)
    want = (H @ (np.where(sg, 1.0, -1.0)[:, None] * P) * 1.5 * scale)[op] + 0.5 * C0
    assert np.max(np.abs(got - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))
    # textbook special case: s = n_pad, identity sample, all signs +: scaled Hadamard
    if m == s == n_pad:
        got2 = R.srht_left_mul(A, np.arange(n_pad), np.ones(n_pad, bool), 1.0, np.float64)
        assert np.allclose(got2, H @ A, rtol=0, atol=1e-9)


def test_srht_adjoint_is_transpose_of_forward():
    rng = np.random.default_rng(11)
    m, s = 300, 40
    n_pad = R.padded_size(m)
    op = rng.choice(n_pad, s, replace=False)
    sg = _signs(n_pad, 12)
    scale = math.sqrt(s) / math.sqrt(n_pad)
    T = R.srht_left_mul(np.eye(m), op, sg, scale, np.float64)   # s x m operator
    X = rng.standard_normal((s, 3))
    got = R.srht_adjoint_mul(X, op, sg, scale, m, np.float64, None, 2.0, 0.0)
    assert np.max(np.abs(got - 2.0 * T.T @ X)) < 1e-12


@pytest.mark.parametrize("m,ncols,s", [(4096, 5, 256), (1000, 3, 64), (70000, 2, 2048)])
def test_srht_left_c_restatement_equals_numpy(coracle, m, ncols, s):
    # oracle_srht_left_f32 (the CPU-baseline leg) against the numpy restatement,
    # which test_srht_left_matches_reference_test_formula pins to the reference
    # test formula: same operation order, so the bits must agree, including
    # padded rows (m not a power of two), alpha and beta
    rng = np.random.default_rng(m + s)
    A = np.asfortranarray(rng.standard_normal((m, ncols)).astype(np.float32))
    n_pad = R.padded_size(m)
    op = rng.choice(n_pad, s, replace=False).astype(np.int32)
    sg = _signs(n_pad, m + 3)
    scale = float(np.float32(math.sqrt(s) / math.sqrt(n_pad)))
    C0 = np.asfortranarray(rng.standard_normal((s, ncols)).astype(np.float32))
    chunks, neg_chunks = _pack(sg), _pack(~sg)  # kept alive across the foreign calls
    for alpha, beta in ((1.0, 0.0), (1.5, 0.5)):
        want = R.srht_left_mul(A, op, sg, scale, np.float32, C0.copy(), alpha, beta)
        C = C0.copy(order="F")
        coracle.oracle_srht_left_f32(A.ctypes.data, m, ncols, m, n_pad, s, op.ctypes.data,
                                     chunks.ctypes.data, scale, C.ctypes.data, s, alpha,
                                     beta, 2)
        assert np.array_equal(C, want)
    # a planted mistake (the sign vector negated) must be seen
    C = C0.copy(order="F")
    coracle.oracle_srht_left_f32(A.ctypes.data, m, ncols, m, n_pad, s, op.ctypes.data,
                                 neg_chunks.ctypes.data, scale, C.ctypes.data, s, 1.0, 0.0, 2)
    assert not np.array_equal(C, R.srht_left_mul(A, op, sg, scale, np.float32))


@pytest.mark.parametrize("m,s,nnz", [(10, 6, 3), (300, 40, 8), (64, 64, 64)])
def test_sparse_rowgather_matches_dense_product(m, s, nnz):
    idx, val = R.sparse_sign_gen(99, 0, m, s, nnz)
    M = np.zeros((s, m))
    for j in range(m):
        M[idx[j * nnz:(j + 1) * nnz], j] = val[j * nnz:(j + 1) * nnz]
    rng = np.random.default_rng(m)
    X = rng.standard_normal((m, 5))
    C0 = rng.standard_normal((s, 5))
    got = R.sparse_mul(M, X, 0.75, 2.0, C0)
    want = 0.75 * M @ X + 2.0 * C0
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    # M' * Y (the Right / adjoint orientation) through the CSC lists
    Y = rng.standard_normal((s, 2))
    gotT = R.sparse_mul(M.T, Y)
    assert np.max(np.abs(gotT - M.T @ Y)) < 1e-12 * max(1, np.max(np.abs(M.T @ Y)))


def test_sparse_gather_equals_c_scatter_loop(coracle):
    # the gather restatement and the CSC scatter loop of SparseArrays produce
    # the same sums in the same order, hence identical bits
    m, s, nnz, k = 2000, 128, 8, 6
    idx, val = R.sparse_sign_gen(5, 1, m, s, nnz)
    M = np.zeros((s, m))
    for j in range(m):
        M[idx[j * nnz:(j + 1) * nnz], j] = val[j * nnz:(j + 1) * nnz]
    X = np.asfortranarray(np.random.default_rng(1).standard_normal((m, k)))
    C = np.asfortranarray(np.random.default_rng(2).standard_normal((s, k)))
    want = R.sparse_mul(M, X, 1.25, 0.5, C)
    colptr = np.arange(0, m * nnz + 1, nnz, dtype=np.int32)
    coracle.oracle_sparse_left_f64(colptr.ctypes.data, idx.ctypes.data, val.ctypes.data, s, m,
                                   X.ctypes.data, k, m, C.ctypes.data, s, 1.25, 0.5, 2)
    assert np.array_equal(C, want)


def test_sparse_sign_gen_structure():
    idx, val = R.sparse_sign_gen(1234, 0, 500, 32, 8)
    for j in range(500):
        r = idx[j * 8:(j + 1) * 8]
        assert np.all(np.diff(r) > 0) and r.min() >= 0 and r.max() < 32   # distinct, sorted
    assert set(np.unique(np.abs(val))) == {1 / math.sqrt(8)}
    frac = np.mean(val > 0)
    assert 0.45 < frac < 0.55
    # rows uniform: chi-square over 32 bins, 4000 draws
    counts = np.bincount(idx, minlength=32)
    chi2 = ((counts - 125.0) ** 2 / 125.0).sum()
    assert chi2 < 70  # 31 dof, p ~ 1e-4


def test_count_sketch_gen_distribution():
    # count_sketch.jl:124-127: each column has one entry, at a row uniform on
    # 1:s, with a Rademacher sign; rows and signs independent of each other
    s, n = 64, 12800
    idx, val = R.count_sketch_gen(1234, 0, n, s)
    assert idx.shape == (n,) and val.shape == (n,)
    assert idx.min() >= 0 and idx.max() < s
    assert set(np.unique(val)) == {-1.0, 1.0}
    # uniform rows: chi-square over s bins (63 dof; 120 is p ~ 1e-5)
    counts = np.bincount(idx, minlength=s)
    e = n / s
    assert ((counts - e) ** 2 / e).sum() < 120
    # sign balance: |#(+1) - n/2| within 4.5 sigma
    assert abs((val > 0).sum() - n / 2) < 4.5 * math.sqrt(n / 4)
    # independence of row and sign: 2 x s contingency chi-square (63 dof)
    pos = np.bincount(idx[val > 0], minlength=s)
    neg = counts - pos
    exp_pos = counts * (val > 0).mean()
    exp_neg = counts - exp_pos
    chi2 = ((pos - exp_pos) ** 2 / exp_pos).sum() + ((neg - exp_neg) ** 2 / exp_neg).sum()
    assert chi2 < 120
    # successive columns are not correlated (a stream reused per column would be)
    assert abs(np.corrcoef(idx[:-1], idx[1:])[0, 1]) < 0.05
    # another epoch draws another operator
    idx2, _ = R.count_sketch_gen(1234, 1, n, s)
    assert np.mean(idx2 == idx) < 0.1


def test_floyd_is_uniform_over_subsets():
    # all C(5,2) = 10 subsets of a 2-sample from 5 should appear equally often
    from collections import Counter
    c = Counter(tuple(R.floyd_sorted(R.Stream(7, 0, 1, j), 5, 2)) for j in range(5000))
    assert len(c) == 10
    chi2 = sum((v - 500.0) ** 2 / 500.0 for v in c.values())
    assert chi2 < 33  # 9 dof, p ~ 1e-4


def test_lemire_below_is_unbiased():
    st = R.Stream(3, 0, 1, 0)
    draws = np.array([st.below(7) for _ in range(14000)])
    counts = np.bincount(draws, minlength=7)
    assert counts.min() > 1800 and counts.max() < 2200


def test_srht_gen_structure():
    op, sg = R.srht_gen(1, 2, 1024, 100)
    assert len(set(op.tolist())) == 100 and op.min() >= 0 and op.max() < 1024
    assert not np.all(np.diff(op) > 0)  # unordered, like sample(..., replace=false)
    assert 0.4 < sg.mean() < 0.6


@pytest.mark.parametrize("n,want", [(1, 1), (2, 2), (3, 4), (8, 8), (1000, 1024), (65536, 65536),
                                    (2 ** 29, 2 ** 30), (2 ** 29 + 1, 2 ** 30),
                                    (2 ** 28, 2 ** 28)])
def test_padded_size_matches_log_ratio_quirk(n, want):
    assert R.padded_size(n) == want


def _lowrank(m, n, r, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((m, r)) @ rng.standard_normal((r, n))


@pytest.mark.parametrize("q", [0, 1, 2])
def test_rangefinder_oracle_exact_recovery(q):
    # test/Approximators/RangeApproximator/rangefinder.jl:190,247: with k >= rank(A),
    # ||A - Q Q'A|| < 1e-10 and Q'Q = I
    A = _lowrank(60, 40, 5, q)
    Om = np.random.default_rng(9).standard_normal((40, 8))
    for fn in (R.rand_power_it, R.rand_ortho_it):
        Q = fn(A, Om, q)
        assert Q.shape == (60, 8)
        assert np.linalg.norm(Q.T @ Q - np.eye(8)) < 1e-12
        assert np.linalg.norm(A - Q @ (Q.T @ A)) < 1e-10 * np.linalg.norm(A)


def test_rangefinder_oracle_formula():
    # range_finder_helpers.jl:66-68: rand_power_it spans qr(A*A'*A*A'*A*Omega)
    rng = np.random.default_rng(4)
    A = rng.standard_normal((50, 30))
    Om = rng.standard_normal((30, 6))
    Q = R.rand_power_it(A, Om, 2)
    Qf, _ = np.linalg.qr(A @ A.T @ A @ A.T @ A @ Om)
    assert np.linalg.norm(Q @ Q.T - Qf @ Qf.T) < 1e-10
    Qo = R.rand_ortho_it(A, Om, 2)  # same subspace in exact arithmetic
    assert np.linalg.norm(Qo @ Qo.T - Qf @ Qf.T) < 1e-8


def test_randsvd_oracle_exact_recovery():
    # randsvd tests :272-273: exact recovery of a rank-5 matrix
    A = _lowrank(40, 30, 5, 2)
    Q = R.rand_ortho_it(A, np.random.default_rng(1).standard_normal((30, 7)), 1)
    U, S, V = R.randsvd(A, Q)
    assert np.linalg.norm(A - (U * S) @ V.T) < 1e-10 * np.linalg.norm(A)
    assert np.allclose(S[:5], np.linalg.svd(A, compute_uv=False)[:5], rtol=1e-10)


# ---------------------------------------------------------------------------
# Kaczmarz oracle (kaczmarz.jl:262-311, lq.jl:35-46)


@pytest.mark.parametrize("s,n", [(1, 7), (5, 12), (16, 40), (40, 40)])
def test_kaczmarz_block_update_is_pseudoinverse_projection(s, n):
    rng = np.random.default_rng(s + n)
    At, x, bt = rng.standard_normal((s, n)), rng.standard_normal(n), rng.standard_normal(s)
    got = R.kaczmarz_update_block(x, At, bt, 0.9)
    want = x + 0.9 * np.linalg.pinv(At) @ (bt - At @ x)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    # alpha = 1 lands on the block's solution set, orthogonally (minimum norm)
    x1 = R.kaczmarz_update_block(x, At, bt, 1.0)
    assert np.allclose(At @ x1, bt, atol=1e-10)
    u = x1 - x
    assert np.allclose(u, At.T @ np.linalg.lstsq(At.T, u, rcond=None)[0], atol=1e-10)


def test_kaczmarz_single_row_matches_block_and_zero_block_is_skipped():
    rng = np.random.default_rng(3)
    a, x, c = rng.standard_normal(9), rng.standard_normal(9), rng.standard_normal(1)
    assert np.allclose(R.kaczmarz_update(x, a, c, 1.4),
                       R.kaczmarz_update_block(x, a[None, :], c, 1.4), rtol=1e-13)
    # x - alpha*(a.x - c)/||a||^2 a, written out
    assert np.allclose(R.kaczmarz_update(x, a, c, 1.4),
                       x - 1.4 * (a @ x - c[0]) / (a @ a) * a, rtol=1e-14)
    assert np.array_equal(R.kaczmarz_update_block(x, np.zeros((3, 9)), np.ones(3), 1.0), x)


def test_successive_inclusion_closed_forms():
    w = np.array([1.0, 2.0, 3.0, 4.0, 0.0])
    W = w.sum()
    assert np.allclose(R.successive_inclusion(w, 1), w / W)
    # k = 2: P(i) = w_i/W + sum_{j != i} (w_j/W) (w_i/(W - w_j))
    p2 = np.array([w[i] / W + sum(w[j] / W * w[i] / (W - w[j]) for j in range(5) if j != i)
                   for i in range(5)])
    assert np.allclose(R.successive_inclusion(w, 2), p2)
    assert np.isclose(R.successive_inclusion(w, 3).sum(), 3.0)
    assert R.successive_inclusion(w, 3)[4] == 0.0


def test_ihs_step_with_exact_hessian_is_the_least_squares_solution():
    # with S'S = I (S = I, or S with orthonormal columns) one IHS step from any
    # x is the Newton step of ||Ax - b||^2, i.e. the least-squares solution
    rng = np.random.default_rng(11)
    A = rng.standard_normal((60, 7))
    b = rng.standard_normal(60)
    x0 = rng.standard_normal(7)
    xls = np.linalg.lstsq(A, b, rcond=None)[0]
    assert np.allclose(R.ihs_step(A, b, x0, A), xls, rtol=1e-10, atol=1e-12)
    Q, _ = np.linalg.qr(rng.standard_normal((60, 60)))
    assert np.allclose(R.ihs_step(A, b, x0, Q[:, :60] @ A), xls, rtol=1e-10, atol=1e-12)
    # a sketched Hessian moves x towards xls; alpha scales the step exactly
    SA = rng.standard_normal((30, 60)) @ A / np.sqrt(30)
    x1 = R.ihs_step(A, b, x0, SA)
    assert np.linalg.norm(A @ (x1 - xls)) < np.linalg.norm(A @ (x0 - xls))
    assert np.allclose(R.ihs_step(A, b, x0, SA, 0.5) - x0, 0.5 * (x1 - x0))


def test_colproj_step_closed_forms():
    # S square and invertible: the projection onto the whole column space is
    # the least-squares solution from any x (alpha = 1)
    rng = np.random.default_rng(12)
    A = rng.standard_normal((50, 6))
    b = rng.standard_normal(50)
    x0 = rng.standard_normal(6)
    xls = np.linalg.lstsq(A, b, rcond=None)[0]
    Q, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    assert np.allclose(R.colproj_step(A, b, x0, Q), xls, rtol=1e-10, atol=1e-12)
    assert np.allclose(R.colproj_step(A, b, x0, 3.0 * np.eye(6)), xls, rtol=1e-10, atol=1e-12)
    # one column e_j: exact coordinate minimisation, x_j += a_j'r / a_j'a_j, and
    # the new residual is orthogonal to a_j
    e = np.zeros(6)
    e[2] = 1.0
    x1 = R.colproj_step(A, b, x0, e)
    r0 = b - A @ x0
    want = x0.copy()
    want[2] += A[:, 2] @ r0 / (A[:, 2] @ A[:, 2])
    assert np.allclose(x1, want, rtol=1e-12)
    assert abs(A[:, 2] @ (b - A @ x1)) < 1e-10
    # a block of columns: the residual is orthogonal to A S; alpha scales the step
    S = rng.standard_normal((6, 3))
    x2 = R.colproj_step(A, b, x0, S)
    assert np.allclose((A @ S).T @ (b - A @ x2), 0.0, atol=1e-10)
    assert np.allclose(R.colproj_step(A, b, x0, S, 0.5) - x0, 0.5 * (x2 - x0))
    # the LSGradient error vanishes at the least-squares solution
    assert R.ls_gradient(A, xls, b) < 1e-10
    assert np.isclose(R.ls_gradient(A, x0, b), np.linalg.norm(A.T @ r0))


def test_agmon_select_known_answers_and_brute_force():
    # the reference's own known answers (test/Compressors/Distributions/agmon.jl:
    # 301-315, 382-398, 400-424, 426-443, 445-477), 0-based here
    def left(A, x, b):
        return np.asarray(A, float) @ np.asarray(x, float) - np.asarray(b, float)
    r = left([[1, 0], [0, 1], [0, 0]], [0, 0], [1, 5, 0.1])
    assert R.agmon_select(r, range(3), 1) == [1]
    A = np.array([[1.0, 0.0], [0.0, 1.0], [1.0, 1.0]])
    g = A.T @ left(A, [1, 2], [0, 0, 0])              # |A'(Ax - b)| = [4, 5]
    assert R.agmon_select(g, range(2), 1) == [1]
    r = left([[2, 1], [1, 3], [0.5, 0.5], [1, 1]], [0.5, 0.5], [1, 2, 3, 0.5])
    assert R.agmon_select(r, range(4), 1) == [int(np.argmax(np.abs(r)))] == [2]
    A = np.array([[1.0, 0.0], [0.0, 1.0]])
    assert R.agmon_select(left(A, [0, 0], [2, 3]), range(2), 1) == [1]
    assert R.agmon_select(left(A, [10, 0], [2, 3]), range(2), 1) == [0]
    r = left([[1, 0], [0, 1], [1, 0], [0, 1]], [0, 0], [2, 2, 2, 1])
    assert R.agmon_select(r, range(4), 2) == [0, 1]     # tie at the cut: smaller index
    assert R.agmon_select([0.1, 0.2, 9.9, 0.3], range(4), 1) == [2]
    # brute force: every subset order on small random cases with ties
    rng = np.random.default_rng(5)
    for _ in range(200):
        v = rng.integers(-3, 4, 9).astype(float)
        cand = rng.choice(9, 6, replace=False)
        k = int(rng.integers(1, 7))
        got = R.agmon_select(v, cand, k)
        for a, b in zip(got, got[1:]):
            assert abs(v[a]) > abs(v[b]) or (abs(v[a]) == abs(v[b]) and a < b)
        rest = [c for c in cand if c not in got]
        assert all(abs(v[got[-1]]) > abs(v[c]) or
                   (abs(v[got[-1]]) == abs(v[c]) and got[-1] < c) for c in rest)


def test_full_residual_closed_form():
    A = np.array([[1.0, 2.0], [3.0, 4.0], [0.0, 1.0]])
    assert R.full_residual(A, [1.0, 1.0], [3.0, 7.0, 2.0]) == 1.0


# ---------------------------------------------------------------------------
# FJLT (fjlt.jl): pinned to the closed forms the reference's own tests use
# (test/Compressors/fjlt.jl:343,371,402,428,460) with scipy's dense Hadamard


def _fjlt_case(m, s, q, seed):
    from scipy.linalg import hadamard
    n_pad = R.padded_size(m)
    cp, ri, va = R.fjlt_gen(seed, 1, s, n_pad, q)
    K = R.csc_dense(s, n_pad, cp, ri, va)
    rng = np.random.default_rng(seed)
    sg = rng.random(n_pad) < 0.5
    scale = 1.0 / (math.sqrt(n_pad) * math.sqrt(q) * math.sqrt(s))
    return K, sg, scale, hadamard(n_pad).astype(np.float64), n_pad


@pytest.mark.parametrize("m,s", [(8, 2), (10, 2), (37, 5), (64, 9)])
def test_fjlt_left_matches_reference_test_formula(m, s):
    K, sg, scale, H, n_pad = _fjlt_case(m, s, 0.4, 11 + m)
    A = np.random.default_rng(m).standard_normal((m, 6))
    C0 = np.random.default_rng(m + 1).standard_normal((s, 6))
    pad = np.zeros((n_pad, 6))
    pad[:m] = A
    # SA = SAc + S.op * (H * (ifelse.(S.signs, 1, -1) .* pad) * 2 * S.scale)
    want = C0 + K @ (H @ (np.where(sg, 1, -1)[:, None] * pad) * 2 * scale)
    got = R.fjlt_left_mul(A, K, sg, scale, np.float64, C0, 2.0, 1.0)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("m,s", [(8, 2), (10, 3), (33, 4)])
def test_fjlt_adjoint_matches_reference_test_formula(m, s):
    K, sg, scale, H, n_pad = _fjlt_case(m, s, 0.5, 5 + m)
    A = np.random.default_rng(m).standard_normal((7, s))
    C0 = np.random.default_rng(m + 2).standard_normal((7, m))
    # SA = SAc + 2 * (A * S.op * H * Diagonal(ifelse.(S.signs, 1, -1)) * S.scale)[:, 1:m]
    want = C0 + 2 * (A @ K @ H @ np.diag(np.where(sg, 1.0, -1.0)) * scale)[:, :m]
    got = R.fjlt_adjoint_mul(A.T, K, sg, scale, m, np.float64, C0.T, 2.0, 1.0).T
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # and the adjoint is the transpose of the forward operator
    T = R.fjlt_left_mul(np.eye(m), K, sg, scale, np.float64)
    X = np.random.default_rng(3).standard_normal((s, 4))
    assert np.allclose(R.fjlt_adjoint_mul(X, K, sg, scale, m, np.float64), T.T @ X,
                       rtol=1e-12, atol=1e-13)


def test_fjlt_gen_structure():
    cp, ri, va = R.fjlt_gen(9, 4, 12, 32, 0.3)
    assert cp[0] == 0 and np.all(np.diff(cp) >= 0) and cp[-1] == len(ri) == len(va)
    for j in range(32):
        rows = ri[cp[j]:cp[j + 1]]
        assert np.all(np.diff(rows) > 0) and np.all((rows >= 0) & (rows < 12))
    # q = 1: every entry; q = 0: none; a pure function of (seed, epoch)
    cp1, ri1, _ = R.fjlt_gen(9, 4, 5, 8, 1.0)
    assert cp1[-1] == 40 and np.array_equal(ri1, np.tile(np.arange(5), 8))
    assert R.fjlt_gen(9, 4, 5, 8, 0.0)[0][-1] == 0
    assert np.array_equal(R.fjlt_gen(9, 4, 12, 32, 0.3)[1], ri)
    assert not np.array_equal(R.fjlt_gen(9, 5, 12, 32, 0.3)[1], ri)
