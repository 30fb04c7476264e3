"""GPU parity: Sampling compressor (Uniform / L2Norm), the device row-set
sampler, the Kaczmarz block projection, the FP64 residual and the full
Kaczmarz solve, against the oracle and closed forms."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import restate as R

pytestmark = pytest.mark.gpu
F32 = torch.float32


def colmajor(a, dev, dtype=F32):
    t = torch.as_tensor(np.asarray(a)).to(dtype)
    return t.T.contiguous().T.to(dev) if t.dim() == 2 else t.to(dev)


def rnd(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def wsample(rla, dev, n, k, w=None, replace=False, nsets=1, seed=7, epoch=0):
    out = torch.empty(nsets, k, dtype=torch.int32, device=dev)
    wt = None if w is None else torch.as_tensor(np.asarray(w, np.float64)).to(dev)
    rla._lib.call("rla_wsample", rla._lib.ctx(dev), seed, epoch, n,
                  rla._lib.ptr(wt) if wt is not None else None, k, 1 if replace else 0, nsets,
                  rla._lib.ptr(out))
    return out.cpu().numpy()


# ---------------------------------------------------------------------------
# sampler


@pytest.mark.parametrize("n,k", [(10, 10), (50, 5), (5000, 64), (100000, 64), (9000, 1024)])
def test_wsample_uniform_distinct_sorted(rla, dev, n, k):
    S = wsample(rla, dev, n, k, nsets=8)
    for row in S:
        assert np.all(np.diff(row) > 0) and row[0] >= 0 and row[-1] < n
    assert not np.array_equal(S[0], S[1]) or k == n
    # a pure function of (seed, epoch): the same call again gives the same sets
    assert np.array_equal(S, wsample(rla, dev, n, k, nsets=8))
    # set r of a batch equals a single draw with epoch + r
    assert np.array_equal(S[3], wsample(rla, dev, n, k, epoch=3)[0])


def _host_topk(n, k, w, seed, epoch):
    """The k items with the largest A-Res keys log(u_i) / w_i (u_i from the
    sampler's counter stream, domain 5), sorted by index."""
    keys = np.full(n, -np.inf)
    for i in range(n):
        if w is None or w[i] > 0:
            u = R.Stream(seed, epoch, 5, i).open01()
            keys[i] = np.log(u) / (1.0 if w is None else w[i])
    return np.sort(np.argsort(-keys, kind="stable")[:k])


@pytest.mark.parametrize("n,k,wkind", [(5000, 64, "none"), (9000, 700, "rand"),
                                       (12000, 64, "sparse"), (3000, 1024, "rand")])
def test_wsample_is_exact_topk_of_the_keys(rla, dev, n, k, wkind):
    # the device selection (per-chunk radix select, then a merge pass) keeps
    # exactly the k largest keys, whatever their spread over the chunks
    rng = np.random.default_rng(n + k)
    if wkind == "none":
        w = None
    elif wkind == "rand":
        w = rng.uniform(0.1, 5.0, n)
    else:  # mostly zero weights: many equal keys (-inf) below the k kept
        w = np.zeros(n)
        w[rng.choice(n, 150, replace=False)] = rng.uniform(0.5, 2.0, 150)
    S = wsample(rla, dev, n, k, w=w, nsets=2, seed=21, epoch=4)
    for r in range(2):
        assert np.array_equal(S[r], _host_topk(n, k, w, 21, 4 + r))


def test_wsample_uniform_inclusion_frequencies(rla, dev):
    n, k, sets = 40, 6, 6000
    S = wsample(rla, dev, n, k, nsets=sets)
    freq = np.bincount(S.ravel(), minlength=n) / sets
    p = k / n
    z = (freq - p) / np.sqrt(p * (1 - p) / sets)
    assert np.max(np.abs(z)) < 5.0


@pytest.mark.parametrize("k", [1, 2, 3])
def test_wsample_weighted_matches_successive_sampling(rla, dev, k):
    w = np.array([1.0, 2.0, 3.0, 4.0, 0.0, 10.0, 0.5])
    sets = 20000
    S = wsample(rla, dev, len(w), k, w=w, nsets=sets, seed=11)
    assert not np.any(S == 4)  # weight 0 never drawn
    freq = np.bincount(S.ravel(), minlength=len(w)) / sets
    p = R.successive_inclusion(w, k)
    sd = np.sqrt(np.maximum(p * (1 - p), 1e-12) / sets)
    assert np.max(np.abs(freq - p) / sd) < 5.0


def test_wsample_with_replacement_frequencies(rla, dev):
    w = np.array([1.0, 3.0, 0.0, 6.0])
    sets, k = 4000, 5
    S = wsample(rla, dev, len(w), k, w=w, replace=True, nsets=sets, seed=3)
    assert np.all(np.diff(S, axis=1) >= 0)
    freq = np.bincount(S.ravel(), minlength=4) / (sets * k)
    p = w / w.sum()
    assert np.max(np.abs(freq - p) / np.sqrt(p * (1 - p) / (sets * k) + 1e-12)) < 5.0


def test_wsample_rejects_bad_arguments(rla, dev):
    with pytest.raises(rla._lib.ArgumentError):
        wsample(rla, dev, 5, 6)
    with pytest.raises(rla._lib.ArgumentError):
        wsample(rla, dev, 5000, 2000)


# ---------------------------------------------------------------------------
# Sampling compressor (sampling.jl:125-218)


@pytest.mark.parametrize("dist", ["uniform", "l2norm"])
def test_sampling_left_and_right_products(rla, dev, dist):
    A = rnd((300, 40), 1)
    A[7] = 0.0
    Ad = colmajor(A, dev, torch.float64)
    rla.seed(5)
    D = rla.Uniform() if dist == "uniform" else rla.L2Norm()
    S = rla.complete_compressor(rla.Sampling(compression_dim=12, distribution=D), Ad)
    idx = S.idx.cpu().numpy()
    assert S.size() == (12, 300) and np.all(np.diff(idx) > 0)
    if dist == "l2norm":
        assert 7 not in idx
    C = torch.full((12, 40), 7.0, dtype=torch.float64, device=dev)
    rla.mul(C, S, Ad, 2.0, 0.5)
    assert np.array_equal(C.cpu().numpy(), 2.0 * A[idx] + 3.5)
    # vector form (sampling.jl:156-157)
    b = rnd(300, 2)
    bt = torch.zeros(12, dtype=torch.float64, device=dev)
    rla.mul(bt, S, colmajor(b, dev, torch.float64))
    assert np.array_equal(bt.cpu().numpy(), b[idx])
    # A' * S' via the adjoint: (S A)'
    Ct = torch.zeros(40, 12, dtype=torch.float64, device=dev)
    rla.mul(Ct, colmajor(A.T, dev, torch.float64), S.T)
    assert np.array_equal(Ct.cpu().numpy(), A[idx].T)
    Sr = rla.complete_compressor(rla.Sampling(cardinality=rla.Right(), compression_dim=9,
                                              distribution=D), Ad)
    j = Sr.idx.cpu().numpy()
    Cr = torch.zeros(300, 9, dtype=torch.float64, device=dev)
    rla.mul(Cr, Ad, Sr)
    assert np.array_equal(Cr.cpu().numpy(), A[:, j])
    old = S.idx.clone()
    rla.update_compressor(S, None, Ad, None)
    assert not torch.equal(old, S.idx)


def test_sampling_errors(rla):
    with pytest.raises(rla._lib.ArgumentError, match="compression_dim"):
        rla.Sampling(compression_dim=0)
    with pytest.raises(rla._lib.ArgumentError, match="Undef"):
        rla.Sampling(cardinality=rla.Undef())


# ---------------------------------------------------------------------------
# block projection and residual


def block_update(rla, dev, At, bt, x, alpha):
    s, n = At.shape
    nb = rla._dense.query_bytes("rla_kaczmarz_workspace_bytes", s, n, s, 1)
    work = rla._dense.workspace(nb, dev)
    Atd, btd, xd = colmajor(At, dev), colmajor(bt, dev), colmajor(x, dev)
    rla._lib.call("rla_kaczmarz_block_update_f32", rla._lib.ctx(dev), s, n, rla._lib.ptr(Atd),
                  1, s, rla._lib.ptr(btd), 1, rla._lib.ptr(xd), alpha, rla._lib.ptr(work),
                  work.numel() * 8)
    return xd.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("s,n,alpha", [(1, 100, 1.0), (8, 100, 1.0), (64, 1000, 1.0),
                                       (64, 130, 0.7), (30, 64, 1.5), (3, 3, 1.0)])
def test_block_update_matches_oracle(rla, dev, s, n, alpha):
    At = rnd((s, n), s).astype(np.float32)
    x = rnd(n, n).astype(np.float32)
    bt = rnd(s, 2 * s).astype(np.float32)
    got = block_update(rla, dev, At, bt, x, alpha)
    want = R.kaczmarz_update_block(x, At, bt, alpha)
    # FP32 x and A, FP64 Gram/solve: within 1e-5 of the FP64 oracle
    assert np.linalg.norm(got - want) <= 1e-5 * max(1.0, np.linalg.norm(want))
    if s == 1:
        assert np.allclose(got, R.kaczmarz_update(x, At[0], bt, alpha), rtol=1e-5, atol=1e-5)


def test_block_update_skips_zero_block_and_dependent_rows(rla, dev):
    x = rnd(20, 1).astype(np.float32)
    got = block_update(rla, dev, np.zeros((4, 20), np.float32), np.ones(4, np.float32), x, 1.0)
    assert np.array_equal(got, x.astype(np.float64))
    # a row drawn twice (sampling with replacement) of a consistent system:
    # the pseudoinverse projection, which is the projection onto the other rows.
    # (The reference's LQ sub-solve divides by a zero diagonal of R here; the
    # oracle's restatement of it is rounding noise for such a block.)
    At = rnd((4, 20), 2).astype(np.float32)
    At[3] = At[1]
    bt = rnd(4, 3).astype(np.float32)
    bt[3] = bt[1]
    got = block_update(rla, dev, At, bt, x, 1.0)
    A64 = At.astype(np.float64)
    want = x + np.linalg.pinv(A64) @ (bt - A64 @ x)
    assert np.linalg.norm(got - want) <= 1e-5 * np.linalg.norm(want)
    want3 = R.kaczmarz_update_block(x, At[:3], bt[:3], 1.0)
    assert np.linalg.norm(got - want3) <= 1e-5 * np.linalg.norm(want3)
    # a row that is a combination of two others
    At[3] = 0.5 * At[0] - 2.0 * At[2]
    bt[3] = 0.5 * bt[0] - 2.0 * bt[2]
    got = block_update(rla, dev, At, bt, x, 1.0)
    want3 = R.kaczmarz_update_block(x, At[:3], bt[:3], 1.0)
    assert np.linalg.norm(got - want3) <= 1e-5 * np.linalg.norm(want3)
    # a non-finite block is rejected
    At[2, 5] = np.nan
    with pytest.raises(rla._lib.DomainError, match="not finite"):
        block_update(rla, dev, At, bt, x, 1.0)


@pytest.mark.parametrize("m,n", [(1, 1), (1000, 37), (100000, 1000)])
def test_residual_norm_fp64(rla, dev, m, n):
    A = rnd((m, n), 4).astype(np.float32)
    x = rnd(n, 5).astype(np.float32)
    b = rnd(m, 6).astype(np.float32)
    out = ctypes.c_double()
    Ad, xd, bd = colmajor(A, dev), colmajor(x, dev), colmajor(b, dev)  # keep them alive
    rla._lib.call("rla_residual_norm_f32", rla._lib.ctx(dev), m, n, rla._lib.ptr(Ad), 1, m,
                  rla._lib.ptr(xd), rla._lib.ptr(bd), ctypes.byref(out))
    assert abs(out.value - R.full_residual(A, x, b)) <= 1e-12 * R.full_residual(A, x, b)


# ---------------------------------------------------------------------------
# rsolve!


def consistent_system(m, n, seed, dev):
    A = rnd((m, n), seed).astype(np.float32)
    xs = rnd(n, seed + 1).astype(np.float32)
    Ad = colmajor(A, dev)
    b = Ad @ colmajor(xs, dev)
    return A, xs, Ad, b


@pytest.mark.parametrize("dist", ["l2norm", "uniform"])
def test_kaczmarz_device_loop_converges(rla, dev, dist):
    A, xs, Ad, b = consistent_system(3000, 120, 10, dev)
    bn = float(torch.linalg.vector_norm(b.double()))
    D = rla.L2Norm() if dist == "l2norm" else rla.Uniform()
    solver = rla.Kaczmarz(compressor=rla.Sampling(compression_dim=32, distribution=D),
                          log=rla.BasicLogger(max_it=3000, threshold=1e-6 * bn),
                          error=rla.FullResidual(every=25))
    x = torch.zeros(120, dtype=F32, device=dev)
    rla.seed(1)
    x, rec = rla.rsolve(solver, x, Ad, b)
    assert rec.log.converged and rec.log.iteration < 3000
    assert rec.log.error < 1e-6 * bn
    assert (rec.log.iteration - 1) % 25 == 0
    xe = x.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(xe - xs) <= 1e-5 * np.linalg.norm(xs)
    assert abs(R.full_residual(A, xe, b.cpu().numpy()) - rec.log.error) <= 1e-9 * bn
    # the recorded history decreases (geometric convergence)
    h = rec.log.hist[:rec.log.record_location]
    assert h[0] > h[-1]


def test_kaczmarz_device_loop_with_replacement_converges(rla, dev):
    # Uniform(replace=True) draws a row twice in some blocks (here about one
    # block in three); those blocks project onto their distinct rows
    A, xs, Ad, b = consistent_system(2000, 100, 12, dev)
    bn = float(torch.linalg.vector_norm(b.double()))
    solver = rla.Kaczmarz(
        compressor=rla.Sampling(compression_dim=32, distribution=rla.Uniform(replace=True)),
        log=rla.BasicLogger(max_it=4000, threshold=1e-6 * bn), error=rla.FullResidual(every=25))
    x = torch.zeros(100, dtype=F32, device=dev)
    rla.seed(3)
    x, rec = rla.rsolve(solver, x, Ad, b)
    assert rec.log.converged and rec.log.iteration < 4000
    assert np.linalg.norm(x.cpu().numpy().astype(np.float64) - xs) <= 1e-5 * np.linalg.norm(xs)


def test_kaczmarz_recipe_sees_changed_and_reallocated_matrix(rla, dev):
    # the recipe caches a row-major copy of A and the L2Norm weights; an A
    # changed in place, or a new A of the same shape, must not reuse them
    m, n = 1500, 80
    solver = rla.Kaczmarz(compressor=rla.Sampling(compression_dim=16, distribution=rla.L2Norm()),
                          log=rla.BasicLogger(max_it=3000, threshold=0.0),
                          error=rla.FullResidual(every=50))
    A1, xs1, Ad, b1 = consistent_system(m, n, 40, dev)
    x = torch.zeros(n, dtype=F32, device=dev)
    rla.seed(5)
    rec = rla.complete_solver(solver, x, Ad, b1)
    rec.log.threshold_info = 1e-6 * float(torch.linalg.vector_norm(b1.double()))
    rla.rsolve(rec, x, Ad, b1)
    assert np.linalg.norm(x.cpu().numpy() - xs1) <= 1e-4 * np.linalg.norm(xs1)
    # in place: A <- 2 A with the same b, so the solution halves
    Ad.mul_(2.0)
    x2 = torch.zeros(n, dtype=F32, device=dev)
    rla.rsolve(rec, x2, Ad, b1)
    assert np.linalg.norm(x2.cpu().numpy() - xs1 / 2) <= 1e-4 * np.linalg.norm(xs1)
    # reallocated: drop A and make another matrix of the same shape
    del Ad
    torch.cuda.synchronize()
    A3, xs3, Ad3, b3 = consistent_system(m, n, 41, dev)
    x3 = torch.zeros(n, dtype=F32, device=dev)
    rec.log.threshold_info = 1e-6 * float(torch.linalg.vector_norm(b3.double()))
    rla.rsolve(rec, x3, Ad3, b3)
    assert np.linalg.norm(x3.cpu().numpy() - xs3) <= 1e-4 * np.linalg.norm(xs3)


def test_kaczmarz_device_loop_stops_at_max_it(rla, dev):
    A, xs, Ad, b = consistent_system(500, 50, 20, dev)
    solver = rla.Kaczmarz(compressor=rla.Sampling(compression_dim=8, distribution=rla.L2Norm()),
                          log=rla.BasicLogger(max_it=7, threshold=0.0),
                          error=rla.FullResidual(every=3))
    x = torch.zeros(50, dtype=F32, device=dev)
    x, rec = rla.rsolve(solver, x, Ad, b)
    # max_it - 1 projections, then the logger stops at iteration max_it
    assert rec.log.iteration == 7 and rec.log.converged
    # errors were evaluated at 1, 4, 7
    assert np.isclose(rec.log.error, R.full_residual(A, x.cpu().numpy(), b.cpu().numpy()),
                      rtol=1e-9)
    # max_it = 0: nothing runs
    solver.log = rla.BasicLogger(max_it=0)
    x0 = torch.zeros(50, dtype=F32, device=dev)
    rla.rsolve(solver, x0, Ad, b)
    assert not x0.any()


def test_kaczmarz_generic_path_sparse_sign(rla, dev):
    A, xs, Ad, b = consistent_system(400, 30, 30, dev)
    bn = float(torch.linalg.vector_norm(b.double()))
    solver = rla.Kaczmarz(compressor=rla.SparseSign(cardinality=rla.Left(), compression_dim=12,
                                                    type=torch.float32),
                          log=rla.BasicLogger(max_it=2000, threshold=1e-5 * bn))
    x = torch.zeros(30, dtype=F32, device=dev)
    x, rec = rla.rsolve(solver, x, Ad, b)
    assert rec.log.converged and rec.log.iteration < 2000
    assert np.linalg.norm(x.cpu().numpy() - xs) <= 1e-4 * np.linalg.norm(xs)


def test_kaczmarz_dimension_errors(rla, dev):
    A, xs, Ad, b = consistent_system(50, 10, 1, dev)
    solver = rla.Kaczmarz(compressor=rla.Sampling(compression_dim=4),
                          log=rla.BasicLogger(max_it=5))
    rec = rla.complete_solver(solver, torch.zeros(10, dtype=F32, device=dev), Ad, b)
    with pytest.raises(rla._lib.DimensionMismatch, match="Dimension of `x`, 9"):
        rla.rsolve(rec, torch.zeros(9, dtype=F32, device=dev), Ad, b)
    with pytest.raises(rla._lib.DimensionMismatch, match="Dimension of `b`, 49"):
        rla.rsolve(rec, torch.zeros(10, dtype=F32, device=dev), Ad, b[:49])


def test_kaczmarz_c5_system_reaches_1e6(rla, dev):
    # the served Kaczmarz system: 100000 x 1000 N(0,1), b = A x*, 64-row blocks
    # drawn by squared row norm, FP64 residual every 100 iterations
    g = torch.Generator(device=dev).manual_seed(1234)
    Ad = torch.randn(1000, 100000, generator=g, device=dev, dtype=F32).T
    xs = torch.randn(1000, generator=g, device=dev, dtype=F32)
    b = Ad @ xs
    bn = float(torch.linalg.vector_norm(b.double()))
    solver = rla.Kaczmarz(compressor=rla.Sampling(compression_dim=64, distribution=rla.L2Norm()),
                          log=rla.BasicLogger(max_it=20000, threshold=1e-6 * bn),
                          error=rla.FullResidual(every=100))
    x = torch.zeros(1000, dtype=F32, device=dev)
    x, rec = rla.rsolve(solver, x, Ad, b)
    assert rec.log.converged and rec.log.iteration < 20000
    assert rec.log.error / bn <= 1e-6
    assert float(torch.linalg.vector_norm((x - xs).double())) <= 1e-4 * float(
        torch.linalg.vector_norm(xs.double()))


# ---------------------------------------------------------------------------
# any block size, Float64 data: the QR-based LQ projection


def lq_update(rla, dev, At, bt, x, alpha, dtype):
    s, n = At.shape
    nb = rla._dense.query_bytes("rla_kaczmarz_lq_workspace_bytes", s, n)
    work = rla._dense.workspace(nb, dev)
    Atd, btd, xd = colmajor(At, dev, dtype), colmajor(bt, dev, dtype), colmajor(x, dev, dtype)
    rla._lib.call(f"rla_kaczmarz_lq_update_{rla._lib.sfx(dtype)}", rla._lib.ctx(dev), s, n,
                  rla._lib.ptr(Atd), 1, s, rla._lib.ptr(btd), 1, rla._lib.ptr(xd), alpha,
                  rla._lib.ptr(work), work.numel() * 8)
    return xd.double().cpu().numpy()


@pytest.mark.parametrize("s,n,alpha", [(1, 50, 1.0), (8, 100, 1.0), (64, 1000, 0.7),
                                       (100, 300, 1.0), (300, 300, 1.25), (257, 2000, 1.0)])
def test_lq_update_f64_matches_oracle(rla, dev, s, n, alpha):
    # kaczmarz_update_block! with the LQ sub-solver (qr!(At') \ r) in FP64:
    # within 1e-12 of the oracle's restatement, for any block size
    At = rnd((s, n), s + n)
    x = rnd(n, n)
    bt = rnd(s, 2 * s)
    got = lq_update(rla, dev, At, bt, x, alpha, torch.float64)
    want = R.kaczmarz_update_block(x, At, bt, alpha)
    assert np.linalg.norm(got - want) <= 1e-12 * max(1.0, np.linalg.norm(want))
    if s == 1:
        assert np.allclose(got, R.kaczmarz_update(x, At[0], bt, alpha), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("s,n", [(100, 400), (65, 130)])
def test_lq_update_f32_matches_oracle(rla, dev, s, n):
    At = rnd((s, n), 3).astype(np.float32)
    x = rnd(n, 4).astype(np.float32)
    bt = rnd(s, 5).astype(np.float32)
    got = lq_update(rla, dev, At, bt, x, 1.0, F32)
    want = R.kaczmarz_update_block(x, At, bt, 1.0)
    assert np.linalg.norm(got - want) <= 1e-5 * max(1.0, np.linalg.norm(want))


def test_lq_update_zero_block_and_dependent_rows(rla, dev):
    x = rnd(40, 1)
    got = lq_update(rla, dev, np.zeros((30, 40)), np.ones(30), x, 1.0, torch.float64)
    assert np.array_equal(got, x)
    x = rnd(100, 1)
    At = rnd((70, 100), 2)
    At[50] = At[3]                      # a row drawn twice
    At[60] = 0.5 * At[1] - 2.0 * At[7]  # a combination of two others
    bt = rnd(70, 3)
    bt[50], bt[60] = bt[3], 0.5 * bt[1] - 2.0 * bt[7]
    got = lq_update(rla, dev, At, bt, x, 1.0, torch.float64)
    want = x + np.linalg.pinv(At) @ (bt - At @ x)
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)


def test_residual_norm_f64(rla, dev):
    A = rnd((3000, 77), 6)
    x, b = rnd(77, 7), rnd(3000, 8)
    out = ctypes.c_double()
    Ad, xd, bd = colmajor(A, dev, torch.float64), colmajor(x, dev, torch.float64), \
        colmajor(b, dev, torch.float64)
    rla._lib.call("rla_residual_norm_f64", rla._lib.ctx(dev), 3000, 77, rla._lib.ptr(Ad), 1, 3000,
                  rla._lib.ptr(xd), rla._lib.ptr(bd), ctypes.byref(out))
    assert abs(out.value - R.full_residual(A, x, b)) <= 1e-12 * R.full_residual(A, x, b)


@pytest.mark.parametrize("dtype,block,dist", [(torch.float64, 32, "l2norm"),
                                              (torch.float64, 100, "uniform"),
                                              (F32, 128, "l2norm")])
def test_kaczmarz_float64_and_wide_blocks_converge(rla, dev, dtype, block, dist):
    # Float64 data, and blocks over 64 rows, run the LQ projection per block
    m, n = 2000, 150
    A = rnd((m, n), 31)
    xs = rnd(n, 32)
    Ad = colmajor(A, dev, dtype)
    b = Ad @ colmajor(xs, dev, dtype)
    bn = float(torch.linalg.vector_norm(b.double()))
    D = rla.L2Norm() if dist == "l2norm" else rla.Uniform()
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    solver = rla.Kaczmarz(compressor=rla.Sampling(compression_dim=block, distribution=D),
                          log=rla.BasicLogger(max_it=3000, threshold=tol * bn),
                          error=rla.FullResidual(every=10))
    x = torch.zeros(n, dtype=dtype, device=dev)
    rla.seed(2)
    x, rec = rla.rsolve(solver, x, Ad, b)
    assert rec.log.converged and rec.log.iteration < 3000
    err = np.linalg.norm(x.double().cpu().numpy() - xs) / np.linalg.norm(xs)
    assert err <= (1e-8 if dtype == torch.float64 else 1e-4), err


def test_kaczmarz_float64_sparse_sign(rla, dev):
    m, n = 600, 40
    A = rnd((m, n), 41)
    xs = rnd(n, 42)
    Ad = colmajor(A, dev, torch.float64)
    b = Ad @ colmajor(xs, dev, torch.float64)
    bn = float(torch.linalg.vector_norm(b))
    solver = rla.Kaczmarz(compressor=rla.SparseSign(cardinality=rla.Left(), compression_dim=20),
                          log=rla.BasicLogger(max_it=4000, threshold=1e-10 * bn))
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    x, rec = rla.rsolve(solver, x, Ad, b)
    assert rec.log.converged
    assert np.linalg.norm(x.cpu().numpy() - xs) <= 1e-8 * np.linalg.norm(xs)
