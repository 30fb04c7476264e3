"""GPU parity: Gaussian compressor, BF16 MFMA GEMM, FP64 CholeskyQR2,
RangeFinder and RandSVD against the oracle and closed forms."""
import numpy as np
import pytest
import torch

from oracle import restate as R

pytestmark = pytest.mark.gpu
F64, F32 = torch.float64, torch.float32


def colmajor(a, dev, dtype=None):
    t = torch.as_tensor(np.asarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.T.contiguous().T.to(dev) if t.dim() == 2 else t.to(dev)


def rnd(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("dtype,tol", [(F64, 1e-12), (F32, 1e-5)])
@pytest.mark.parametrize("card", ["left", "right"])
def test_gaussian_all_orientations(rla, dev, dtype, tol, card):
    # test/Compressors/gaussian.jl:216-280: mul! against S.op * A with the recipe's op
    rla.seed(5)
    cardinality = rla.Left() if card == "left" else rla.Right()
    A = rnd((50, 40), 1)
    S = rla.complete_compressor(rla.Gaussian(cardinality=cardinality, compression_dim=9,
                                             type=dtype), colmajor(A, dev, dtype))
    op = S.op.double().cpu().numpy()
    assert op.shape == S.size()
    r, c = op.shape
    # column scale ~ 1/sqrt(s): variance of the entries
    assert abs(np.var(op) * 9 - 1) < 0.3
    X = rnd((c, 4), 2)
    C = torch.zeros(r, 4, dtype=dtype, device=dev)
    rla.mul(C, S, colmajor(X, dev, dtype), 2.0, 0.0)
    assert np.max(np.abs(C.double().cpu().numpy() - 2 * op @ X)) < tol * np.max(np.abs(op @ X)) * 4
    Y = rnd((3, r), 3)
    C2 = torch.zeros(3, c, dtype=dtype, device=dev)
    rla.mul(C2, colmajor(Y, dev, dtype), S)
    assert np.max(np.abs(C2.double().cpu().numpy() - Y @ op)) < tol * np.max(np.abs(Y @ op)) * 4
    C3 = torch.zeros(c, 2, dtype=dtype, device=dev)
    Z = rnd((r, 2), 4)
    rla.mul(C3, S.T, colmajor(Z, dev, dtype))
    assert np.max(np.abs(C3.double().cpu().numpy() - op.T @ Z)) < tol * np.max(np.abs(op.T @ Z)) * 4


@pytest.mark.parametrize("dtype,tol", [(F64, 1e-12), (F32, 1e-5)])
@pytest.mark.parametrize("card", ["left", "right"])
def test_gaussian_all_orientations_c3_size(rla, dev, dtype, tol, card):
    # the Gaussian apply at the C3 sketch shape (8192 x 266 operator) on the
    # MFMA GEMM, all four orientations, against FP64 numpy products
    rla.seed(6)
    cardinality = rla.Left() if card == "left" else rla.Right()
    m, n, s = (8192, 300, 266) if card == "left" else (700, 8192, 266)
    A = colmajor(rnd((m, n), 1), dev, dtype)
    S = rla.complete_compressor(rla.Gaussian(cardinality=cardinality, compression_dim=s,
                                             type=dtype), A)
    op = S.op.double().cpu().numpy()
    r, c = op.shape

    def check(got, want):
        err = np.linalg.norm(got.double().cpu().numpy() - want) / np.linalg.norm(want)
        assert err <= tol, err

    X = rnd((c, 33), 2)
    C = torch.full((r, 33), 3.0, dtype=dtype, device=dev)
    rla.mul(C, S, colmajor(X, dev, dtype), 1.5, 0.5)
    check(C, 1.5 * op @ X + 1.5)
    Y = rnd((31, r), 3)
    C2 = torch.zeros(31, c, dtype=dtype, device=dev)
    rla.mul(C2, colmajor(Y, dev, dtype), S)
    check(C2, Y @ op)
    Z = rnd((r, 5), 4)
    C3 = torch.zeros(c, 5, dtype=dtype, device=dev)
    rla.mul(C3, S.T, colmajor(Z, dev, dtype))
    check(C3, op.T @ Z)
    W = rnd((7, c), 5)
    C4 = torch.zeros(7, r, dtype=dtype, device=dev)
    rla.mul(C4, colmajor(W, dev, dtype), S.T)
    check(C4, W @ op.T)
    # a row-major A and a row-major C (Julia adjoint views) give the same result
    Xr = torch.as_tensor(X).to(dtype).to(dev).contiguous()
    Cr = torch.zeros(r, 33, dtype=dtype, device=dev)
    rla.mul(Cr, S, Xr)
    check(Cr, op @ X)


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (64, 64, 16), (65, 63, 17), (300, 7, 20000),
                                   (5, 3000, 2), (1000, 266, 8192), (129, 1, 4096)])
@pytest.mark.parametrize("dtype,tol", [(F64, 1e-13), (F32, 2e-6)])
@pytest.mark.parametrize("layout", ["nn", "tn", "nt", "tt"])
def test_gemm_strided_matches_numpy(rla, dev, M, N, K, dtype, tol, layout):
    # C = alpha A B + beta C through rla_gemm_strided_*: transposed operands
    # are stride swaps; long K splits over workgroups (deterministic sums)
    A = rnd((M, K), M + K)
    B = rnd((K, N), N)
    C0 = rnd((M, N), 7)
    At = colmajor(A, dev, dtype) if layout[0] == "n" else colmajor(A.T, dev, dtype).T
    Bt = colmajor(B, dev, dtype) if layout[1] == "n" else colmajor(B.T, dev, dtype).T
    for beta in (0.0, -0.5):
        C = colmajor(C0, dev, dtype)
        rla._dense.gemm(rla.as_view(C), rla.as_view(At), rla.as_view(Bt), 1.25, beta)
        want = 1.25 * A @ B + beta * C0
        err = np.abs(C.double().cpu().numpy() - want).max() / max(1.0, np.abs(want).max())
        assert err <= tol * max(1.0, np.sqrt(K)), err
        # same inputs, same bits
        C2 = colmajor(C0, dev, dtype)
        rla._dense.gemm(rla.as_view(C2), rla.as_view(At), rla.as_view(Bt), 1.25, beta)
        assert torch.equal(C, C2)


def test_gaussian_bf16_apply(rla, dev):
    # precision="bf16": operands rounded to BF16, FP32 accumulation; equals the
    # FP64 product of the BF16-rounded operands to FP32 accumulation error
    rla.seed(8)
    A = colmajor(rnd((4000, 300), 1), dev, F32)
    for card in (rla.Left(), rla.Right()):
        S = rla.complete_compressor(rla.Gaussian(cardinality=card, compression_dim=266,
                                                 type=F32, precision="bf16"), A)
        opb = S.op.to(torch.bfloat16).double().cpu().numpy()
        r, c = opb.shape
        X = torch.as_tensor(rnd((c, 40), 2)).float()
        C = torch.zeros(r, 40, dtype=F32, device=dev)
        rla.mul(C, S, X.to(dev))
        want = opb @ X.to(torch.bfloat16).double().numpy()
        err = np.linalg.norm(C.double().cpu().numpy() - want) / np.linalg.norm(want)
        assert err < 1e-5, err
        Y = torch.as_tensor(rnd((25, r), 3)).float()
        C2 = torch.zeros(25, c, dtype=F32, device=dev)
        rla.mul(C2, Y.to(dev), S)
        want2 = Y.to(torch.bfloat16).double().numpy() @ opb
        err2 = np.linalg.norm(C2.double().cpu().numpy() - want2) / np.linalg.norm(want2)
        assert err2 < 1e-5, err2
        # and it is the bf16 product, not the plain one (BF16 rounding is visible)
        plain = S.op.double().cpu().numpy() @ X.double().numpy()
        assert np.linalg.norm(C.double().cpu().numpy() - plain) / np.linalg.norm(plain) > 1e-4
        # update_compressor! regenerates op; the cached BF16 copy follows it
        rla.update_compressor(S)
        opb2 = S.op.to(torch.bfloat16).double().cpu().numpy()
        rla.mul(C, S, X.to(dev))
        want3 = opb2 @ X.to(torch.bfloat16).double().numpy()
        assert np.linalg.norm(C.double().cpu().numpy() - want3) < 1e-5 * np.linalg.norm(want3)
    with pytest.raises(rla._lib.ArgumentError):
        rla.Gaussian(precision="fp8")


@pytest.mark.parametrize("M,N,K", [(256, 266, 8192), (100, 37, 300), (8192, 266, 4096),
                                   (33, 300, 1000), (128, 64, 64), (300, 266, 1000),
                                   (1000, 288, 8200), (129, 100, 64), (5000, 33, 130),
                                   (256, 266, 32768), (70, 200, 40), (4096, 160, 5000)])
def test_gemm_bf16_matches_fp32_reference(rla, dev, M, N, K):
    # inputs rounded to bf16 on the host; products exact, FP32 accumulation
    A = rnd((M, K), 1).astype(np.float32)
    B = rnd((K, N), 2).astype(np.float32)
    Ab = torch.as_tensor(A).to(torch.bfloat16).float().numpy().astype(np.float64)
    Bb = torch.as_tensor(B).to(torch.bfloat16).float().numpy().astype(np.float64)
    want = Ab @ Bb
    Ad = colmajor(A, dev)
    Bd = colmajor(B, dev)
    rows = rla._dense.bf16_rows(rla.as_view(Ad))        # M x ld8(K)
    cols = rla._dense.bf16_rows(rla.as_view(Bd).T)      # N x ld8(K)
    C = torch.zeros(N, M, dtype=F32, device=dev).T
    rla._dense.gemm_bf16(M, N, K, rows, cols, rla.as_view(C))
    got = C.double().cpu().numpy()
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert err < 2e-6, err


def _tiled_host(Ab):
    """The tiled BF16 layout from its definition: 128 x 64 tiles, tile (t, s)
    at (t * nkt + s) * 8192, the 8 entries of row 32 g + r and columns
    16 q + 8 h .. + 8 at ((4 g + q) * 64 + r + 32 h) * 8; zero padding."""
    rows, cols = Ab.shape
    nmt, nkt = -(-rows // 128), -(-cols // 64)
    out = np.zeros(nmt * nkt * 8192, Ab.dtype)
    i, l = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    r, c = i % 128, l % 64
    idx = ((i // 128) * nkt + l // 64) * 8192 + ((4 * (r // 32) + c // 16) * 64 + r % 32
                                                  + 32 * ((c // 8) % 2)) * 8 + c % 8
    out[idx] = Ab
    return out


@pytest.mark.parametrize("dtype", [F32, F64])
def test_bf16_tiled_copy_layout(rla, dev, dtype):
    # a strided 130 x 70 view (two row tiles, two K tiles, both partial)
    rows, cols = 130, 70
    X = rnd((rows + 5, cols), 3)
    Xd = colmajor(X, dev, dtype)[:rows]
    got = rla._dense.bf16_tiled(rla.as_view(Xd)).view(torch.int16).cpu().numpy()
    # the same rounding as the row copies (through float, nearest even)
    Xb = torch.as_tensor(X[:rows]).to(dtype).float().to(torch.bfloat16).view(torch.int16).numpy()
    assert np.array_equal(got, _tiled_host(Xb))


@pytest.mark.parametrize("M,N,K", [(1000, 266, 3000), (130, 40, 70), (1000, 266, 40003),
                                   (8192, 266, 32768), (300, 288, 1000)])
def test_gemm_bf16_tiled_equals_row_layout(rla, dev, M, N, K):
    # the same products in the same order whatever A's layout: identical bits,
    # with and without split K, partial K steps and partial row tiles
    A = rnd((M, K), 1).astype(np.float32)
    B = rnd((K, N), 2).astype(np.float32)
    Ad, Bd = colmajor(A, dev), colmajor(B, dev)
    cols = rla._dense.bf16_rows(rla.as_view(Bd).T)
    C1 = torch.zeros(N, M, dtype=F32, device=dev).T
    rla._dense.gemm_bf16(M, N, K, rla._dense.bf16_rows(rla.as_view(Ad)), cols, rla.as_view(C1))
    C2 = torch.zeros(N, M, dtype=F32, device=dev).T
    rla._dense.gemm_bf16_tiled(M, N, K, rla._dense.bf16_tiled(rla.as_view(Ad)), cols,
                               rla.as_view(C2))
    assert torch.equal(C1, C2)


def test_rangefinder_tiled_copies_give_identical_q(rla, dev):
    # rla_rangefinder with the row BF16 copies (precision 1) and the tiled
    # ones (precision 2) on the same sketch: the same Q, bit for bit
    m, n, k = 3000, 700, 40
    A = R._q(rnd((m, 60), 4)) @ np.diag(np.exp(-np.arange(60) / 8.0)) @ rnd((60, n), 5)
    Ad = colmajor(A, dev, F32)
    v = rla.as_view(Ad)
    Y0 = colmajor(rnd((m, k), 6), dev, F64)
    lib = rla._lib
    nb = rla._dense.query_bytes("rla_rangefinder_workspace_bytes", 1, 4, m, n, k)
    work = rla._dense.workspace(nb, dev)
    Qs = []
    for prec, rm, cm, ld_rm, ld_cm in (
            (1, rla._dense.bf16_rows(v), rla._dense.bf16_rows(v.T), None, None),
            (2, rla._dense.bf16_tiled(v), rla._dense.bf16_tiled(v.T), 0, 0)):
        Y = Y0.clone()
        lib.call("rla_rangefinder", lib.ctx(dev), prec, 1, m, n, k, 2, 4, lib.ptr(Ad), m,
                 lib.ptr(rm), rm.shape[1] if ld_rm is None else ld_rm, lib.ptr(cm),
                 cm.shape[1] if ld_cm is None else ld_cm, lib.ptr(Y), m, lib.ptr(work),
                 work.numel() * 8)
        Qs.append(Y)
    assert torch.equal(Qs[0], Qs[1])
    Q = Qs[1].cpu().numpy()
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 1e-10


@pytest.mark.parametrize("ta,tb,M,N,K", [(0, 0, 64, 64, 16), (1, 0, 64, 64, 16),
                                         (1, 0, 266, 266, 5000), (0, 0, 5000, 266, 266),
                                         (0, 1, 70, 33, 100), (1, 1, 17, 65, 1030)])
def test_gemm_f64_mfma(rla, dev, ta, tb, M, N, K):
    A = rnd((K, M) if ta else (M, K), 1)
    B = rnd((N, K) if tb else (K, N), 2)
    C0 = rnd((M, N), 3)
    want = 1.5 * (A.T if ta else A) @ (B.T if tb else B) + 0.5 * C0
    Ad, Bd, Cd = colmajor(A, dev), colmajor(B, dev), colmajor(C0, dev)
    rla._lib.call("rla_gemm_f64_mfma", rla._lib.ctx(dev), ta, tb, M, N, K, 1.5,
                  rla._lib.ptr(Ad), Ad.stride(1), rla._lib.ptr(Bd), Bd.stride(1), 0.5,
                  rla._lib.ptr(Cd), Cd.stride(1))
    got = Cd.cpu().numpy()
    err = np.abs(got - want)
    assert err.max() <= 1e-11 * np.abs(want).max()


def test_bf16_conversion_rounds_to_nearest_even(rla, dev):
    x = torch.tensor([[1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-30]], dtype=F32)
    got = rla._dense.bf16_rows(rla.as_view(x.to(dev)))[:, :6].cpu().float()
    want = x.to(torch.bfloat16).float()
    assert torch.equal(got, want)


@pytest.mark.parametrize("m,k,ld", [(5000, 266, 5000), (300, 40, 300), (64, 64, 64),
                                    (33000, 17, 33000), (2000, 288, 2000), (2000, 289, 2000),
                                    (1000, 1, 1000), (4100, 130, 4100), (70, 65, 70),
                                    (500, 100, 512), (8192, 272, 8192), (3000, 200, 3001)])
def test_qr_f64_cholqr2(rla, dev, m, k, ld):
    # k <= 288: one-read Gram, one-workgroup Cholesky/inverse, in-place Y R^-1;
    # k = 289: tiled GEMM path.  ld > m checks the leading dimension.
    Y = rnd((m, k), m) @ np.diag(np.logspace(0, -4, k))  # cond 1e4
    Yp = np.zeros((ld, k))
    Yp[:m] = Y
    Yfull = colmajor(Yp, dev)
    Yd = Yfull[:m]
    Rd = torch.zeros(k, k, dtype=F64, device=dev).T
    nb = rla._dense.query_bytes("rla_qr_workspace_bytes", m, k)
    work = rla._dense.workspace(nb, dev)
    import ctypes
    hh = ctypes.c_int(-1)
    rla._lib.call("rla_qr_f64", rla._lib.ctx(dev), m, k, rla._lib.ptr(Yd), ld, rla._lib.ptr(Rd),
                  k, rla._lib.ptr(work), work.numel() * 8, ctypes.byref(hh))
    Q, Rm = Yd.cpu().numpy(), Rd.cpu().numpy()
    assert hh.value == 0
    assert np.all(Yfull[m:].cpu().numpy() == 0)  # rows past m untouched
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 1e-12 * k
    assert np.linalg.norm(Q @ Rm - Y) < 1e-12 * np.linalg.norm(Y)
    assert np.allclose(np.triu(Rm), Rm) and np.all(np.diag(Rm) > 0)
    # same Q as a positive-diagonal Householder QR
    assert np.linalg.norm(Q - R._q(Y)) < 1e-8


def _chol_inv(rla, dev, G):
    import ctypes
    k = G.shape[0]
    Gd = colmajor(G, dev)
    Rd = torch.full((k, k), np.nan, dtype=F64, device=dev).T
    Xd = torch.full((k, k), np.nan, dtype=F64, device=dev).T
    info = ctypes.c_int(-1)
    rla._lib.call("rla_chol_inv_f64", rla._lib.ctx(dev), k, rla._lib.ptr(Gd), k,
                  rla._lib.ptr(Rd), k, rla._lib.ptr(Xd), k, ctypes.byref(info))
    return Rd.cpu().numpy(), Xd.cpu().numpy(), info.value


@pytest.mark.parametrize("k", [1, 2, 15, 16, 17, 31, 33, 100, 200, 255, 266, 272, 288])
def test_chol_inv_matches_lapack(rla, dev, k):
    B = rnd((k + 8, k), k) @ np.diag(np.logspace(0, -2, k))
    G = B.T @ B
    Rm, Xm, info = _chol_inv(rla, dev, G)
    assert info == 0
    Rref = np.linalg.cholesky(G).T            # upper, positive diagonal
    assert np.array_equal(np.triu(Rm), Rm) and np.array_equal(np.triu(Xm), Xm)
    assert np.linalg.norm(Rm - Rref) < 1e-12 * np.linalg.norm(Rref) * max(1, np.sqrt(k))
    # X R = I to the conditioning of R (cond(R) <= 1e2 * cond(B)-ish)
    assert np.linalg.norm(Xm @ Rm - np.eye(k)) < 1e-11 * np.sqrt(k)
    assert np.linalg.norm(Xm - np.linalg.inv(Rref)) < 1e-10 * np.linalg.norm(Xm)


def test_chol_inv_rejects_k_above_288(rla, dev):
    with pytest.raises(rla._lib.ArgumentError, match="288"):
        _chol_inv(rla, dev, np.eye(289))


@pytest.mark.parametrize("k,bad", [(5, 3), (40, 1), (40, 17), (288, 280), (266, 266)])
def test_chol_inv_reports_breakdown_column(rla, dev, k, bad):
    # G = diag(d) with d[bad-1] <= 0: the factorisation stops at column `bad`
    d = np.linspace(1.0, 2.0, k)
    d[bad - 1] = -1.0 if bad % 2 else 0.0
    _, _, info = _chol_inv(rla, dev, np.diag(d))
    assert info == bad


def test_chol_inv_flags_numerically_singular(rla, dev):
    # rank 10 of 30 in FP64: pivots at rounding level are a breakdown
    B = rnd((400, 10), 1) @ rnd((10, 30), 2)
    _, _, info = _chol_inv(rla, dev, B.T @ B)
    assert 10 < info <= 30


def test_qr_f64_householder_fallback_on_rank_deficient(rla, dev):
    m, k = 400, 30
    Y = rnd((m, 10), 1) @ rnd((10, k), 2)        # rank 10 < k: Cholesky breaks down
    Yd = colmajor(Y, dev)
    Rd = torch.zeros(k, k, dtype=F64, device=dev).T
    nb = rla._dense.query_bytes("rla_qr_workspace_bytes", m, k)
    work = rla._dense.workspace(nb, dev)
    import ctypes
    hh = ctypes.c_int(-1)
    rla._lib.call("rla_qr_f64", rla._lib.ctx(dev), m, k, rla._lib.ptr(Yd), m, rla._lib.ptr(Rd),
                  k, rla._lib.ptr(work), work.numel() * 8, ctypes.byref(hh))
    Q, Rm = Yd.cpu().numpy(), Rd.cpu().numpy()
    assert hh.value == 1
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 1e-12 * k
    assert np.linalg.norm(Q @ Rm - Y) < 1e-12 * np.linalg.norm(Y)
    assert np.all(np.diag(Rm) >= 0)


def _spectrum_matrix(m, n, decay, seed):
    rng = np.random.default_rng(seed)
    r = min(m, n)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    s = np.exp(-np.arange(r) / decay)
    return (U * s) @ V.T, s


@pytest.mark.parametrize("ortho", [True, False])
@pytest.mark.parametrize("q", [0, 2])
@pytest.mark.parametrize("dtype,tol", [(F64, 1e-9), (F32, 2e-4)])
def test_rangefinder_plain_matches_oracle(rla, dev, ortho, q, dtype, tol):
    A, _ = _spectrum_matrix(300, 200, 20.0, 3)
    Ad = colmajor(A, dev, dtype)
    rla.seed(11)
    rf = rla.RangeFinder(compressor=rla.Gaussian(cardinality=rla.Right(), compression_dim=25,
                                                 type=dtype),
                         power_its=q, orthogonalize=ortho)
    rec = rla.rapproximate(rf, Ad)
    Om = rec.compressor.op.double().cpu().numpy()
    want = (R.rand_ortho_it if ortho else R.rand_power_it)(A, Om, q)
    Q = rec.range.double().cpu().numpy()
    assert rec.range.dtype == dtype and Q.shape == (300, 25)
    assert np.linalg.norm(Q.T @ Q - np.eye(25)) < (1e-12 if dtype == F64 else 1e-5) * 25
    assert np.linalg.norm(Q @ Q.T - want @ want.T) < tol * 25


@pytest.mark.parametrize("dtype,tol", [(F64, 1e-12), (F32, 1e-5)])
def test_rangefinder_plain_row_copy_matches_transposed_read(rla, dev, dtype, tol):
    # the plain path's A'X through the rows-contiguous copy (an NN product) and
    # through a transposed read of A give the same range; the recipe keeps the
    # copy of a column-major A and uses a row-major A as its own copy
    import ctypes
    A, _ = _spectrum_matrix(500, 300, 20.0, 5)
    m, n, k = 500, 300, 24
    Ad = colmajor(A, dev, dtype)
    Om = torch.tensor(rnd((n, k), 8), dtype=F64, device=dev)
    Y0 = (Ad.double() @ Om).T.contiguous().T
    a_elem = 4 if dtype == F32 else 8
    nb = ctypes.c_size_t(0)
    rla._lib.call("rla_rangefinder_workspace_bytes", 0, a_elem, m, n, k, ctypes.byref(nb))
    work = torch.empty(nb.value // 8 + 1, dtype=F64, device=dev)
    rows = Ad.contiguous()
    out = {}
    for name, copy in (("transposed", None), ("copy", rows)):
        Y = Y0.clone().T.contiguous().T
        rla._lib.call("rla_rangefinder", rla._lib.ctx(dev), 0, 1, m, n, k, 2, a_elem,
                      rla._lib.ptr(Ad), m, rla._lib.ptr(copy), n if copy is not None else 0,
                      None, 0, rla._lib.ptr(Y), m, rla._lib.ptr(work), work.numel() * 8)
        out[name] = Y.cpu().numpy()
    P0 = out["transposed"] @ out["transposed"].T
    P1 = out["copy"] @ out["copy"].T
    assert np.linalg.norm(P0 - P1) < tol * k
    with pytest.raises(rla.ArgumentError):
        rla._lib.call("rla_rangefinder", rla._lib.ctx(dev), 0, 1, m, n, k, 2, a_elem,
                      rla._lib.ptr(Ad), m, rla._lib.ptr(rows), n - 1, None, 0,
                      rla._lib.ptr(Y0), m, rla._lib.ptr(work), work.numel() * 8)
    rf = rla.RangeFinder(compressor=rla.Gaussian(cardinality=rla.Right(), compression_dim=k,
                                                 type=dtype), power_its=2, orthogonalize=True)
    rec = rla.complete_approximator(rf, Ad)
    rla.rapproximate_(rec, Ad)
    assert rec._rows.rows is not None and rec._rows.rows.is_contiguous()
    Ar = Ad.contiguous()                   # row-major: its own row copy
    rec2 = rla.complete_approximator(rf, Ar)
    rec2.compressor = rec.compressor
    rla.rapproximate_(rec2, Ar)
    assert rec2._rows.rows is None
    Q, Q2 = rec.range.double().cpu().numpy(), rec2.range.double().cpu().numpy()
    assert np.linalg.norm(Q @ Q.T - Q2 @ Q2.T) < tol * k


def test_rangefinder_exact_recovery_sparse_sign(rla, dev):
    # default compressor SparseSign(Right()), rank 6 < k: A = QQ'A to 1e-10
    A = R._q(rnd((400, 6), 1)) @ rnd((6, 150), 2)
    Ad = colmajor(A, dev)
    rec = rla.rapproximate(rla.RangeFinder(compressor=rla.SparseSign(
        cardinality=rla.Right(), compression_dim=12)), Ad)
    Q = rec.range.cpu().numpy()
    assert np.linalg.norm(A - Q @ (Q.T @ A)) < 1e-10 * np.linalg.norm(A)
    C = torch.zeros(400, 150, dtype=F64, device=dev)
    B = torch.zeros(12, 150, dtype=F64, device=dev)
    rla.mul(B, rec.T, Ad)       # Q'A
    rla.mul(C, rec, B)          # Q Q'A
    assert np.linalg.norm(C.cpu().numpy() - A) < 1e-10 * np.linalg.norm(A)


def test_rangefinder_bf16_error_close_to_plain(rla, dev):
    # the C3 recipe at a reduced size: sigma_i = exp(-i/64), k = 266, q = 2
    A, s = _spectrum_matrix(4096, 1024, 64.0, 7)
    Ad = colmajor(A, dev)
    opt = np.sqrt(np.sum(s[266:] ** 2) / np.sum(s ** 2))
    errs = {}
    for prec in ("plain", "bf16"):
        rla.seed(3)
        rf = rla.RangeFinder(compressor=rla.Gaussian(cardinality=rla.Right(),
                                                     compression_dim=266),
                             power_its=2, orthogonalize=True, precision=prec)
        Q = rla.rapproximate(rf, Ad).range.cpu().numpy()
        errs[prec] = np.linalg.norm(A - Q @ (Q.T @ A)) / np.linalg.norm(A)
    assert errs["plain"] < 1.05 * opt
    assert errs["bf16"] <= 1.1 * errs["plain"], errs


def test_rangefinder_bf16_recipe_sees_changed_and_reallocated_matrix(rla, dev):
    # the bf16 copies of A are cached per recipe; a row-major A (a fresh
    # column-major temporary would be made from it on every call), changed in
    # place or replaced by another matrix of the same shape, must not reuse them
    def rowmajor(a):
        return torch.as_tensor(a).to(dev).contiguous()

    def proj_err(Q, A):
        return np.linalg.norm(A - Q @ (Q.T @ A)) / np.linalg.norm(A)

    rf = rla.RangeFinder(compressor=rla.Gaussian(cardinality=rla.Right(), compression_dim=12),
                         power_its=1, orthogonalize=True, precision="bf16")
    A1 = R._q(rnd((500, 10), 1)) @ rnd((10, 200), 2)
    Ad = rowmajor(A1)
    rla.seed(4)
    rec = rla.rapproximate(rf, Ad)
    assert proj_err(rec.range.cpu().numpy(), A1) < 2e-2
    A2 = R._q(rnd((500, 10), 3)) @ rnd((10, 200), 4)
    Ad.copy_(torch.as_tensor(A2))              # in place
    rla.rapproximate_(rec, Ad)
    assert proj_err(rec.range.cpu().numpy(), A2) < 2e-2
    del Ad
    torch.cuda.synchronize()
    A3 = R._q(rnd((500, 10), 5)) @ rnd((10, 200), 6)
    Ad3 = rowmajor(A3)                         # reallocated, same shape
    rla.rapproximate_(rec, Ad3)
    assert proj_err(rec.range.cpu().numpy(), A3) < 2e-2
    # a fresh recipe with the same compressor gives the same range
    rec2 = rla.complete_approximator(rf, Ad3)
    rec2.compressor = rec.compressor
    rla.rapproximate_(rec2, Ad3)
    Q, Q2 = rec.range.cpu().numpy(), rec2.range.cpu().numpy()
    assert np.linalg.norm(Q @ Q.T - Q2 @ Q2.T) < 1e-12


@pytest.mark.parametrize("prec", ["plain", "bf16"])
@pytest.mark.parametrize("ortho", [True, False])
def test_rangefinder_rank_deficient_reruns_with_householder(rla, dev, prec, ortho):
    # rank 8 < k = 30: the device-side Cholesky breakdown is caught after the
    # run, which is redone from the saved sketch with host-checked QRs
    A = R._q(rnd((700, 8), 4)) @ rnd((8, 260), 5)
    Ad = colmajor(A, dev)
    rla.seed(9)
    rf = rla.RangeFinder(compressor=rla.Gaussian(cardinality=rla.Right(), compression_dim=30),
                         power_its=2, orthogonalize=ortho, precision=prec)
    Q = rla.rapproximate(rf, Ad).range.cpu().numpy()
    assert np.isfinite(Q).all()
    assert np.linalg.norm(Q.T @ Q - np.eye(30)) < 1e-10
    tol = 1e-10 if prec == "plain" else 5e-2   # BF16 rounding of A is full rank
    assert np.linalg.norm(A - Q @ (Q.T @ A)) < tol * np.linalg.norm(A)


@pytest.mark.parametrize("q", [1, 2])
def test_rangefinder_bf16_breakdown_reruns_from_untouched_sketch(rla, dev, q):
    # only 8 rows of A are nonzero, so even its BF16 copy has rank 8 and the
    # sketch's Gram matrix is exactly singular: the orthogonalised BF16 run
    # (which keeps no copy of the sketch) breaks down and must leave the
    # sketch intact for the host-checked rerun
    A = np.zeros((700, 260))
    A[[3, 50, 51, 200, 333, 480, 610, 699]] = rnd((8, 260), 6)
    Ad = colmajor(A, dev)
    rla.seed(11)
    rf = rla.RangeFinder(compressor=rla.Gaussian(cardinality=rla.Right(), compression_dim=30),
                         power_its=q, orthogonalize=True, precision="bf16")
    Q = rla.rapproximate(rf, Ad).range.cpu().numpy()
    assert np.isfinite(Q).all()
    assert np.linalg.norm(Q.T @ Q - np.eye(30)) < 1e-10
    # range(A) lies in span(Q) up to FP64 rounding: the zero rows stay zero
    assert np.linalg.norm(A - Q @ (Q.T @ A)) < 1e-10 * np.linalg.norm(A)


def small_svd(rla, dev, X, transpose=False):
    import ctypes
    k = X.shape[0]
    Xd = colmajor(X, dev, F64)
    U = torch.zeros(k, k, dtype=F64, device=dev).T
    V = torch.zeros(k, k, dtype=F64, device=dev).T
    S = torch.zeros(k, dtype=F64, device=dev)
    nb = rla._dense.query_bytes("rla_svd_small_workspace_bytes", k)
    work = rla._dense.workspace(nb, dev)
    sweeps = ctypes.c_int(0)
    rla._lib.call("rla_svd_small_f64", rla._lib.ctx(dev), k, rla._lib.ptr(Xd), k,
                  1 if transpose else 0, rla._lib.ptr(U), k, rla._lib.ptr(S), rla._lib.ptr(V), k,
                  rla._lib.ptr(work), work.numel() * 8, ctypes.byref(sweeps))
    return U.cpu().numpy(), S.cpu().numpy(), V.cpu().numpy(), sweeps.value


@pytest.mark.parametrize("k,decay", [(1, 1.0), (2, 1.0), (17, 4.0), (40, 8.0), (266, 64.0),
                                     (288, 20.0), (100, 1e9)])
def test_small_svd_one_sided_jacobi(rla, dev, k, decay):
    # against LAPACK (numpy) on X = P diag(s) W' with known s: singular values
    # within 1e-12 * s_max (forming X in FP64 already moves them by ~eps s_max;
    # relative 1e-12 where s >= 1e-4), U and V orthogonal, X = U S V'
    rng = np.random.default_rng(k)
    P = np.linalg.qr(rng.standard_normal((k, k)))[0]
    W = np.linalg.qr(rng.standard_normal((k, k)))[0]
    s = np.exp(-np.arange(k) / decay)
    X = (P * s) @ W.T
    for tr in (False, True):
        U, S, V, sweeps = small_svd(rla, dev, X.T if tr else X, tr)
        assert 0 < sweeps < 60
        assert np.all(np.diff(S) <= 0)
        tol = 1e-12 * max(1.0, k / 40)
        big = s >= 1e-4
        assert np.max(np.abs(S - s)) < tol
        assert np.max(np.abs(S - s)[big] / s[big]) < tol
        assert np.max(np.abs(S - np.linalg.svd(X, compute_uv=False))) < tol
        assert np.linalg.norm(U.T @ U - np.eye(k)) < 1e-12 * k
        assert np.linalg.norm(V.T @ V - np.eye(k)) < 1e-12 * k
        assert np.linalg.norm((U * S) @ V.T - X) < 1e-13 * k * np.linalg.norm(X)


def test_small_svd_triangular_and_rank_deficient(rla, dev):
    # the RandSVD case: the transposed triangular factor; and a matrix of rank
    # 5 (zero singular values come out as zeros, the rest exact)
    rng = np.random.default_rng(3)
    R = np.triu(rng.standard_normal((120, 120))) + 10 * np.eye(120)
    U, S, V, _ = small_svd(rla, dev, R, True)
    ref = np.linalg.svd(R, compute_uv=False)
    assert np.max(np.abs(S - ref) / ref) < 1e-12
    assert np.linalg.norm((U * S) @ V.T - R.T) < 1e-12 * np.linalg.norm(R)
    X = rng.standard_normal((60, 5)) @ rng.standard_normal((5, 60))
    U, S, V, _ = small_svd(rla, dev, X)
    ref = np.linalg.svd(X, compute_uv=False)
    assert np.max(np.abs(S[:5] - ref[:5]) / ref[:5]) < 1e-12
    assert np.max(np.abs(S[5:])) < 1e-12 * ref[0]
    assert np.linalg.norm(V.T @ V - np.eye(60)) < 1e-12 * 60
    assert np.linalg.norm((U * S) @ V.T - X) < 1e-12 * np.linalg.norm(X)


@pytest.mark.parametrize("prec", ["plain", "bf16"])
def test_randsvd(rla, dev, prec):
    A, s = _spectrum_matrix(600, 300, 8.0, 5)
    Ad = colmajor(A, dev)
    rla.seed(2)
    rec = rla.rapproximate(rla.RandSVD(compressor=rla.Gaussian(cardinality=rla.Right(),
                                                               compression_dim=40),
                                       power_its=2, orthogonalize=True, precision=prec), Ad)
    U, S, V = rec.U.cpu().numpy(), rec.S.cpu().numpy(), rec.V.cpu().numpy()
    assert U.shape == (600, 40) and S.shape == (40,) and V.shape == (300, 40)
    tol = 1e-10 if prec == "plain" else 5e-3
    assert np.max(np.abs(S[:20] - s[:20]) / s[:20]) < tol
    assert np.linalg.norm(U.T @ U - np.eye(40)) < 1e-10 and np.linalg.norm(V.T @ V - np.eye(40)) < 1e-10
    if prec == "plain":
        Qw = rla.rapproximate(rla.RangeFinder(compressor=rla.Gaussian(cardinality=rla.Right(),
                                                                      compression_dim=40)),
                              Ad)  # sanity of the recipe path
        assert Qw.range.shape == (600, 40)
        # oracle finish from the same Q reproduces U S V'
        rla.seed(2)
        rec2 = rla.rapproximate(rla.RandSVD(compressor=rla.Gaussian(cardinality=rla.Right(),
                                                                    compression_dim=40),
                                            power_its=2, orthogonalize=True), Ad)
        Om = rec2.compressor.op.cpu().numpy()
        Q = R.rand_ortho_it(A, Om, 2)
        Uo, So, Vo = R.randsvd(A, Q)
        assert np.max(np.abs(rec2.S.cpu().numpy() - So) / So) < 1e-9
        lowrank = (U * S) @ V.T
        assert np.linalg.norm(lowrank - (Uo * So) @ Vo.T) < 1e-8 * np.linalg.norm(lowrank)
    # low-rank mul!: C = U S V' X
    X = rnd((300, 3), 1)
    C = torch.zeros(600, 3, dtype=F64, device=dev)
    rla.mul(C, rec, colmajor(X, dev))
    assert np.linalg.norm(C.cpu().numpy() - (U * S) @ (V.T @ X)) < 1e-10 * np.linalg.norm(C.cpu().numpy())
