"""GPU parity at the full RangeFinder size (32768 x 8192, sigma_i = exp(-i/64),
k = 266, q = 2, orthogonalised): the BF16 and plain paths against an
independent FP64 run of the same algorithm with the same Omega (torch FP64
products and Householder QR, power_its.jl:53-80), the Eckart-Young bound, and
the BF16 GEMM at the two product shapes of that run against FP64 products of
the BF16-rounded operands."""
import pytest
import torch

pytestmark = pytest.mark.gpu
F64, F32 = torch.float64, torch.float32
M, N, K, RANK = 32768, 8192, 266, 2560


@pytest.fixture(scope="module")
def c3(dev):
    g = torch.Generator(device=dev).manual_seed(1234)
    U, _ = torch.linalg.qr(torch.randn(M, RANK, generator=g, device=dev, dtype=F64))
    V, _ = torch.linalg.qr(torch.randn(N, RANK, generator=g, device=dev, dtype=F64))
    sig = torch.exp(-torch.arange(RANK, device=dev, dtype=F64) / 64.0)
    A64 = ((U * sig) @ V.T).T.contiguous().T       # column-major
    del U, V
    A32 = A64.float().T.contiguous().T
    full = torch.exp(-torch.arange(N, device=dev, dtype=F64) / 64.0)
    opt = float(torch.sqrt((full[K:] ** 2).sum() / (full ** 2).sum()))
    yield A64, A32, opt
    del A64, A32
    torch.cuda.empty_cache()


def _q(Y):
    Q, R = torch.linalg.qr(Y)
    return Q * torch.sign(torch.diagonal(R))      # positive-diagonal R, like the kernels


def oracle_range(A64, Om):
    """rand_ortho_it (power_its.jl:53-80) in FP64 with torch: Y = A Om, QR,
    then q times Z = A'Q, QR, Y = A Z, QR."""
    Q = _q(A64 @ Om)
    for _ in range(2):
        Q = _q(A64 @ _q(A64.T @ Q))
    return Q


def rel_err(A64, Q):
    return float(torch.linalg.norm(A64 - Q @ (Q.T @ A64)) / torch.linalg.norm(A64))


def test_c3_rangefinder_bf16_and_plain_vs_fp64_same_omega(rla, dev, c3):
    A64, A32, opt = c3
    res = {}
    for prec in ("bf16", "plain"):
        rla.seed(7)
        rf = rla.RangeFinder(compressor=rla.Gaussian(cardinality=rla.Right(), compression_dim=K,
                                                     type=F32),
                             power_its=2, orthogonalize=True, precision=prec)
        rec = rla.rapproximate(rf, A32)
        res[prec] = (rec.range.double(), rec.compressor.op.double())
    Om = res["bf16"][1]
    assert torch.equal(Om, res["plain"][1])       # the same Omega for both paths
    Q64 = oracle_range(A64, Om)
    e64 = rel_err(A64, Q64)
    eb, ep = rel_err(A64, res["bf16"][0]), rel_err(A64, res["plain"][0])
    # the FP64 run is near the optimum; both paths are near the FP64 run
    assert opt <= e64 <= 1.05 * opt, (e64, opt)
    assert ep <= 1.01 * e64, (ep, e64)
    assert eb <= 1.1 * ep and eb <= 1.1 * e64, (eb, ep, e64)
    print(f"C3 errors: fp64 {e64:.6f} plain {ep:.6f} bf16 {eb:.6f} optimum {opt:.6f}")
    for prec in ("bf16", "plain"):
        Q = res[prec][0]
        # Q is returned in A's type (Float32): orthonormal to Float32 rounding
        assert float(torch.linalg.norm(Q.T @ Q - torch.eye(K, device=dev, dtype=F64))) < 1e-6
        # distance of the two approximations of A, relative to ||A||
        d = Q @ (Q.T @ A64) - Q64 @ (Q64.T @ A64)
        dist = float(torch.linalg.norm(d) / torch.linalg.norm(A64))
        print(f"C3 {prec}: ||(P - P64) A|| / ||A|| = {dist:.6f}")
        # measured 1e-6 (plain) and 3.6e-4 (bf16), against an error of 0.0163
        assert dist <= (1e-4 if prec == "plain" else 2e-3), dist


@pytest.mark.parametrize("m_,n_,k_", [(M, K, N), (N, K, M)])
def test_gemm_bf16_at_c3_shapes(rla, dev, m_, n_, k_):
    # C = A B with A (m x k) and B (k x n) rounded to BF16: the kernel's FP32
    # sums against the FP64 product of the rounded operands
    g = torch.Generator(device=dev).manual_seed(m_)
    A = torch.randn(m_, k_, generator=g, device=dev, dtype=F32)
    B = torch.randn(k_, n_, generator=g, device=dev, dtype=F32)
    want = A.to(torch.bfloat16).double() @ B.to(torch.bfloat16).double()
    rows = rla._dense.bf16_rows(rla.as_view(A))          # m x ld8(k), rows contiguous
    cols = rla._dense.bf16_rows(rla.as_view(B).T)        # n x ld8(k)
    C = torch.zeros(n_, m_, dtype=F32, device=dev).T
    rla._dense.gemm_bf16(m_, n_, k_, rows, cols, rla.as_view(C))
    err = float((C.double() - want).abs().max() / want.abs().max())
    assert err < 2e-6, err
