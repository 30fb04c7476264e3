"""GPU parity of the iterative Hessian sketch solver (src/Solvers/ihs.jl):
one update against the oracle's step with the same sketched matrix
(SparseSign, SRHT and Gaussian compressors), convergence to the
least-squares solution (the isometric SparseSign and Gaussian sketches), the
threshold stop, and the reference's argument checks."""
import numpy as np
import pytest
import torch

from oracle import restate as R

pytestmark = pytest.mark.gpu
F32 = torch.float32


def colmajor(a, dev):
    t = torch.as_tensor(np.asarray(a, np.float32))
    return t.T.contiguous().T.to(dev) if t.dim() == 2 else t.to(dev)


def problem(m, n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) * 10.0 ** rng.uniform(-1, 0, n)
    b = A @ rng.standard_normal(n) + 0.1 * rng.standard_normal(m)  # inconsistent
    return A.astype(np.float32), b.astype(np.float32)


@pytest.mark.parametrize("kind", ["sparse_sign", "srht", "gaussian"])
def test_ihs_one_update_matches_oracle_step(rla, dev, kind):
    # max_it = 2: iteration 1 updates x, iteration 2 only logs (ihs.jl:223-231)
    A, b = problem(3000, 40, 1)
    x0 = np.random.default_rng(2).standard_normal(40).astype(np.float32)
    Ad, bd, xd = colmajor(A, dev), colmajor(b, dev), colmajor(x0, dev)
    rla.seed(5)
    comp = {"sparse_sign": rla.SparseSign(compression_dim=320, type=F32),
            "srht": rla.SRHT(compression_dim=320, type=F32),
            "gaussian": rla.Gaussian(compression_dim=320, type=F32)}[kind]
    solver = rla.IHS(compressor=comp, log=rla.BasicLogger(max_it=2), alpha=0.75)
    x, rec = rla.rsolve(solver, xd, Ad, bd)
    SA = rec.compressed_mat[:320].cpu().numpy()
    want = R.ihs_step(A, b, x0, SA, 0.75)
    got = x.double().cpu().numpy()
    assert np.linalg.norm(got - want) <= 2e-4 * np.linalg.norm(want)
    assert rec.log.iteration == 2 and rec.log.converged


# SparseSign and Gaussian sketches are isometric in expectation (E[S'S] = I).
# The SRHT's stored scale sqrt(s/n_pad) (srht.jl:152) gives E[S'S] =
# (s^2/n_pad) I, so IHS steps with it are s^2/n_pad times too short, in the
# reference as here; its update is checked against the oracle step above.
@pytest.mark.parametrize("kind", ["sparse_sign", "gaussian"])
def test_ihs_converges_to_least_squares(rla, dev, kind):
    A, b = problem(4096, 50, 3)
    xls = np.linalg.lstsq(A.astype(np.float64), b.astype(np.float64), rcond=None)[0]
    Ad, bd = colmajor(A, dev), colmajor(b, dev)
    x = torch.zeros(50, dtype=F32, device=dev)
    comp = {"sparse_sign": rla.SparseSign(compression_dim=400, type=F32),
            "gaussian": rla.Gaussian(compression_dim=400, type=F32)}[kind]
    rla.seed(9)
    solver = rla.IHS(compressor=comp, log=rla.BasicLogger(max_it=40))
    rla.rsolve(solver, x, Ad, bd)
    got = x.double().cpu().numpy()
    assert np.linalg.norm(got - xls) <= 1e-4 * np.linalg.norm(xls), kind


def test_ihs_stops_at_threshold(rla, dev):
    # consistent system: the residual falls below the threshold and the loop stops early
    rng = np.random.default_rng(4)
    A = rng.standard_normal((2000, 30)).astype(np.float32)
    b = (A @ rng.standard_normal(30)).astype(np.float32)
    x = torch.zeros(30, dtype=F32, device=dev)
    rla.seed(1)
    solver = rla.IHS(compressor=rla.SparseSign(compression_dim=240, type=F32),
                     log=rla.BasicLogger(max_it=100, threshold=1e-3 * float(np.linalg.norm(b))))
    _, rec = rla.rsolve(solver, x, colmajor(A, dev), colmajor(b, dev))
    assert rec.log.converged and rec.log.iteration < 100
    assert rec.log.error < 1e-3 * float(np.linalg.norm(b))


def test_ihs_argument_checks(rla, dev):
    A = colmajor(np.ones((100, 10)), dev)
    b = colmajor(np.ones(100), dev)
    x = torch.zeros(10, dtype=F32, device=dev)
    with pytest.raises(rla.ArgumentError, match="not larger than column dimension"):
        rla.complete_solver(rla.IHS(compressor=rla.SparseSign(compression_dim=8, type=F32)),
                            x, A, b)
    with pytest.raises(rla.ArgumentError, match="larger row dimension"):
        rla.complete_solver(rla.IHS(compressor=rla.SparseSign(compression_dim=200, type=F32)),
                            x, A, b)
    Aw = colmajor(np.ones((10, 12)), dev)
    with pytest.raises(rla.ArgumentError):
        rla.complete_solver(rla.IHS(compressor=rla.SparseSign(compression_dim=10, type=F32)),
                            torch.zeros(12, dtype=F32, device=dev), Aw, colmajor(np.ones(10), dev))
    rec = rla.complete_solver(rla.IHS(compressor=rla.SparseSign(compression_dim=40, type=F32)),
                              x, A, b)
    with pytest.raises(rla.DimensionMismatch):
        rla.rsolve(rec, torch.zeros(9, dtype=F32, device=dev), A, b)
