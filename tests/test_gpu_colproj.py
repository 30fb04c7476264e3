"""GPU parity of the column projection solver (src/Solvers/col_projection.jl):
one update against the oracle's step with the same compression S (SparseSign,
Gaussian and SRHT from the right, and a single column), convergence to the
least-squares solution, the LSGradient error and threshold stop, and the
reference's argument checks."""
import numpy as np
import pytest
import torch

from oracle import restate as R

pytestmark = pytest.mark.gpu
F32 = torch.float32


def colmajor(a, dev):
    t = torch.as_tensor(np.asarray(a, np.float32))
    return t.T.contiguous().T.to(dev) if t.dim() == 2 else t.to(dev)


def problem(m, n, seed, spread=0.5):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) * 10.0 ** rng.uniform(-spread, 0, n)
    b = A @ rng.standard_normal(n) + 0.1 * rng.standard_normal(m)  # inconsistent
    return A.astype(np.float32), b.astype(np.float32)


def dense_right(rla, S, n, dev):
    """The n x k operator of a Right recipe: I_n * S."""
    C = torch.zeros(S.size(2), n, dtype=F32, device=dev).T
    rla.mul(C, torch.eye(n, dtype=F32, device=dev), S)
    return C.double().cpu().numpy()


def compressor(rla, kind, k):
    return {"sparse_sign": rla.SparseSign(cardinality=rla.Right(), compression_dim=k, type=F32),
            "gaussian": rla.Gaussian(cardinality=rla.Right(), compression_dim=k, type=F32),
            "srht": rla.SRHT(cardinality=rla.Right(), compression_dim=k, type=F32)}[kind]


@pytest.mark.parametrize("kind,k", [("sparse_sign", 8), ("gaussian", 8), ("srht", 8),
                                    ("sparse_sign", 1), ("gaussian", 1)])
def test_colproj_one_update_matches_oracle_step(rla, dev, kind, k):
    # max_it = 1: one update at iteration 0, then the terminal error (:356-371)
    A, b = problem(2000, 40, 1)
    x0 = np.random.default_rng(2).standard_normal(40).astype(np.float32)
    Ad, bd, xd = colmajor(A, dev), colmajor(b, dev), colmajor(x0, dev)
    rla.seed(5)
    solver = rla.ColumnProjection(compressor=compressor(rla, kind, k),
                                  log=rla.BasicLogger(max_it=1), alpha=0.75)
    x, rec = rla.rsolve(solver, xd, Ad, bd)
    S = dense_right(rla, rec.compressor, 40, dev)
    want = R.colproj_step(A, b, x0, S, 0.75)
    got = x.double().cpu().numpy()
    assert np.linalg.norm(got - want) <= 2e-5 * np.linalg.norm(want)
    assert rec.log.iteration == 1 and rec.log.converged
    # the terminal error is the LSGradient norm at the new iterate (Float32 residual)
    assert rec.log.error == pytest.approx(R.ls_gradient(A, got, b), rel=1e-3)


@pytest.mark.parametrize("kind", ["sparse_sign", "gaussian"])
def test_colproj_converges_to_least_squares(rla, dev, kind):
    # well-conditioned columns: a random 12 of 48 directions per step contracts
    # the error by about 1 - (12/48)/cond(A)^2 per iteration
    A, b = problem(3000, 48, 3, spread=0.0)
    xls = np.linalg.lstsq(A.astype(np.float64), b.astype(np.float64), rcond=None)[0]
    Ad, bd = colmajor(A, dev), colmajor(b, dev)
    x = torch.zeros(48, dtype=F32, device=dev)
    rla.seed(9)
    solver = rla.ColumnProjection(compressor=compressor(rla, kind, 12),
                                  log=rla.BasicLogger(max_it=300))
    rla.rsolve(solver, x, Ad, bd)
    got = x.double().cpu().numpy()
    assert np.linalg.norm(got - xls) <= 1e-4 * np.linalg.norm(xls), kind


def test_colproj_stops_at_gradient_threshold(rla, dev):
    rng = np.random.default_rng(4)
    A = rng.standard_normal((1500, 30)).astype(np.float32)
    b = (A @ rng.standard_normal(30)).astype(np.float32)
    x = torch.zeros(30, dtype=F32, device=dev)
    rla.seed(1)
    thr = 1e-3 * float(np.linalg.norm(A.T.astype(np.float64) @ b))
    solver = rla.ColumnProjection(
        compressor=compressor(rla, "sparse_sign", 10),
        log=rla.BasicLogger(max_it=500, threshold=thr))
    _, rec = rla.rsolve(solver, x, colmajor(A, dev), colmajor(b, dev))
    assert rec.log.converged and rec.log.iteration < 500
    assert rec.log.error < thr


def test_colproj_argument_checks(rla, dev):
    A = colmajor(np.ones((100, 10)), dev)
    b = colmajor(np.ones(100), dev)
    x = torch.zeros(10, dtype=F32, device=dev)
    with pytest.raises(rla.ArgumentError, match="does not contain the field 'gradient'"):
        rla.complete_solver(rla.ColumnProjection(compressor=compressor(rla, "sparse_sign", 4),
                                                 error=rla.FullResidual()), x, A, b)
    rec = rla.complete_solver(rla.ColumnProjection(compressor=compressor(rla, "sparse_sign", 4)),
                              x, A, b)
    with pytest.raises(rla.DimensionMismatch):
        rla.rsolve(rec, torch.zeros(9, dtype=F32, device=dev), A, b)
    with pytest.raises(rla.DimensionMismatch):
        rla.rsolve(rec, x, A, torch.zeros(99, dtype=F32, device=dev))
