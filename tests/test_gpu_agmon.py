"""GPU parity of the Agmon distribution (src/Compressors/Distributions/agmon.jl):
the reference's own sampling tests (beta cases, greedy correctness, update,
top-k tie-break, stored residuals; test/Compressors/Distributions/agmon.jl:
286-477, 1-based answers shifted to 0-based), and the top-k selection on
larger inputs against the oracle."""
import numpy as np
import pytest
import torch

from oracle import restate as R

pytestmark = pytest.mark.gpu
F64 = torch.float64


def dev_t(a, dev, dtype=F64):
    t = torch.as_tensor(np.asarray(a, np.float64)).to(dtype)
    return (t.T.contiguous().T if t.dim() == 2 else t).to(dev)


def draw(rla, dev, card, beta, A, b, x, k=1, r=None, dtype=F64):
    Ad, bd, xd = dev_t(A, dev, dtype), dev_t(b, dev, dtype), dev_t(x, dev, dtype)
    mr = rla.complete_distribution(rla.Agmon(cardinality=card, beta=beta), xd, Ad, bd)
    rla.update_distribution(mr, xd, Ad, bd, r=None if r is None else dev_t(r, dev, dtype))
    out = torch.zeros(k, dtype=torch.int32, device=dev)
    rla.sample_distribution(out, mr)
    return out.cpu().tolist(), mr


def test_agmon_beta_cases(rla, dev):
    rng = np.random.default_rng(0)
    rla.seed(3)
    got, _ = draw(rla, dev, rla.Left(), 1, rng.standard_normal((10, 3)), rng.standard_normal(10),
                  rng.standard_normal(3))
    assert 0 <= got[0] < 10
    got, _ = draw(rla, dev, rla.Left(), 3, [[1, 0], [0, 1], [0, 0]], [1, 5, 0.1], [0, 0])
    assert got == [1]
    got, _ = draw(rla, dev, rla.Left(), 10, rng.standard_normal((100, 5)),
                  rng.standard_normal(100), rng.standard_normal(5))
    assert 0 <= got[0] < 100
    A = np.array([[1.0, 0.0], [0.0, 1.0], [1.0, 1.0]])
    got, _ = draw(rla, dev, rla.Right(), 2, A, [0, 0, 0], [1, 2])
    assert got == [int(np.argmax(np.abs(A.T @ (A @ [1, 2]))))] == [1]
    got, _ = draw(rla, dev, rla.Right(), 1, rng.standard_normal((5, 8)), rng.standard_normal(5),
                  rng.standard_normal(8))
    assert 0 <= got[0] < 8
    A, b, x = rng.standard_normal((20, 50)), rng.standard_normal(20), rng.standard_normal(50)
    got, mr = draw(rla, dev, rla.Right(), 8, A, b, x)
    cand = mr.sample_buffer.cpu().numpy()
    assert got[0] in cand and 0 <= got[0] < 50
    assert got == R.agmon_select(A.T @ (A @ x - b), cand, 1)


def test_agmon_greedy_correctness(rla, dev):
    A = [[2.0, 1.0], [1.0, 3.0], [0.5, 0.5], [1.0, 1.0]]
    b, x = [1.0, 2.0, 3.0, 0.5], [0.5, 0.5]
    res = np.abs(np.asarray(A) @ x - b)
    assert draw(rla, dev, rla.Left(), 4, A, b, x)[0] == [int(np.argmax(res))]
    # update changes the selection
    A2 = [[1.0, 0.0], [0.0, 1.0]]
    assert draw(rla, dev, rla.Left(), 2, A2, [2.0, 3.0], [0.0, 0.0])[0] == [1]
    assert draw(rla, dev, rla.Left(), 2, A2, [2.0, 3.0], [10.0, 0.0])[0] == [0]
    # top-k with a tie at the cut: smaller index first
    got, _ = draw(rla, dev, rla.Left(), 4, [[1, 0], [0, 1], [1, 0], [0, 1]], [2, 2, 2, 1],
                  [0, 0], k=2)
    assert got == [0, 1]
    # Right top-k
    A3 = np.array([[1.0, 0.0, 2.0], [0.0, 1.0, 1.0], [1.0, 1.0, 0.0]])
    got, _ = draw(rla, dev, rla.Right(), 3, A3, [0, 0, 0], [1, 1, 1], k=2)
    scores = np.abs(A3.T @ (A3 @ np.ones(3)))
    assert sorted(got) == sorted(np.argsort(-scores, kind="stable")[:2].tolist())
    # stored residuals are read instead of A x - b
    got, _ = draw(rla, dev, rla.Left(), 4, [[1, 0], [0, 1], [1, 1], [0, 0]], [0, 0, 0, 0],
                  [0, 0], r=[0.1, 0.2, 9.9, 0.3])
    assert got == [2]
    got, _ = draw(rla, dev, rla.Right(), 3, np.eye(3), [0, 0, 0], [0, 0, 0], r=[1.0, 2.0, 7.0])
    assert got == [2]


@pytest.mark.parametrize("dtype", [torch.float32, F64])
@pytest.mark.parametrize("card,beta,k", [("Left", 5000, 16), ("Left", 700, 64),
                                         ("Right", 300, 9), ("Right", 64, 64)])
def test_agmon_topk_matches_oracle(rla, dev, dtype, card, beta, k):
    # integer-valued A and stored residual r: A'r is exact in both precisions,
    # so the scores, and their many ties, are the oracle's exactly
    rng = np.random.default_rng(beta + k)
    m, n = 5000, 300
    A = rng.integers(-2, 3, (m, n)).astype(np.float64)
    x = rng.standard_normal(n)
    b = rng.standard_normal(m)
    r = rng.integers(-6, 7, m).astype(np.float64)
    rla.seed(beta)
    cardinality = rla.Left() if card == "Left" else rla.Right()
    got, mr = draw(rla, dev, cardinality, beta, A, b, x, k=k, r=r, dtype=dtype)
    d = m if card == "Left" else n
    cand = np.arange(d) if beta >= d else mr.sample_buffer.cpu().numpy()
    v = r if card == "Left" else A.T @ r
    assert got == R.agmon_select(v, cand, k)


def test_agmon_on_the_fly_residual_matches_oracle(rla, dev):
    # no stored r: the device forms A x - b (rows) and A'(A x - b) (columns)
    rng = np.random.default_rng(7)
    A = rng.integers(-2, 3, (400, 30)).astype(np.float64)
    x = rng.integers(-2, 3, 30).astype(np.float64)
    b = rng.integers(-5, 6, 400).astype(np.float64)
    got, _ = draw(rla, dev, rla.Left(), 400, A, b, x, k=12)
    assert got == R.agmon_select(A @ x - b, np.arange(400), 12)
    got, _ = draw(rla, dev, rla.Right(), 30, A, b, x, k=5)
    assert got == R.agmon_select(A.T @ (A @ x - b), np.arange(30), 5)
