"""GPU parity: batched CountSketch sketch-and-solve least squares (one hipGraph
per batch) and the QR sub-solver, against the oracle's sparse product and
FP64 least squares."""
import numpy as np
import pytest
import torch

from oracle import restate as R

pytestmark = pytest.mark.gpu
F32 = torch.float32


def colmajor(a, dev):
    return torch.as_tensor(np.asarray(a, np.float32)).T.contiguous().T.to(dev)


def make_problem(m, n, seed, scale_cols=True):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n))
    if scale_cols:
        A = A * 10.0 ** rng.uniform(-3, 0, n)
    A = A.astype(np.float32)
    xs = rng.standard_normal(n).astype(np.float32)
    b = (A.astype(np.float64) @ xs + 1e-6 * rng.standard_normal(m)).astype(np.float32)
    return A, b, xs


def sketch_oracle(S, A, b):
    """[SA Sb] in the SparseArrays loop order, FP32 (count_sketch.jl mul!)."""
    ptr = S.op.rowptr.cpu().numpy()
    idx = S.op.colidx.cpu().numpy()
    val = S.op.val_t.cpu().numpy()
    Ab = np.concatenate([A, b[:, None]], axis=1)
    return R.sparse_rowgather(ptr, idx, val, Ab, dtype=np.float32)


SHAPES = [(2000, 32), (5000, 100), (3001, 17), (700, 1), (12000, 256), (4096, 64), (999, 250)]


@pytest.mark.parametrize("cfg", [2, 3, 4])
def test_batch_sketch_bit_exact_and_solution(rla, dev, cfg):
    rla.seed(2024)
    probs = [make_problem(m, n, i) for i, (m, n) in enumerate(SHAPES)]
    As = [colmajor(A, dev) for A, _, _ in probs]
    bs = [torch.as_tensor(b).to(dev) for _, b, _ in probs]
    xs = [torch.zeros(A.shape[1], dtype=F32, device=dev) for A, _, _ in probs]
    batch = rla.SketchSolveBatch(As, bs, xs, autotune=cfg)  # forced tile shape
    batch.run()
    assert batch.info() == [0] * len(probs)
    for p, (A, b, xstar) in enumerate(probs):
        got = batch.sketch(p).cpu().numpy()
        want = sketch_oracle(batch.recipes[p], A, b)
        assert np.array_equal(got, want), f"problem {p}: sketch differs from the oracle"
        SA, Sb = want[:, :-1].astype(np.float64), want[:, -1].astype(np.float64)
        xo = np.linalg.lstsq(SA, Sb, rcond=None)[0]
        x = xs[p].cpu().numpy().astype(np.float64)
        cond = np.linalg.cond(SA)
        # FP64 normal equations of the FP32 sketch: ~cond^2 * 1e-16, plus FP32 output
        assert np.linalg.norm(x - xo) <= (1e-6 + cond ** 2 * 1e-15) * np.linalg.norm(xo)
    # replaying the graph reproduces the result exactly
    before = [x.clone() for x in xs]
    for x in xs:
        x.zero_()
    batch.run()
    assert all(torch.equal(a, b) for a, b in zip(before, xs))
    batch.close()


def test_autotuned_plan_and_custom_sketch_sizes(rla, dev, monkeypatch, tmp_path):
    # s = 123 is not a multiple of 4 (the Gram kernel's unaligned-column path);
    # an empty tile cache, so all three shapes are timed
    monkeypatch.setenv("RLA_TILE_CACHE", str(tmp_path / "tiles.json"))
    monkeypatch.setattr(rla.batched, "_TILE_CACHE_DEFAULT", str(tmp_path / "none.json"))
    probs = [make_problem(m, n, 50 + i)
             for i, (m, n) in enumerate([(3000, 40), (800, 64), (2000, 33)])]
    As = [colmajor(A, dev) for A, _, _ in probs]
    bs = [torch.as_tensor(b).to(dev) for _, b, _ in probs]
    xs = [torch.zeros(A.shape[1], dtype=F32, device=dev) for A, _, _ in probs]
    batch = rla.SketchSolveBatch(As, bs, xs, compression_dims=[120, 1024, 123], autotune=1)
    cfg = batch.config()
    assert cfg["cols_per_tile"] in (8, 16, 32) and min(cfg["tuned_ms"]) > 0
    batch.run()
    assert batch.info() == [0, 0, 0]
    for p, (A, b, xstar) in enumerate(probs):
        want = sketch_oracle(batch.recipes[p], A, b)
        assert want.shape == (batch.shapes[p][2], A.shape[1] + 1)
        assert np.array_equal(batch.sketch(p).cpu().numpy(), want)
        # x solves the sketched problem (FP64 normal equations, cond(SA) ~ 1e3)
        SA, Sb = want[:, :-1].astype(np.float64), want[:, -1].astype(np.float64)
        xls = np.linalg.lstsq(SA, Sb, rcond=None)[0]
        got = xs[p].double().cpu().numpy()
        assert np.linalg.norm(got - xls) <= 1e-5 * np.linalg.norm(xls), p


def test_tile_cache_written_then_used(rla, dev, monkeypatch, tmp_path):
    # the first plan of a bucket times the three shapes and records the best;
    # a second batch of the same bucket takes the recorded shape (timing only
    # that one) and still sketches bit-exactly
    import json
    path = tmp_path / "tiles.json"
    monkeypatch.setenv("RLA_TILE_CACHE", str(path))
    monkeypatch.setattr(rla.batched, "_TILE_CACHE_DEFAULT", str(tmp_path / "none.json"))
    probs = [make_problem(m, n, 70 + i) for i, (m, n) in enumerate([(3000, 40), (2500, 48)])]
    As = [colmajor(A, dev) for A, _, _ in probs]
    bs = [torch.as_tensor(b).to(dev) for _, b, _ in probs]
    x1 = [torch.zeros(A.shape[1], dtype=F32, device=dev) for A, _, _ in probs]
    x2 = [torch.zeros(A.shape[1], dtype=F32, device=dev) for A, _, _ in probs]
    b1 = rla.SketchSolveBatch(As, bs, x1, autotune=1)
    assert not b1.tile_cached and min(b1.config()["tuned_ms"]) > 0
    saved = json.loads(path.read_text())
    assert b1.tile_key in saved and saved[b1.tile_key]["shape"] in (0, 1, 2)
    shape = saved[b1.tile_key]["shape"]
    rla.seed(5)
    b2 = rla.SketchSolveBatch(As, bs, x2, autotune=1)
    assert b2.tile_cached and b2.tile_key == b1.tile_key
    ms = b2.config()["tuned_ms"]
    assert ms[shape] > 0 and all(ms[c] == -1 for c in range(3) if c != shape)
    assert b2.config()["cols_per_tile"] == b1.config()["cols_per_tile"]
    b1.run()
    b2.run()
    assert b1.info() == [0, 0] and b2.info() == [0, 0]
    for p, (A, b, _) in enumerate(probs):
        assert np.array_equal(b2.sketch(p).cpu().numpy(), sketch_oracle(b2.recipes[p], A, b))


def test_rank_deficient_problem_reports_breakdown(rla, dev):
    A, b, _ = make_problem(1000, 20, 7, scale_cols=False)
    A[:, 5] = A[:, 3]
    A2, b2, _ = make_problem(1500, 30, 8)
    As = [colmajor(A, dev), colmajor(A2, dev)]
    bs = [torch.as_tensor(b).to(dev), torch.as_tensor(b2).to(dev)]
    xs = [torch.full((20,), 7.0, dtype=F32, device=dev), torch.zeros(30, dtype=F32, device=dev)]
    batch = rla.SketchSolveBatch(As, bs, xs, autotune=0)
    batch.run()
    info = batch.info()
    assert info[0] > 0 and info[1] == 0
    assert torch.all(xs[0] == 7.0)  # left untouched


def test_argument_checks(rla, dev):
    A, b, _ = make_problem(100, 10, 1)
    with pytest.raises(rla._lib.ArgumentError):
        rla.SketchSolveBatch([colmajor(A, dev)], [torch.as_tensor(b).to(dev)],
                             [torch.zeros(9, dtype=F32, device=dev)])
    with pytest.raises(rla._lib.ArgumentError, match="fewer rows"):
        rla.SketchSolveBatch([colmajor(A, dev)], [torch.as_tensor(b).to(dev)],
                             [torch.zeros(10, dtype=F32, device=dev)], compression_dims=[8])


def test_qr_sub_solver_ldiv(rla, dev):
    A, b, _ = make_problem(400, 60, 9)
    Ad, bd = colmajor(A, dev), torch.as_tensor(b).to(dev)
    x = torch.zeros(60, dtype=F32, device=dev)
    rec = rla.complete_sub_solver(rla.QRSolver(), Ad, bd)
    rla.ldiv(x, rec, bd)
    xo = np.linalg.lstsq(A.astype(np.float64), b.astype(np.float64), rcond=None)[0]
    cond = np.linalg.cond(A.astype(np.float64))
    assert np.linalg.norm(x.cpu().numpy() - xo) <= (1e-6 + cond ** 2 * 1e-15) * np.linalg.norm(xo)


@pytest.mark.parametrize("m,n,dt", [(3000, 300, torch.float64), (2000, 280, F32),
                                    (1500, 1, torch.float64), (64, 64, torch.float64)])
def test_qr_sub_solver_any_size_and_type(rla, dev, m, n, dt):
    # qr.jl:36-43 on any Float32/Float64 matrix: sizes beyond the batch limits
    # (n + 1 > 272, m > 1024) and square systems, against LAPACK lstsq
    A, b, _ = make_problem(m, n, 11)
    A = A.astype(np.float64) if dt == torch.float64 else A
    b = b.astype(np.float64) if dt == torch.float64 else b
    Ad = torch.as_tensor(A).to(dev).T.contiguous().T
    bd = torch.as_tensor(b).to(dev)
    x = torch.zeros(n, dtype=dt, device=dev)
    rec = rla.complete_sub_solver(rla.QRSolver(), Ad, bd)
    rla.ldiv(x, rec, bd)
    xo = np.linalg.lstsq(A.astype(np.float64), b.astype(np.float64), rcond=None)[0]
    cond = np.linalg.cond(A.astype(np.float64))
    tol = (1e-12 + cond ** 2 * 1e-16) if dt == torch.float64 else (1e-6 + cond ** 2 * 1e-15)
    assert np.linalg.norm(x.cpu().numpy() - xo) <= tol * np.linalg.norm(xo)
    # A unchanged, and a second solve with another b reuses the factor
    assert np.array_equal(Ad.cpu().numpy(), A)
    b2 = torch.as_tensor(np.roll(b, 3)).to(dev)
    rla.ldiv(x, rec, b2)
    xo2 = np.linalg.lstsq(A.astype(np.float64), np.roll(b, 3).astype(np.float64), rcond=None)[0]
    assert np.linalg.norm(x.cpu().numpy() - xo2) <= tol * np.linalg.norm(xo2)


def test_qr_sub_solver_strided_x_and_changed_a(rla, dev):
    A, b, _ = make_problem(500, 40, 12)
    A = A.astype(np.float64)
    b = b.astype(np.float64)
    Ad = torch.as_tensor(A).to(dev).T.contiguous().T
    bd = torch.as_tensor(b).to(dev)
    buf = torch.zeros(80, dtype=torch.float64, device=dev)
    x = buf[::2]
    rec = rla.complete_sub_solver(rla.QRSolver(), Ad, bd)
    rla.ldiv(x, rec, bd)
    xo = np.linalg.lstsq(A, b, rcond=None)[0]
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-10 * np.linalg.norm(xo)
    assert float(buf[1::2].abs().max()) == 0.0
    Ad.mul_(2.0)  # changed in place: the cached factor must not be reused
    rla.ldiv(x, rec, bd)
    assert np.linalg.norm(x.cpu().numpy() - xo / 2) <= 1e-10 * np.linalg.norm(xo)


def test_qr_sub_solver_rank_deficient_raises(rla, dev):
    A, b, _ = make_problem(300, 20, 13)
    A = A.astype(np.float64)
    A[:, 7] = 0.0
    Ad = torch.as_tensor(A).to(dev).T.contiguous().T
    bd = torch.as_tensor(b.astype(np.float64)).to(dev)
    x = torch.zeros(20, dtype=torch.float64, device=dev)
    rec = rla.complete_sub_solver(rla.QRSolver(), Ad, bd)
    with pytest.raises(rla._lib.DomainError, match="SingularException"):
        rla.ldiv(x, rec, bd)
    with pytest.raises(rla._lib.ArgumentError, match="same element type"):
        rla.ldiv(torch.zeros(20, dtype=F32, device=dev), rec, bd)


def test_c4_batch_512_problems(rla, dev):
    # the batched least-squares workload: 512 problems, m log-uniform in
    # [2000, 50000], n uniform in [32, 256], columns scaled by 10^U(-3,0),
    # b = A x* + 1e-6 noise, CountSketch to 4n rows
    rng = np.random.default_rng(1234)
    g = torch.Generator(device=dev).manual_seed(1234)
    ms = np.floor(np.exp(rng.uniform(np.log(2000), np.log(50000), 512))).astype(int)
    ns = rng.integers(32, 257, 512)
    As, bs, xs, xstars = [], [], [], []
    for m, n in zip(ms, ns):
        sc = torch.as_tensor(10.0 ** rng.uniform(-3, 0, n), dtype=F32, device=dev)
        A = (torch.randn(int(n), int(m), generator=g, device=dev, dtype=F32).T) * sc
        xsr = torch.randn(int(n), generator=g, device=dev, dtype=F32)
        b = A @ xsr + 1e-6 * torch.randn(int(m), generator=g, device=dev, dtype=F32)
        As.append(A)
        bs.append(b)
        xs.append(torch.zeros(int(n), dtype=F32, device=dev))
        xstars.append(xsr)
    rla.seed(1234)
    batch = rla.SketchSolveBatch(As, bs, xs, autotune=1)
    batch.run()
    assert batch.info() == [0] * 512
    err = np.array([float(torch.linalg.vector_norm((x - xsr).double()) /
                          torch.linalg.vector_norm(xsr.double())) for x, xsr in zip(xs, xstars)])
    assert np.all(np.isfinite(err)) and np.median(err) < 1e-3 and err.max() < 1e-1
    # FP64 oracle on a few problems: the same sketch solved by numpy lstsq
    for p in (0, 100, 511):
        Sab = batch.sketch(p).cpu().numpy().astype(np.float64)
        xo = np.linalg.lstsq(Sab[:, :-1], Sab[:, -1], rcond=None)[0]
        cond = np.linalg.cond(Sab[:, :-1])
        x = xs[p].cpu().numpy()
        assert np.linalg.norm(x - xo) <= (1e-6 + cond ** 2 * 1e-15) * np.linalg.norm(xo)
