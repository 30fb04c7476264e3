"""GPU: the serving loop — continuous batching of RangeFinder and Kaczmarz
requests on several streams, admission by HBM budget, and the same result
for a request alone and in the mix."""
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
F32 = torch.float32


def lowrank(m, n, r, decay, g, dev):
    U = torch.randn(r, m, generator=g, device=dev, dtype=F32).T
    V = torch.randn(n, r, generator=g, device=dev, dtype=F32)
    s = torch.exp(-torch.arange(r, device=dev, dtype=F32) / decay)
    return ((U * s) @ V.T / np.sqrt(n)).T.contiguous().T  # column-major


@pytest.fixture(scope="module")
def system(dev):
    g = torch.Generator(device=dev).manual_seed(99)
    A1 = lowrank(4096, 1024, 400, 40.0, g, dev)
    A2 = torch.randn(200, 20000, generator=g, device=dev, dtype=F32).T
    bs = [A2 @ torch.randn(200, generator=g, device=dev, dtype=F32) for _ in range(4)]
    return A1, A2, bs


def requests(rla, bs):
    reqs = []
    for i in range(4):
        reqs.append(rla.RangeFinderRequest("rf", k=64 + 32 * (i % 2), power_its=2,
                                           precision="bf16" if i % 2 else "plain", seed=100 + i))
        reqs.append(rla.KaczmarzRequest("kz", bs[i], block=32, rtol=1e-5, every=20, seed=200 + i))
    return reqs


def run_alone(rla, A1, A2, reqs):
    out = []
    with rla.Server(slots=1) as srv:
        srv.register("rf", A1, "rangefinder")
        srv.register("kz", A2, "kaczmarz")
        for r in reqs:
            out.append(srv.submit(r).result())
    return out


def test_results_same_alone_and_in_the_mix(rla, dev, system):
    A1, A2, bs = system
    alone = run_alone(rla, A1, A2, requests(rla, bs))
    rng = np.random.default_rng(5)
    with rla.Server(slots=4) as srv:
        srv.register("rf", A1, "rangefinder")
        srv.register("kz", A2, "kaczmarz")
        reqs = requests(rla, bs)
        t0 = time.perf_counter()
        offs = rla.poisson_arrivals(200.0, len(reqs), rng)
        futs = []
        for r, o in zip(reqs, offs):
            while time.perf_counter() < t0 + o:
                time.sleep(0.0005)
            futs.append(srv.submit(r))
        mixed = [f.result() for f in futs]
        st = srv.stats()
        assert srv.max_in_flight >= 2  # requests really overlapped
    assert st["requests"] == 8 and st["p99_ms"] >= st["p50_ms"] > 0 and st["solves_per_s"] > 0
    for r, a, b in zip(reqs, alone, mixed):
        if r.kind == "rangefinder":
            Pa, Pb = a.double() @ a.double().T, b.double() @ b.double().T
            assert float(torch.linalg.matrix_norm(Pa - Pb)) <= 1e-5
        else:
            xa, xb = a[0].double(), b[0].double()
            assert float(torch.linalg.vector_norm(xa - xb)) <= 1e-6 * float(
                torch.linalg.vector_norm(xa))
            assert a[1].converged and b[1].converged and a[1].iteration == b[1].iteration


def test_admission_by_hbm_budget(rla, dev, system):
    A1, A2, bs = system
    resident = A2.numel() * 4 * 2 + A2.shape[0] * 8
    probe = rla.Server(slots=1)
    probe.register("kz", A2, "kaczmarz")
    need = probe._need(rla.KaczmarzRequest("kz", bs[0], block=32, every=20))
    probe.close()
    # room for exactly one Kaczmarz request at a time: four slots, one in flight
    with rla.Server(slots=4, hbm_budget=resident + need + need // 2) as srv:
        srv.register("kz", A2, "kaczmarz")
        futs = [srv.submit(rla.KaczmarzRequest("kz", b, block=32, rtol=1e-5, every=20, seed=i))
                for i, b in enumerate(bs)]
        for f in futs:
            f.result()
        assert srv.max_in_flight == 1
        with pytest.raises(rla._lib.AdmissionError):
            srv.submit(rla.KaczmarzRequest("kz", bs[0], block=64, every=4000))
    with pytest.raises(rla._lib.AdmissionError):
        with rla.Server(slots=1, hbm_budget=1000) as srv:
            srv.register("kz", A2, "kaczmarz")
