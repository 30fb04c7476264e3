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


def test_same_alone_and_in_the_mix_at_c5_sizes(rla, dev):
    # the served stream's own sizes: the 100000 x 1000 Kaczmarz system (64-row
    # L2Norm blocks, rtol 1e-6, FP64 residual every 100) beside a 16384 x 4096
    # RangeFinder (k = 266, q = 2), both precisions
    g = torch.Generator(device=dev).manual_seed(5)
    Arf = lowrank(16384, 4096, 1024, 64.0, g, dev)
    Akz = torch.randn(1000, 100000, generator=g, device=dev, dtype=F32).T
    bs = [Akz @ torch.randn(1000, generator=g, device=dev, dtype=F32) for _ in range(2)]

    def reqs():
        out = []
        for i in range(2):
            out.append(rla.RangeFinderRequest("rf", k=266, power_its=2,
                                              precision="bf16" if i else "plain", seed=10 + i))
            out.append(rla.KaczmarzRequest("kz", bs[i], block=64, rtol=1e-6, every=100,
                                           seed=20 + i))
        return out

    alone = run_alone(rla, Arf, Akz, reqs())
    with rla.Server(slots=4) as srv:
        srv.register("rf", Arf, "rangefinder")
        srv.register("kz", Akz, "kaczmarz")
        futs = [srv.submit(r) for r in reqs()]
        mixed = [f.result() for f in futs]
        assert srv.max_in_flight >= 2
    for r, a, b in zip(reqs(), alone, mixed):
        if r.kind == "rangefinder":
            Pa, Pb = a.double() @ a.double().T, b.double() @ b.double().T
            assert float(torch.linalg.matrix_norm(Pa - Pb)) <= 1e-5
        else:
            xa, xb = a[0].double(), b[0].double()
            assert float(torch.linalg.vector_norm(xa - xb)) <= 1e-6 * float(
                torch.linalg.vector_norm(xa))
            assert a[1].converged and b[1].converged and a[1].iteration == b[1].iteration
            assert a[1].error <= 1e-6 * float(torch.linalg.vector_norm(r.b.double()))


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


def test_entry_point_workspace_honours_hbm_budget(rla, dev):
    # an entry point whose context workspace would grow past the device's HBM
    # budget fails with RLA_EADMIT before it launches anything
    import ctypes
    import threading
    lib = rla._lib.load()
    used, budget = ctypes.c_size_t(), ctypes.c_size_t()
    rla._lib.check(lib.rla_hbm_query(0, ctypes.byref(used), ctypes.byref(budget)))
    out = {}

    def work():
        idx1 = torch.arange(1, 11, dtype=torch.int64, device=dev)
        idx0 = torch.full((10,), -1, dtype=torch.int32, device=dev)
        c = rla._lib.ctx(dev)  # this thread's new context holds no workspace yet
        try:
            lib.rla_hbm_set_budget(0, used.value + 4096)
            out["rc"] = lib.rla_index_from_julia(c, rla._lib.ptr(idx1), 10, 100,
                                                 rla._lib.ptr(idx0))
            out["msg"] = lib.rla_last_error()
            out["untouched"] = idx0.cpu().clone()
            lib.rla_hbm_set_budget(0, budget.value)
            out["rc2"] = lib.rla_index_from_julia(c, rla._lib.ptr(idx1), 10, 100,
                                                  rla._lib.ptr(idx0))
            out["idx0"] = idx0.cpu()
        finally:
            lib.rla_hbm_set_budget(0, budget.value)
            rla._lib.drop_thread_contexts()

    t = threading.Thread(target=work)
    t.start()
    t.join()
    assert out["rc"] == rla._lib.RLA_EADMIT and b"HBM budget" in out["msg"]
    assert (out["untouched"] == -1).all()
    assert out["rc2"] == 0 and torch.equal(out["idx0"], torch.arange(10, dtype=torch.int32))
    after = ctypes.c_size_t()
    lib.rla_hbm_query(0, ctypes.byref(after), None)
    assert after.value == used.value  # the dropped context released its workspace
