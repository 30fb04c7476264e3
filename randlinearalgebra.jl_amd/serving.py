"""A request loop that serves RangeFinder and Kaczmarz solves on one GPU.

Continuous batching: requests are admitted one by one, in arrival order, as
soon as a slot (a worker with its own HIP stream) is free and their device
memory fits the HBM budget; admitted requests run concurrently on their
streams, so one request's one-workgroup factorisations and latency-bound
loops overlap another's bandwidth-bound products.  Matrices are registered
once and stay resident with the derived data the kernels read (the two BF16
copies of a RangeFinder matrix; the row-major copy and squared row norms of a
Kaczmarz matrix).  Every request carries its own random seed, so its result
does not depend on what else is in flight.

Latency is measured from a request's arrival time (given, or the submit
time) to the completion of its last kernel.
"""
import collections
import ctypes
import threading
import time
from concurrent.futures import Future

import numpy as np
import torch

from . import _dense, _lib
from ._lib import AdmissionError, ArgumentError
from .approximators import (PRECISIONS, RangeFinder, _Bf16Copies, complete_approximator,
                            rapproximate_)
from .compressors import Gaussian, Right, as_view, seed
from .solvers import (BasicLogger, FullResidual, Kaczmarz, L2Norm, Sampling, complete_solver,
                      rsolve)

_setup_lock = threading.Lock()  # recipe construction draws from the global stream
# per request: growth of the slot context's own library workspace
_SLOT_SCRATCH = 8 << 20


class RangeFinderRequest:
    """Q (m x k, FP64) of a registered matrix: Gaussian sketch, power_its
    orthogonalised power iterations, products in `precision`."""

    kind = "rangefinder"

    def __init__(self, matrix, k, power_its=2, orthogonalize=True, precision="bf16", seed=0,
                 arrival=None):
        if precision not in PRECISIONS:
            raise ArgumentError(f"precision must be one of {sorted(PRECISIONS)}")
        self.matrix, self.k, self.power_its = matrix, int(k), int(power_its)
        self.orthogonalize, self.precision, self.seed = orthogonalize, precision, int(seed)
        self.arrival = arrival


class KaczmarzRequest:
    """x with A x = b for a registered matrix: Sampling(L2Norm) blocks of
    `block` rows, FP64 residual every `every` iterations, stop at
    ||b - A x|| <= rtol ||b|| or max_it."""

    kind = "kaczmarz"

    def __init__(self, matrix, b, block=64, rtol=1e-6, max_it=20000, every=100, seed=0,
                 arrival=None):
        self.matrix, self.b, self.block, self.rtol = matrix, b, int(block), float(rtol)
        self.max_it, self.every, self.seed = int(max_it), int(every), int(seed)
        self.arrival = arrival


class _Resident:
    def __init__(self, name, A, kind):
        self.name, self.A, self.kind = name, A, kind
        self.bf16 = None
        self.A_rm = None
        self.weights = None
        self.key = None
        self.bytes = A.numel() * A.element_size()


class Server:
    """Serve requests on `device` with at most `slots` in flight and
    `hbm_budget` bytes for resident matrices plus in-flight working memory."""

    def __init__(self, device=None, hbm_budget=None, slots=4):
        self.device = torch.device("cuda", 0) if device is None else torch.device(device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.85 * free)
        self.budget = int(hbm_budget)
        self.slots = int(slots)
        if self.slots < 1:
            raise ArgumentError("slots must be positive")
        # Admission runs on the library's HBM account of the device (rla_hbm_*),
        # which also carries the library's own workspaces: the account's budget
        # is what the library already holds plus this server's budget, so the
        # growth of workspaces while serving counts against the budget.  One
        # server per device at a time.
        self._dev = self.device.index if self.device.index is not None else 0
        used0, prev = ctypes.c_size_t(0), ctypes.c_size_t(0)
        _lib.call("rla_hbm_query", self._dev, ctypes.byref(used0), ctypes.byref(prev))
        self._prev_budget = prev.value
        _lib.call("rla_hbm_set_budget", self._dev, used0.value + self.budget)
        self.resident = {}
        self._resident_bytes = 0
        self.in_flight = 0
        self.max_in_flight = 0
        self._cv = threading.Condition()
        self._queue = collections.deque()
        self._closed = False
        self.records = []
        self._workers = [threading.Thread(target=self._worker, daemon=True)
                         for _ in range(self.slots)]
        self._streams = [torch.cuda.Stream(device=self.device) for _ in range(self.slots)]
        for i, w in enumerate(self._workers):
            w.slot = i
            w.start()

    # -- resident matrices -----------------------------------------------------

    def register(self, name, A, kind):
        """Keep A (column-major Float32) resident for `kind` requests."""
        if kind not in ("rangefinder", "kaczmarz"):
            raise ArgumentError("kind must be 'rangefinder' or 'kaczmarz'")
        v = as_view(A)
        if v.rs != 1 or A.dtype != torch.float32:
            raise ArgumentError("register a column-major Float32 matrix")
        r = _Resident(name, A, kind)
        with torch.cuda.device(self.device):
            if kind == "rangefinder":
                r.bf16 = _Bf16Copies()
                rm, cm = r.bf16.get(A)
                r.bytes += rm.numel() * 2 + cm.numel() * 2
            else:
                r.A_rm = torch.empty(v.rows, v.cols, dtype=A.dtype, device=A.device)
                r.A_rm.copy_(A)
                w = torch.empty(v.rows, dtype=torch.float64, device=A.device)
                _lib.call("rla_row_norms2_f32", _lib.ctx(A.device), v.rows, v.cols,
                          _lib.ptr(A), 1, v.cs, _lib.ptr(w))
                r.weights = w
                r.key = _dense.SourceKey(A)  # what A_rm and the weights were made from
                r.bytes += r.A_rm.numel() * 4 + w.numel() * 8
            torch.cuda.current_stream(self.device).synchronize()
        with self._cv:
            if _lib.call_rc("rla_hbm_admit", self._dev, r.bytes) != _lib.RLA_OK:
                raise AdmissionError(f"matrix {name!r} ({r.bytes} B) exceeds the HBM budget: "
                                     + _lib.last_error())
            self._resident_bytes += r.bytes
            self.resident[name] = r

    # -- requests ------------------------------------------------------------------

    def _need(self, req):
        r = self.resident[req.matrix]
        m, n = r.A.shape
        if req.kind == "rangefinder":
            p = PRECISIONS[req.precision]
            ws = _dense.query_bytes("rla_rangefinder_workspace_bytes", p, 4, m, n, req.k)
            return ws + 8 * m * req.k * 3 + 4 * n * req.k * 2 + _SLOT_SCRATCH
        ws = _dense.query_bytes("rla_kaczmarz_workspace_bytes", m, n, req.block, req.every)
        return ws + 4 * n * 2 + _SLOT_SCRATCH

    def submit(self, req):
        """Queue a request; returns a Future resolving to its result."""
        if req.matrix not in self.resident:
            raise ArgumentError(f"no resident matrix {req.matrix!r}")
        if self.resident[req.matrix].kind != req.kind:
            raise ArgumentError(f"matrix {req.matrix!r} is not registered for {req.kind}")
        need = self._need(req)
        resident = self._resident_bytes
        if need > self.budget - resident:
            raise AdmissionError(f"request needs {need} B, more than the budget leaves "
                                 f"({self.budget - resident} B)")
        fut = Future()
        if req.arrival is None:
            req.arrival = time.perf_counter()
        with self._cv:
            self._queue.append((req, need, fut))
            self._cv.notify_all()
        return fut

    def _worker(self):
        slot = threading.current_thread().slot
        stream = self._streams[slot]
        while True:
            with self._cv:
                while True:
                    if self._closed and not self._queue:
                        _lib.drop_thread_contexts()
                        return
                    if self._queue:
                        if _lib.call_rc("rla_hbm_admit", self._dev,
                                        self._queue[0][1]) == _lib.RLA_OK:
                            req, need, fut = self._queue.popleft()
                            self.in_flight += 1
                            self.max_in_flight = max(self.max_in_flight, self.in_flight)
                            break
                        if self.in_flight == 0:
                            # nothing in flight will free memory: the request
                            # can never be admitted
                            _, _, dead = self._queue.popleft()
                            dead.set_exception(AdmissionError(_lib.last_error()))
                            continue
                    self._cv.wait()
            t_admit = time.perf_counter()
            try:
                with torch.cuda.device(self.device), torch.cuda.stream(stream):
                    out = self._run(req)
                    stream.synchronize()
                t_done = time.perf_counter()
                self.records.append({"kind": req.kind, "matrix": req.matrix,
                                     "arrival": req.arrival, "admit": t_admit, "done": t_done,
                                     "latency": t_done - req.arrival, "slot": slot})
                fut.set_result(out)
            except Exception as e:  # noqa: BLE001 — delivered to the caller
                fut.set_exception(e)
            finally:
                with self._cv:
                    _lib.call("rla_hbm_release", self._dev, need)
                    self.in_flight -= 1
                    self._cv.notify_all()

    def _run(self, req):
        r = self.resident[req.matrix]
        A = r.A
        if req.kind == "rangefinder":
            with _setup_lock:
                seed(req.seed)
                rec = complete_approximator(
                    RangeFinder(compressor=Gaussian(cardinality=Right(), compression_dim=req.k,
                                                    type=torch.float32),
                                power_its=req.power_its, orthogonalize=req.orthogonalize,
                                precision=req.precision), A)
            rec._bf16 = r.bf16
            rapproximate_(rec, A)
            return rec.range
        b = req.b
        bn = float(torch.linalg.vector_norm(b.double()))
        x = torch.zeros(A.shape[1], dtype=torch.float32, device=A.device)
        solver = Kaczmarz(compressor=Sampling(compression_dim=req.block, distribution=L2Norm()),
                          log=BasicLogger(max_it=req.max_it, threshold=req.rtol * bn),
                          error=FullResidual(every=req.every))
        with _setup_lock:
            seed(req.seed)
            rec = complete_solver(solver, x, A, b)
        rec.rng = (req.seed, 0)
        rec.A_rm, rec._rm_key = r.A_rm, None
        d = rec.compressor.distribution_recipe
        d.weights, d._key = r.weights, r.key
        rec._rm_key = r.key
        rsolve(rec, x, A, b)
        return x, rec.log

    # -- reporting -------------------------------------------------------------------

    def stats(self):
        if not self.records:
            return {}
        lat = np.array([r["latency"] for r in self.records])
        t0 = min(r["arrival"] for r in self.records)
        t1 = max(r["done"] for r in self.records)
        out = {"requests": len(lat), "p50_ms": float(np.percentile(lat, 50) * 1e3),
               "p99_ms": float(np.percentile(lat, 99) * 1e3), "solves_per_s": len(lat) / (t1 - t0)}
        for kind in ("rangefinder", "kaczmarz"):
            lk = np.array([r["latency"] for r in self.records if r["kind"] == kind])
            if len(lk):
                out[kind] = {"requests": len(lk), "p50_ms": float(np.percentile(lk, 50) * 1e3),
                             "p99_ms": float(np.percentile(lk, 99) * 1e3)}
        return out

    def close(self):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for w in self._workers:
            w.join()
        with self._cv:
            if self._resident_bytes:
                _lib.call("rla_hbm_release", self._dev, self._resident_bytes)
                self._resident_bytes = 0
            if self._prev_budget is not None:
                _lib.call("rla_hbm_set_budget", self._dev, self._prev_budget)
                self._prev_budget = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def poisson_arrivals(qps, n, rng):
    """n arrival offsets (seconds) of a Poisson process with rate qps."""
    return np.cumsum(rng.exponential(1.0 / qps, n))
