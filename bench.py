"""Benchmark of the sketching hot path on one MI355X.

Headline (one step = one pass of the hot path over one batch): C = S*A with
S an SRHT (Left) recipe, A 65536 x 512 Float32 N(0,1) (seed 1234), 2048
sampled rows, scale sqrt(2048/65536) as the reference stores it.  Inputs are
resident in HBM; steps rotate over 4 copies of A and C (512 MiB + 16 MiB, more
than the 256 MB Infinity Cache) so every step streams A from HBM.

Prints one JSON line: value = whole-job throughput in algorithmic GB/s
((65536*512 + 2048*512) * 4 bytes per step / time) of the K steps issued as
independent applies alternating over two HIP streams and replayed as one HIP
graph (an apply's workgroups then start on the CUs the previous apply's last
workgroups free, instead of waiting for its whole grid to drain); beside it
the one-stream per-apply time, the roofline of the dominant kernel (its
per-launch duration, HIP events on the stream it runs on, one stream) and the
CPU baseline (the oracle's C restatement, all host threads).  Nested
blocks report the other workloads: the SparseSign 8192x256 Float64 apply,
the mixed-precision vs plain RangeFinder 32768x8192 (solves/s and FP64 error),
512 batched sketch-and-solve least-squares problems, the 100000x1000 block
Kaczmarz solve, and a served Poisson stream of RangeFinder and Kaczmarz
requests (p50/p99 latency and solves/s, served vs a plain single queue).
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SPARSE_C1_ARGS = []  # the C1 recipe and one A, for its CPU baseline
M, N, S_DIM = 65536, 512, 2048
COPIES = 4


def srht_bytes():
    return (M * N + S_DIM * N) * 4


def gpu_srht(rla, torch, steps, warmup):
    """The headline apply on the default Float32 path (the pruned kernel,
    within 1e-5 of the reference) and, beside it, the bit-exact entry point
    (the reference's operation order).  Returns ms per step of each and the
    parity figures of the last step's first 4 columns."""
    import numpy as np
    from oracle import restate as R
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(1234)
    As = [torch.randn(N, M, generator=g, device=dev, dtype=torch.float32).T
          for _ in range(COPIES)]  # column-major M x N
    Cs = [torch.zeros(N, S_DIM, device=dev, dtype=torch.float32).T for _ in range(COPIES)]
    rla.seed(1234)
    S = rla.complete_compressor(rla.SRHT(compression_dim=S_DIM, type=torch.float32), As[0])
    stream = torch.cuda.current_stream()

    def timed(exact):
        S.exact = exact
        for i in range(warmup):
            rla.mul(Cs[i % COPIES], S, As[i % COPIES])
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        t0 = time.perf_counter()
        ev[0].record(stream)
        for i in range(steps):
            rla.mul(Cs[i % COPIES], S, As[i % COPIES])
        ev[1].record(stream)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / steps, time.perf_counter() - t0

    def replayed(exact, nstreams):
        """The steps captured as one graph and replayed: on one stream (the
        kernels back to back: per-launch kernel time without the host's call
        path), or as independent applies alternating over two streams (fork
        from and join to the capture stream).  Every replay's outputs are
        checked against the eager ones.  Returns ms per step, and whether
        every replay matched."""
        S.exact = exact
        want = [c.clone() for c in Cs]
        cur = torch.cuda.current_stream()
        cap = torch.cuda.Stream(device=dev)
        sides = [torch.cuda.Stream(device=dev) for _ in range(nstreams)]
        # every stream's library context exists before the capture (one per
        # stream; creating one allocates)
        for sd in sides + [cap]:
            sd.wait_stream(cur)
            with torch.cuda.stream(sd):
                rla.mul(Cs[0], S, As[0])
            cur.wait_stream(sd)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=cap):
            for sd in sides:
                sd.wait_stream(cap)
            for i in range(steps):
                with torch.cuda.stream(sides[i % nstreams]):
                    rla.mul(Cs[i % COPIES], S, As[i % COPIES])
            for sd in sides:
                cap.wait_stream(sd)
        ok = True
        for _ in range(3):
            for c in Cs:
                c.zero_()
            g.replay()
            torch.cuda.synchronize()
            ok = ok and all(bool(torch.equal(c, w)) for c, w in zip(Cs, want))
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(cur)
        g.replay()
        ev[1].record(cur)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / steps, ok

    k = (steps - 1) % COPIES
    A4 = As[k][:, :4].cpu().numpy()
    op, sg = S.op.cpu().numpy(), S.sign_vector()
    ms_exact, _ = timed(True)
    bit_exact = bool(np.array_equal(Cs[k][:, :4].cpu().numpy(),
                                    R.srht_left_mul(A4, op, sg, S.scale, np.float32)))
    ms_eager, wall = timed(False)
    got = Cs[k][:, :4].cpu().numpy().astype(np.float64)
    want = R.srht_left_mul(A4.astype(np.float64), op, sg, S.scale, np.float64)
    rel = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    ms_one, one_ok = replayed(False, 1)
    ms, two_ok = replayed(False, 2)
    parity = {"rel_error_vs_fp64_oracle": rel, "within_1e-5": rel <= 1e-5,
              "graph_replays_equal_eager": one_ok and two_ok,
              "eager_one_stream_ms_per_step": ms_eager,
              "exact_path_ms_per_step": ms_exact,
              "exact_path_GB_s": srht_bytes() / ms_exact / 1e6,
              "exact_path_bit_exact_vs_oracle": bit_exact}
    return ms, ms_one, wall, parity, S, As[0]


def gpu_sparse_sign(rla, torch, steps, warmup):
    """C1: SparseSign 8192x256 Float64 -> 1024x256, nnz 8, rotating 32 copies."""
    dev = torch.device("cuda", 0)
    m, n, s, copies = 8192, 256, 1024, 32
    g = torch.Generator(device=dev).manual_seed(1234)
    As = [torch.randn(n, m, generator=g, device=dev, dtype=torch.float64).T
          for _ in range(copies)]
    Cs = [torch.zeros(n, s, device=dev, dtype=torch.float64).T for _ in range(copies)]
    rla.seed(1234)
    S = rla.complete_compressor(rla.SparseSign(compression_dim=s, nnz=8), As[0])
    for i in range(max(warmup, 1)):
        rla.mul(Cs[i % copies], S, As[i % copies])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(steps, copies)
    e0.record()
    for i in range(reps):
        rla.mul(Cs[i % copies], S, As[i % copies])
    e1.record()
    torch.cuda.synchronize()
    ms_eager = e0.elapsed_time(e1) / reps
    want = Cs[0].clone()
    # the 32 applies captured in one graph and replayed: one apply's kernel
    # time without the per-call host path (a sub-20 us kernel is otherwise
    # paced by the Python call)
    ms_graph, graph_exact = None, None
    try:
        # the capture stream's library context exists before the capture
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for i in range(copies):
                rla.mul(Cs[i], S, As[i])
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            for i in range(copies):
                rla.mul(Cs[i], S, As[i])
        graph_exact = True
        for _ in range(3):
            Cs[0].zero_()
            g.replay()
            torch.cuda.synchronize()
            graph_exact = graph_exact and bool(torch.equal(Cs[0], want))
        nrep = max(2, steps // copies)
        e0.record()
        for _ in range(nrep):
            g.replay()
        e1.record()
        torch.cuda.synchronize()
        ms_graph = e0.elapsed_time(e1) / (nrep * copies)
        # the same 32 applies alternating over two (and four) streams in one
        # graph: an apply's workgroups start on the CUs the previous one frees
        def over_streams(nst):
            nonlocal graph_exact
            sides = [torch.cuda.Stream() for _ in range(nst)]
            for sd in sides:
                sd.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(sd):
                    rla.mul(Cs[0], S, As[0])
                torch.cuda.current_stream().wait_stream(sd)
            torch.cuda.synchronize()
            gn = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gn, stream=side):
                for sd in sides:
                    sd.wait_stream(side)
                for i in range(copies):
                    with torch.cuda.stream(sides[i % nst]):
                        rla.mul(Cs[i], S, As[i])
                for sd in sides:
                    side.wait_stream(sd)
            for _ in range(3):
                Cs[0].zero_()
                gn.replay()
                torch.cuda.synchronize()
                graph_exact = graph_exact and bool(torch.equal(Cs[0], want))
            e0.record()
            for _ in range(nrep):
                gn.replay()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / (nrep * copies)

        ms_graph2 = over_streams(2)
        ms_graph4 = over_streams(4)
    except RuntimeError as exc:  # capture unavailable: report the per-call figure only
        print(f"sparse sign graph capture failed: {exc}", file=sys.stderr)
        ms_graph2 = ms_graph4 = None
    ok = ms_graph is not None and graph_exact
    multi = [v for v in (ms_graph2, ms_graph4) if v is not None]
    ms = min(multi) if ok and multi else (ms_graph if ok else ms_eager)
    SPARSE_C1_ARGS[:] = [S, As[0]]
    byts = (m * n + s * n) * 8
    return {"workload": "SparseSign left apply 8192x256 f64 -> 1024x256, nnz 8",
            "ms_per_apply": ms, "GB_s": byts / ms / 1e6,
            "frac_hbm_peak": byts / ms / 1e6 / HBM_PEAK_GBS, "rotating_copies": copies,
            "timing": ("graph replay of 32 applies over 2 streams" if ms == ms_graph2 else
                       "graph replay of 32 applies over 4 streams" if ms == ms_graph4 else
                       "graph replay of 32 applies" if ms == ms_graph else "per-call loop"),
            "ms_per_apply_per_call_loop": ms_eager, "ms_per_apply_graph": ms_graph,
            "ms_per_apply_graph_two_streams": ms_graph2,
            "ms_per_apply_graph_four_streams": ms_graph4,
            "kernel_GB_s": byts / ms_graph / 1e6 if ms_graph else None,
            "graph_replays_equal_eager": graph_exact}


def fwht_fjlt_rows(rla, torch, reps=10):
    """The FWHT alone (fwht!, helpers/fwht.jl) on 65536x512 f32 in place, and
    the FJLT left apply (fjlt.jl, default sparsity 0.25 log^2(n)/n) of the same
    A to 2048 rows: per-call time and the algorithmic GB/s (one read of X or A,
    one write of X or C)."""
    dev = torch.device("cuda", 0)
    m, n, s = 65536, 512, 2048
    g = torch.Generator(device=dev).manual_seed(1234)
    X = torch.randn(n, m, generator=g, device=dev, dtype=torch.float32).T
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rla.fwht(X)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        rla.fwht(X)
    e1.record()
    torch.cuda.synchronize()
    ms_fwht = e0.elapsed_time(e1) / reps
    A = torch.randn(n, m, generator=g, device=dev, dtype=torch.float32).T
    C = torch.zeros(n, s, device=dev, dtype=torch.float32).T
    rla.seed(1234)
    S = rla.complete_compressor(rla.FJLT(compression_dim=s, type=torch.float32), A)
    rla.mul(C, S, A)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        rla.mul(C, S, A)
    e1.record()
    torch.cuda.synchronize()
    ms_fjlt = e0.elapsed_time(e1) / reps
    fb, jb = 2 * m * n * 4, (m * n + s * n) * 4
    return {"fwht": {"workload": "fwht! 65536x512 f32 in place", "ms_per_call": ms_fwht,
                     "GB_s": fb / ms_fwht / 1e6, "frac_hbm_peak": fb / ms_fwht / 1e6 / HBM_PEAK_GBS},
            "fjlt": {"workload": f"FJLT left apply 65536x512 f32 -> 2048x512, q = {S.sparsity:.3g}",
                     "ms_per_apply": ms_fjlt, "GB_s": jb / ms_fjlt / 1e6,
                     "frac_hbm_peak": jb / ms_fjlt / 1e6 / HBM_PEAK_GBS}}


def tall_sparse_sign_row(rla, torch, reps=10):
    """SparseSign left apply (sparse_sign.jl:339-361) of a tall A, 1000000x64
    f32 to 1024 rows, nnz = 8: the columns of A exceed one workgroup's LDS, so
    the gather walks A's 64 KiB row blocks through the block-interleaved copy
    of the operator.  Per-apply time and the algorithmic GB/s (one read of A,
    one write of C)."""
    dev = torch.device("cuda", 0)
    m, n, s = 1_000_000, 64, 1024
    g = torch.Generator(device=dev).manual_seed(1234)
    A = torch.randn(n, m, generator=g, device=dev, dtype=torch.float32).T
    C = torch.zeros(n, s, device=dev, dtype=torch.float32).T
    rla.seed(1234)
    S = rla.complete_compressor(rla.SparseSign(compression_dim=s, nnz=8, type=torch.float32), A)
    rla.mul(C, S, A)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        rla.mul(C, S, A)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    b = (m * n + s * n) * 4
    return {"workload": "SparseSign left apply 1000000x64 f32 -> 1024x64, nnz 8",
            "ms_per_apply": ms, "GB_s": b / ms / 1e6, "frac_hbm_peak": b / ms / 1e6 / HBM_PEAK_GBS}


def bf16_gemm_c3(rla, torch, reps=20):
    """The two BF16 MFMA products of a C3 RangeFinder solve on their own
    (Y = A Z: 32768 x 266 over K = 8192 from a copy of A; Z = A'Y: 8192 x 266
    over K = 32768 from a copy of A'), N(0,1) operands, with A's copy in the
    row layout and in the tiled layout the RangeFinder keeps resident; wall
    time over `reps` back-to-back products between device synchronisations."""
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(1234)
    out = {"workload": "BF16 GEMM at the RangeFinder 32768x8192, k=266 product shapes"}
    lib = rla._lib
    for name, M, K in (("A_Z", 32768, 8192), ("At_Y", 8192, 32768)):
        N = 266
        Af = torch.randn(M, K, generator=g, device=dev)
        A = Af.to(torch.bfloat16)  # row i = K-run
        B = torch.randn(N, K, generator=g, device=dev).to(torch.bfloat16)
        C = torch.empty(N, M, device=dev)  # column-major M x N
        ctx = lib.ctx(dev)

        At = rla._dense.bf16_tiled(rla.as_view(Af))   # the same A, tiled
        del Af

        def run_rows():
            lib.call("rla_gemm_bf16", ctx, M, N, K, 1.0, lib.ptr(A), K, lib.ptr(B), K, 0.0,
                     lib.ptr(C), 1, M)

        def run_tiled():
            lib.call("rla_gemm_bf16_tiled", ctx, M, N, K, 1.0, lib.ptr(At), lib.ptr(B), K, 0.0,
                     lib.ptr(C), 1, M)
        for layout, run in (("tiled", run_tiled), ("rows", run_rows)):
            for _ in range(3):
                run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                run()
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / reps
            flops = 2.0 * M * N * K
            out[f"{name}_{layout}"] = {"M": M, "N": N, "K": K, "us_per_product": dt * 1e6,
                                       "TFLOP_s": flops / dt / 1e12,
                                       "frac_bf16_peak": flops / dt / 2.5e15,
                                       "a_bytes_GB_s": M * K * 2 / dt / 1e9}
        del A, At, B, C
    return out


def _in_flight(rla, torch, dev, recs, A, n_solves):
    """Seconds for n_solves solves of every recipe in recs, one host thread
    and HIP stream per recipe, all warmed up before the clock starts."""
    streams = [torch.cuda.Stream(device=dev) for _ in recs]
    go = threading.Barrier(len(recs) + 1)

    def work(r, s):
        with torch.cuda.stream(s):
            rla.rapproximate_(r, A)   # this thread's context and workspaces
            s.synchronize()
            go.wait()
            for _ in range(n_solves):
                rla.rapproximate_(r, A)
            s.synchronize()

    threads = [threading.Thread(target=work, args=(r, s)) for r, s in zip(recs, streams)]
    for t in threads:
        t.start()
    go.wait()   # every thread warmed up; the timed solves start together
    t0 = time.perf_counter()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def rangefinder_c3(rla, torch, reps=3, batch=4):
    """RangeFinder mixed precision vs the plain FP32 path (C3): A = U diag(s) V'
    32768 x 8192, s_i = exp(-i/64) (singular values below 1e-17 s_1 dropped),
    Haar U, V from QR of Gaussians (seed 1234); Gaussian sketch k = 256 + 10,
    2 orthogonalised power iterations.  Reports solves/s with A (and its BF16
    copies) resident, one solve at a time and `batch` solves in flight on
    their own streams, and ||A - QQ'A||_F / ||A||_F in FP64."""
    import math
    dev = torch.device("cuda", 0)
    m, n, k, r = 32768, 8192, 266, 2560
    g = torch.Generator(device=dev).manual_seed(1234)
    U, _ = torch.linalg.qr(torch.randn(m, r, generator=g, device=dev, dtype=torch.float64))
    V, _ = torch.linalg.qr(torch.randn(n, r, generator=g, device=dev, dtype=torch.float64))
    sig = torch.exp(-torch.arange(r, device=dev, dtype=torch.float64) / 64.0)
    A64 = ((U * sig) @ V.T).T.contiguous().T
    del U, V
    A32 = A64.float().T.contiguous().T
    full = torch.exp(-torch.arange(n, device=dev, dtype=torch.float64) / 64.0)
    opt = float(torch.sqrt((full[k:] ** 2).sum() / (full ** 2).sum()))
    nA = float(torch.linalg.norm(A64))
    out = {"workload": "RangeFinder 32768x8192, k=266 (256+10), q=2, orthogonalize, Gaussian",
           "eckart_young_optimum": opt}
    for prec in ("bf16", "plain"):
        rla.seed(7)
        rf = rla.RangeFinder(compressor=rla.Gaussian(cardinality=rla.Right(), compression_dim=k,
                                                     type=torch.float32),
                             power_its=2, orthogonalize=True, precision=prec)
        rec = rla.complete_approximator(rf, A32)
        rla.rapproximate_(rec, A32)   # warm-up (and BF16 copies of A for the bf16 path)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            rla.rapproximate_(rec, A32)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        Q = rec.range.double()
        B = torch.zeros(n, k, dtype=torch.float64, device=dev).T
        rla._dense.gemm(rla.as_view(B), rla.as_view(Q).T, rla.as_view(A64), 1.0, 0.0)
        R_ = A64.clone()
        rla._dense.gemm(rla.as_view(R_), rla.as_view(Q), rla.as_view(B), -1.0, 1.0)
        err = float(torch.linalg.norm(R_)) / nA
        out[prec] = {"seconds_per_solve": dt, "solves_per_s": 1.0 / dt, "rel_error": err}
        del Q, B, R_
        # a batch of solves in flight together: one host thread and HIP stream
        # per solve, A and its BF16 copies shared, every solve its own sketch
        recs = [rec]
        for i in range(1, batch):
            rla.seed(7 + i)
            recs.append(rla.complete_approximator(rf, A32))
            recs[-1]._bf16 = rec._bf16
        n_batch = 3 * reps   # solves per stream in the timed batch
        bdt = _in_flight(rla, torch, dev, recs, A32, n_batch)
        out[prec]["batch"] = {"in_flight": batch, "solves": batch * n_batch,
                              "solves_per_s": batch * n_batch / bdt}
        del recs, rec
    # RandSVD (randsvd.jl:174-189) on the same A, BF16 products: the range, then
    # B = Q'A and its FP64 SVD; singular values against the planted ones
    rla.seed(7)
    sv = rla.RandSVD(compressor=rla.Gaussian(cardinality=rla.Right(), compression_dim=k,
                                             type=torch.float32),
                     power_its=2, orthogonalize=True, precision="bf16")
    srec = rla.complete_approximator(sv, A32)
    rla.rapproximate_(srec, A32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        rla.rapproximate_(srec, A32)
    torch.cuda.synchronize()
    sdt = (time.perf_counter() - t0) / reps
    top = 128
    sv_err = float((srec.S[:top] / full[:top] - 1.0).abs().max())
    out["randsvd_bf16"] = {"seconds_per_solve": sdt, "solves_per_s": 1.0 / sdt,
                           f"max_rel_error_top{top}_singular_values": sv_err}
    # RandSVD solves in flight as for the RangeFinder: the Jacobi SVD of one
    # solve's small factor holds few CUs, the other solves' products fill the rest
    srecs = [srec]
    for i in range(1, batch):
        rla.seed(7 + i)
        srecs.append(rla.complete_approximator(sv, A32))
        srecs[-1]._bf16 = srec._bf16
    n_batch = 3 * reps
    bdt = _in_flight(rla, torch, dev, srecs, A32, n_batch)
    out["randsvd_bf16"]["batch"] = {"in_flight": batch, "solves": batch * n_batch,
                                    "solves_per_s": batch * n_batch / bdt}
    out["randsvd_over_rangefinder_seconds_per_solve"] = sdt / out["bf16"]["seconds_per_solve"]
    out["randsvd_over_rangefinder_in_flight"] = (out["bf16"]["batch"]["solves_per_s"] /
                                                 out["randsvd_bf16"]["batch"]["solves_per_s"])
    del srecs, srec
    out["speedup_solves_per_s"] = out["plain"]["seconds_per_solve"] / out["bf16"]["seconds_per_solve"]
    out["batched_speedup_solves_per_s"] = (out["bf16"]["batch"]["solves_per_s"] /
                                           out["plain"]["batch"]["solves_per_s"])
    out["error_ratio_bf16_over_plain"] = out["bf16"]["rel_error"] / out["plain"]["rel_error"]
    return out


def ls_batched_c4(rla, torch, reps=5):
    """Batched sketch-and-solve least squares: 512 problems, m log-uniform in
    [2000, 50000], n uniform in [32, 256], N(0,1) columns scaled by 10^U(-3,0),
    b = A x* + 1e-6 noise (seed 1234); CountSketch to 4n rows, then the FP64
    normal-equation solve of the sketched problem, all one hipGraph replay.
    Reports solves/s over timed replays, the sketch kernel's algorithmic GB/s
    and ||x - x*|| / ||x*|| per problem."""
    import numpy as np
    dev = torch.device("cuda", 0)
    F32 = torch.float32
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
    batch.run()  # warm-up
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        batch.run()
    e1.record()
    torch.cuda.synchronize()
    ms_batch = e0.elapsed_time(e1) / reps
    info = batch.info()
    err = np.array([float(torch.linalg.vector_norm((x - xsr).double()) /
                          torch.linalg.vector_norm(xsr.double())) for x, xsr in zip(xs, xstars)])
    cfg = batch.config()
    sketch_ms = min(t for t in cfg["tuned_ms"] if t > 0)
    # sketch kernel: read A and b, the row's bucket and sign, write [SA Sb]
    sbytes = int(sum(m * (n + 1) * 4 + m * 8 + 4 * n * (n + 1) * 4 for m, n in zip(ms, ns)))
    batch.close()
    return {"workload": "512 CountSketch+solve LS problems, m in [2000,50000] log-uniform, "
                        "n in [32,256], s = 4n, f32 data, FP64 solve",
            "ms_per_batch": ms_batch, "solves_per_s": 512 / ms_batch * 1e3,
            "sketch_kernel_ms": sketch_ms, "sketch_GB_s": sbytes / sketch_ms / 1e6,
            "sketch_frac_hbm_peak": sbytes / sketch_ms / 1e6 / HBM_PEAK_GBS,
            "tile_config": {"cols_per_tile": cfg["cols_per_tile"],
                            "rows_per_chunk": cfg["rows_per_chunk"],
                            "bucket": batch.tile_key, "from_tile_cache": batch.tile_cached},
            "breakdowns": int(sum(1 for i in info if i)),
            "rel_error_median": float(np.median(err)), "rel_error_max": float(err.max())}


def kaczmarz_c5(rla, torch, reps=3):
    """Block Kaczmarz on a consistent 100000 x 1000 N(0,1) system, b = A x*
    (seed 1234): 64-row blocks drawn by squared row norm, FP32 updates, FP64
    residual every 100 iterations, stop at ||r|| <= 1e-6 ||b||."""
    dev = torch.device("cuda", 0)
    F32 = torch.float32
    m, n, s = 100000, 1000, 64
    g = torch.Generator(device=dev).manual_seed(1234)
    A = torch.randn(n, m, generator=g, device=dev, dtype=F32).T
    xs = torch.randn(n, generator=g, device=dev, dtype=F32)
    b = A @ xs
    bn = float(torch.linalg.vector_norm(b.double()))
    solver = rla.Kaczmarz(compressor=rla.Sampling(compression_dim=s, distribution=rla.L2Norm()),
                          log=rla.BasicLogger(max_it=20000, threshold=1e-6 * bn),
                          error=rla.FullResidual(every=100))
    x = torch.zeros(n, dtype=F32, device=dev)
    rla.seed(1234)
    x, rec = rla.rsolve(solver, x, A, b)  # warm-up: row-major copy and row norms
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = 0
    for _ in range(reps):
        x.zero_()
        rla.rsolve(rec, x, A, b)
        iters += rec.log.iteration
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    it = iters / reps
    rel = float(torch.linalg.vector_norm((x - xs).double()) / torch.linalg.vector_norm(xs.double()))
    # per iteration: gather 64 rows (row-major, 64*1000*4 B), read and write x
    byts = s * n * 4 + 2 * n * 4
    return {"workload": "block Kaczmarz 100000x1000 f32, 64-row L2Norm blocks, FP64 residual "
                        "every 100 iterations, to ||r||/||b|| <= 1e-6",
            "seconds_per_solve": dt, "solves_per_s": 1.0 / dt, "iterations": it,
            "us_per_iteration": dt / it * 1e6, "gather_GB_s": byts * it / dt / 1e9,
            "converged": bool(rec.log.converged), "rel_residual": rec.log.error / bn,
            "rel_error_x": rel}


def lowrank(torch, m, n, r, decay, g, dev):
    U = torch.randn(r, m, generator=g, device=dev, dtype=torch.float32).T
    V = torch.randn(n, r, generator=g, device=dev, dtype=torch.float32)
    s = torch.exp(-torch.arange(r, device=dev, dtype=torch.float32) / decay)
    return ((U * s) @ V.T / n ** 0.5).T.contiguous().T


def serving_c5(rla, torch, qps=80.0, n_req=96):
    """Served mixed stream: Poisson arrivals at `qps` mixing RangeFinder
    requests (k = 266, q = 2, on resident low-rank matrices 4096x1024,
    8192x2048, 16384x4096, 32768x8192 with sigma_i = exp(-i/64)) and block
    Kaczmarz solves of a resident 100000x1000 system (64-row L2Norm blocks,
    FP64 residual every 100, rtol 1e-6).  Run three times on the same
    arrival trace: mixed precision with 4 slots (continuous batching), the
    plain FP32 path with 4 slots (batching alone), and the plain path with one
    slot (a queue).  At 80 QPS the single queue saturates."""
    import numpy as np
    dev = torch.device("cuda", 0)
    F32 = torch.float32
    g = torch.Generator(device=dev).manual_seed(1234)
    shapes = [(4096, 1024), (8192, 2048), (16384, 4096), (32768, 8192)]
    mats = {f"rf{m}x{n}": lowrank(torch, m, n, 1024, 64.0, g, dev) for m, n in shapes}
    Akz = torch.randn(1000, 100000, generator=g, device=dev, dtype=F32).T
    bkz = [Akz @ torch.randn(1000, generator=g, device=dev, dtype=F32) for _ in range(4)]
    rng = np.random.default_rng(1234)
    kinds = rng.integers(0, 2, n_req)
    which = rng.integers(0, len(shapes), n_req)
    offs = rla.poisson_arrivals(qps, n_req, rng)
    out = {"workload": f"Poisson {qps:g} QPS, {n_req} requests, RangeFinder 4096..32768 x "
                       "1024..8192 (k=266, q=2) and Kaczmarz 100000x1000 (s=64, rtol 1e-6)",
           "rangefinder_requests": int(n_req - kinds.sum()), "kaczmarz_requests": int(kinds.sum())}
    for label, prec, slots in (("served_bf16_4slots", "bf16", 4),
                               ("served_plain_4slots", "plain", 4),
                               ("queue_plain_1slot", "plain", 1)):
        with rla.Server(device=dev, slots=slots) as srv:
            for name, A in mats.items():
                srv.register(name, A, "rangefinder")
            srv.register("kz", Akz, "kaczmarz")
            # warm-up: one request of each kind and size
            for name in mats:
                srv.submit(rla.RangeFinderRequest(name, k=266, precision=prec, seed=1)).result()
            srv.submit(rla.KaczmarzRequest("kz", bkz[0], seed=1)).result()
            srv.records.clear()
            futs = []
            t0 = time.perf_counter()
            for i in range(n_req):
                while time.perf_counter() < t0 + offs[i]:
                    time.sleep(0.0002)
                if kinds[i]:
                    req = rla.KaczmarzRequest("kz", bkz[i % 4], seed=1000 + i)
                else:
                    name = f"rf{shapes[which[i]][0]}x{shapes[which[i]][1]}"
                    req = rla.RangeFinderRequest(name, k=266, precision=prec, seed=1000 + i)
                futs.append(srv.submit(req))
            res = [f.result() for f in futs]
            st = srv.stats()
            st["max_in_flight"] = srv.max_in_flight
            st["kaczmarz_all_converged"] = all(r[1].converged for r, k in zip(res, kinds) if k)
            out[label] = st
            del res, futs
    # one variable at a time: continuous batching (4 slots vs a 1-slot queue,
    # both plain), then precision (BF16 vs plain products, both 4 slots);
    # the last pair changes both
    q, sp, sb = out["queue_plain_1slot"], out["served_plain_4slots"], out["served_bf16_4slots"]
    out["batching_p99_ratio_queue_over_served_plain"] = q["p99_ms"] / sp["p99_ms"]
    out["batching_solves_per_s_ratio"] = sp["solves_per_s"] / q["solves_per_s"]
    out["precision_p99_ratio_plain_over_bf16_4slots"] = sp["p99_ms"] / sb["p99_ms"]
    out["precision_solves_per_s_ratio_4slots"] = sb["solves_per_s"] / sp["solves_per_s"]
    out["precision_rangefinder_p99_ratio_4slots"] = (sp["rangefinder"]["p99_ms"] /
                                                      sb["rangefinder"]["p99_ms"])
    out["p99_ratio_plain_over_served"] = q["p99_ms"] / sb["p99_ms"]
    out["solves_per_s_ratio_served_over_plain"] = sb["solves_per_s"] / q["solves_per_s"]
    return out


def _host_threads():
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()


def _timed_loop(fn, seconds):
    """Runs fn until `seconds` have passed (at least once); mean seconds per call."""
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            return dt / reps, reps


def cpu_c1(S, A, seconds=4.0):
    """C1 on the host cores: the oracle's C restatement of SparseArrays' CSC
    scatter loop behind sparse_sign.jl:347 (OpenMP over columns), the full
    8192x256 Float64 apply with the same recipe."""
    import ctypes

    import numpy as np
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    vp, i64, f64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
    lib.oracle_sparse_left_f64.argtypes = [vp, vp, vp, i64, i64, vp, i64, i64, vp, i64, f64, f64,
                                           i32]
    op = S.op
    colptr = op.colptr.cpu().numpy().astype(np.int32)
    rowidx = op.rowidx.cpu().numpy().astype(np.int32)
    val = op.val.double().cpu().numpy()
    Ah = np.asfortranarray(A.cpu().numpy())
    m, n = Ah.shape
    C = np.zeros((op.nrows, n), order="F")
    threads = _host_threads()
    per, reps = _timed_loop(lambda: lib.oracle_sparse_left_f64(
        colptr.ctypes.data, rowidx.ctypes.data, val.ctypes.data, op.nrows, m, Ah.ctypes.data, n,
        m, C.ctypes.data, op.nrows, 1.0, 0.0, threads), seconds)
    byts = (m * n + op.nrows * n) * 8
    return {"value": byts / per / 1e9, "unit": "GB/s", "ms_per_apply": per * 1e3,
            "cores": threads, "kind": "port",
            "sample": f"full workload (8192x256 f64 SparseSign apply), {reps} reps, C restatement "
                      "of the SparseArrays CSC scatter loop, OpenMP over columns"}


def cpu_c3(seconds_hint=None):
    """C3 on the host cores: rand_ortho_it (power_its.jl:53-80) restated with
    numpy / its threaded OpenBLAS, FP32 and FP64, one solve each on the
    32768x8192 sigma_i = exp(-i/64) matrix (a host draw of the same
    construction), k = 266, q = 2."""
    import numpy as np
    from oracle import restate as R
    m, n, k, r = 32768, 8192, 266, 2560
    rng = np.random.default_rng(1234)
    U = np.linalg.qr(rng.standard_normal((m, r)))[0]
    V = np.linalg.qr(rng.standard_normal((n, r)))[0]
    A64 = np.asfortranarray((U * np.exp(-np.arange(r) / 64.0)) @ V.T)
    del U, V
    Om64 = rng.standard_normal((n, k)) / np.sqrt(k)
    out = {"cores": _host_threads(), "kind": "port",
           "sample": "one solve per precision, numpy restatement of rand_ortho_it with the "
                     "host BLAS (all threads), Householder QR"}
    for label, A, Om in (("f64", A64, Om64),
                         ("f32", np.asfortranarray(A64.astype(np.float32)),
                          Om64.astype(np.float32))):
        t0 = time.perf_counter()
        Q = R.rand_ortho_it(A, Om, 2)
        dt = time.perf_counter() - t0
        Qd = Q.astype(np.float64)
        err = float(np.linalg.norm(A64 - Qd @ (Qd.T @ A64)) / np.linalg.norm(A64))
        out[label] = {"seconds_per_solve": dt, "solves_per_s": 1.0 / dt, "rel_error": err}
        del Q, Qd
    out["value"] = out["f32"]["solves_per_s"]
    out["unit"] = "solves/s (f32)"
    return out


def cpu_c4(n_sample=64):
    """C4 on the host cores: per problem the CountSketch mul! (scipy.sparse,
    like SparseArrays) and the QRSolver ldiv! (qr.jl:36-43: Householder QR of
    the 4n x n sketch, FP32), over a sample of problems drawn like the 512."""
    import numpy as np
    import scipy.linalg
    import scipy.sparse
    rng = np.random.default_rng(1234)
    ms = np.floor(np.exp(rng.uniform(np.log(2000), np.log(50000), 512))).astype(int)[:n_sample]
    ns = rng.integers(32, 257, 512)[:n_sample]
    probs = []
    for m, n in zip(ms, ns):
        A = (rng.standard_normal((m, n)) * 10.0 ** rng.uniform(-3, 0, n)).astype(np.float32)
        xs = rng.standard_normal(n).astype(np.float32)
        b = A @ xs + np.float32(1e-6) * rng.standard_normal(m).astype(np.float32)
        S = scipy.sparse.csr_matrix((rng.choice([-1.0, 1.0], m).astype(np.float32),
                                     (rng.integers(0, 4 * n, m), np.arange(m))),
                                    shape=(4 * n, m))
        probs.append((np.asfortranarray(A), b, S))
    t0 = time.perf_counter()
    for A, b, S in probs:
        SA, Sb = S @ A, S @ b
        Q, Rr = scipy.linalg.qr(SA, mode="economic")
        scipy.linalg.solve_triangular(Rr, Q.T @ Sb)
    dt = time.perf_counter() - t0
    per = dt / n_sample
    return {"value": 1.0 / per, "unit": "solves/s", "seconds_per_512": per * 512,
            "cores": _host_threads(), "kind": "port",
            "sample": f"{n_sample} of the 512 problem shapes (host draws), scipy.sparse sketch "
                      "and LAPACK Householder QR per problem, extrapolated to solves/s"}


def cpu_c5b(iters=100, written_iters=5):
    """C5b on the host cores: kaczmarz.jl:313-368 restated with numpy on the
    100000x1000 system: 64 rows drawn by squared row norm, the LQ projection
    (lq.jl:35-46) of each block.  Timed with the FP64 residual every 100
    iterations, and as written (residual and L2Norm weights every iteration);
    per-solve figures use the device solve's iteration count (501)."""
    import numpy as np
    import scipy.linalg
    m, n, s = 100000, 1000, 64
    rng = np.random.default_rng(1234)
    A = rng.standard_normal((m, n)).astype(np.float32)
    b = A @ rng.standard_normal(n).astype(np.float32)
    x = np.zeros(n, np.float32)

    def step(weights):
        p = weights / weights.sum()
        rows = np.sort(rng.choice(m, s, replace=False, p=p))
        At, bt = A[rows], b[rows]
        Q, Rr = np.linalg.qr(At.T.astype(np.float64))
        u = Q @ scipy.linalg.solve_triangular(Rr.T, (bt - At @ x).astype(np.float64), lower=True)
        x[:] += u.astype(np.float32)

    # the reference's FullResidual and L2Norm weights are in the data's type
    # (full_residual.jl:32-41, strohmer_vershynin.jl:82)
    w = (A * A).sum(1).astype(np.float64)
    t0 = time.perf_counter()
    for i in range(iters):
        if i % 100 == 0:
            np.linalg.norm(b - A @ x)
        step(w)
    per = (time.perf_counter() - t0) / iters
    t0 = time.perf_counter()
    for _ in range(written_iters):
        np.linalg.norm(b - A @ x)
        w = (A * A).sum(1).astype(np.float64)
        step(w)
    per_written = (time.perf_counter() - t0) / written_iters
    return {"value": 1.0 / (per * 501), "unit": "solves/s", "us_per_iteration": per * 1e6,
            "us_per_iteration_as_written": per_written * 1e6, "cores": _host_threads(),
            "kind": "port",
            "sample": f"{iters} iterations (residual every 100) and {written_iters} as written "
                      "(residual and weights every iteration), numpy restatement with the host "
                      "BLAS; per-solve figures at 501 iterations"}


def cpu_baseline(S, A0):
    """Oracle C restatement (OpenMP over columns, all host threads) on the full
    workload: the same A, op, signs and scale."""
    import ctypes

    import numpy as np
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    vp, i64, f64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
    lib.oracle_srht_left_f32.argtypes = [vp, i64, i64, i64, i64, i64, vp, vp, f64, vp, i64, f64,
                                         f64, i32]
    lib.oracle_max_threads.restype = ctypes.c_int
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    A = np.asfortranarray(A0.cpu().numpy())
    op = S.op.cpu().numpy().astype(np.int32)
    ch = S.signs.cpu().numpy().view(np.uint64).copy()
    C = np.zeros((S_DIM, N), dtype=np.float32, order="F")
    reps, t = 0, 0.0
    while t < 10.0 and reps < 5000:  # about 10 s of host work
        t0 = time.perf_counter()
        lib.oracle_srht_left_f32(A.ctypes.data, M, N, M, M, S_DIM, op.ctypes.data, ch.ctypes.data,
                                 S.scale, C.ctypes.data, S_DIM, 1.0, 0.0, threads)
        t += time.perf_counter() - t0
        reps += 1
    per = t / reps
    return {"value": srht_bytes() / per / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"full workload (65536x512 f32 SRHT apply), {reps} reps, "
                      f"{per * 1e3:.1f} ms each, C restatement of fwht.jl/srht.jl, OpenMP"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (counter calibration)")
    ap.add_argument("--skip-sparse", action="store_true")
    ap.add_argument("--skip-rangefinder", action="store_true")
    ap.add_argument("--skip-ls", action="store_true")
    ap.add_argument("--skip-kaczmarz", action="store_true")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--bf16-gemm-only", action="store_true",
                    help="time only the two BF16 RangeFinder products (counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the SRHT apply "
              "would shard by columns with no exchange). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__
    rla = __graft_entry__.load_package()
    if args.bf16_gemm_only:
        print(json.dumps({"metric": "BF16 GEMM TFLOP/s (RangeFinder product shapes)",
                          "bf16_gemm_c3": bf16_gemm_c3(rla, torch)}))
        return
    ms, ms_one, wall, parity, S, A0 = gpu_srht(rla, torch, args.steps, args.warmup)
    gbs = srht_bytes() / ms / 1e6
    gbs_one = srht_bytes() / ms_one / 1e6  # one stream: the kernel's per-launch time
    extra = None if args.skip_sparse else gpu_sparse_sign(rla, torch, args.steps, args.warmup)
    fwht_fjlt = None if args.skip_sparse else fwht_fjlt_rows(rla, torch)
    tall_sparse = None if args.skip_sparse else tall_sparse_sign_row(rla, torch)
    rf = None if args.skip_rangefinder else rangefinder_c3(rla, torch)
    ls = None if args.skip_ls else ls_batched_c4(rla, torch)
    kz = None if args.skip_kaczmarz else kaczmarz_c5(rla, torch)
    srv = None if args.skip_serving else serving_c5(rla, torch)
    if not args.no_cpu:
        # CPU baselines of the nested workloads (host cores, bounded samples)
        if extra is not None:
            extra["cpu_baseline"] = cpu_c1(*SPARSE_C1_ARGS)
        if rf is not None:
            rf["cpu_baseline"] = cpu_c3()
        if ls is not None:
            ls["cpu_baseline"] = cpu_c4()
        if kz is not None:
            kz["cpu_baseline"] = cpu_c5b()
    if args.calib_copy:
        # counter calibration: a plain streaming kernel of known traffic
        # (rla::convert_kernel<float, double>: reads 256 MiB, writes 512 MiB)
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty(64 << 20, dtype=torch.float64, device="cuda")
        for _ in range(3):
            rla._lib.call("rla_convert_f32_f64", rla._lib.ctx(src.device), 64 << 20, 1,
                          rla._lib.ptr(src), 64 << 20, rla._lib.ptr(dst), 64 << 20)
        torch.cuda.synchronize()
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "srht_f32_65536x512_traffic.txt")
    if os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith("bytes_per_launch:"):
                traffic = float(line.split(":", 1)[1])
    out = {
        "metric": "sketch GB/s (% HBM peak); solves/s + p99 latency vs plain at equal error",
        "value": gbs,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic N(0,1), seed 1234",
        "config": {"workload": "SRHT/FJLT left apply: A 65536x512 f32 -> 2048x512 "
                               "(Rademacher signs, FWHT, 2048 sampled rows)",
                   "rotating_copies": COPIES, "path": "pruned Float32 kernel (default)",
                   "issue": "independent applies alternating over 2 HIP streams, one graph",
                   "one_stream_ms_per_step": ms_one, "one_stream_GB_s": gbs_one,
                   "one_stream_is": "the same steps on one stream, one graph (kernels back to back)",
                   "parity": parity, "wall_s": wall},
        "roofline": {"bound": "hbm", "achieved": gbs_one, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs_one / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": srht_bytes(),
                     "achieved_is": "algorithmic bytes / per-launch kernel time (one stream)",
                     "throughput_frac": gbs / HBM_PEAK_GBS},
        "sparse_sign_c1": extra,
        "fwht_fjlt": fwht_fjlt,
        "tall_sparse_sign": tall_sparse,
        "rangefinder_c3": rf,
        "ls_batched_c4": ls,
        "kaczmarz_c5": kz,
        "serving_c5": srv,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(S, A0)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
