"""The allocation-free Gaussian apply (rla_gaussian_apply_noalloc_f32): the
split partial sums live in the library's file-scope device array instead of a
context workspace.  Through the C ABI, against the allocating entry point
(rla_gaussian_apply_f32) on two calls in a row, the second smaller: the same
kernels and split order, so the results are equal bit for bit; the calls leave
the context workspace and the device's HBM account untouched."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
F32 = torch.float32


def colmajor(a, dev):
    return torch.as_tensor(np.asarray(a, dtype=np.float32)).T.contiguous().T.to(dev)


def apply(rla, dev, name, side, S, A, C, alpha, beta):
    lib = rla._lib
    lib.call(name, lib.ctx(dev), side, S.shape[0], S.shape[1], lib.ptr(S), S.stride(0),
             S.stride(1), A.shape[0], A.shape[1], lib.ptr(A), A.stride(0), A.stride(1),
             alpha, beta, lib.ptr(C), C.stride(0), C.stride(1))


def test_noalloc_gaussian_apply_matches_allocating_entry(rla, dev):
    lib = rla._lib.load()
    rng = np.random.default_rng(7)
    cases = [  # (side, S shape, A shape): the first splits K over 16 workgroups
        (0, (266, 8192), (8192, 300)),
        (1, (64, 40), (3000, 64)),
    ]
    ws0, used0 = ctypes.c_size_t(), ctypes.c_size_t()
    ws1, used1 = ctypes.c_size_t(), ctypes.c_size_t()
    for side, ss, sa in cases:
        S = colmajor(rng.standard_normal(ss) / np.sqrt(ss[0]), dev)
        A = colmajor(rng.standard_normal(sa), dev)
        rows = ss[0] if side == 0 else sa[0]
        cols = sa[1] if side == 0 else ss[1]
        C0 = rng.standard_normal((rows, cols))
        for alpha, beta in ((1.0, 0.0), (0.5, -2.0)):
            Cw = colmajor(C0, dev)
            apply(rla, dev, "rla_gaussian_apply_f32", side, S, A, Cw, alpha, beta)
            torch.cuda.synchronize()
            rla._lib.check(lib.rla_ctx_workspace_bytes(rla._lib.ctx(dev), ctypes.byref(ws0)))
            rla._lib.check(lib.rla_hbm_query(0, ctypes.byref(used0), None))
            Cn = colmajor(C0, dev)
            apply(rla, dev, "rla_gaussian_apply_noalloc_f32", side, S, A, Cn, alpha, beta)
            torch.cuda.synchronize()
            rla._lib.check(lib.rla_ctx_workspace_bytes(rla._lib.ctx(dev), ctypes.byref(ws1)))
            rla._lib.check(lib.rla_hbm_query(0, ctypes.byref(used1), None))
            assert torch.equal(Cn, Cw)
            assert ws1.value == ws0.value and used1.value == used0.value
            # and both are the product (FP64 check, Float32 tolerance)
            Sd, Ad = S.double().cpu().numpy(), A.double().cpu().numpy()
            want = alpha * (Sd @ Ad if side == 0 else Ad @ Sd) + beta * C0.astype(np.float32)
            err = np.linalg.norm(Cn.double().cpu().numpy() - want) / np.linalg.norm(want)
            assert err < 1e-5, err
