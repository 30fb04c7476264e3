import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def rla():
    import __graft_entry__
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


@pytest.fixture(scope="session")
def coracle():
    """The C restatement (oracle/liboracle.so) through ctypes."""
    import ctypes
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(path):
        pytest.skip("oracle/liboracle.so not built")
    lib = ctypes.CDLL(path)
    vp, i64, f64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
    lib.oracle_fwht_f32.argtypes = [vp, i64, i64, i64, vp, f64, i32]
    lib.oracle_fwht_f64.argtypes = [vp, i64, i64, i64, vp, f64, i32]
    lib.oracle_srht_left_f32.argtypes = [vp, i64, i64, i64, i64, i64, vp, vp, f64, vp, i64, f64,
                                         f64, i32]
    lib.oracle_sparse_left_f64.argtypes = [vp, vp, vp, i64, i64, vp, i64, i64, vp, i64, f64, f64,
                                           i32]
    return lib
