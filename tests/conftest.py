import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def assert_parity(got, ref, tol=1e-5, peak=2e-6):
    """Relative L2 within ``tol`` and no element off by more than ``peak``
    times the largest reference magnitude.  FP32 accumulation noise stays
    ~10x below ``peak``; a single flipped rounding decision (a coordinate or
    kernel-table index) moves an element by ~1e-4 of the peak and fails."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    err = rel_l2(got, ref)
    scale = float(np.max(np.abs(ref))) if ref.size else 0.0
    worst = float(np.max(np.abs(got - ref))) if ref.size else 0.0
    assert err < tol, f"relative L2 {err:.3e} >= {tol:.1e}"
    assert worst <= peak * scale, f"max abs error {worst:.3e} > {peak:.1e} * {scale:.3e}"


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o._lib()  # raises if the oracle is not built
    return o


@pytest.fixture(scope="session")
def ker32(oracle):
    return oracle.kb_table().astype(np.float32)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def cuda():
    """Skip-free GPU fixture: a gpu-marked test on a machine without a GPU is
    an error, not a pass."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a machine without a GPU (run with -m 'not gpu')")
    import multigpugridder_amd
    multigpugridder_amd.lib()  # raises LibraryNotBuilt if the HIP library is missing
    return torch.device("cuda", 0)
