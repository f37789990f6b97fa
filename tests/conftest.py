import math
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "accelerate-fft_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X); parity tests through the C ABI")


def tolerance(is_f64, n):
    """Relative-L2 bound against the FP64 oracle: 1e-13 for Complex Double, 1e-6*log2(N) for
    Complex Float, N the full transform size (h*w for 2D, d*h*w for 3D)."""
    if is_f64:
        return 1e-13
    return 1e-6 * max(1.0, math.log2(max(n, 2)))


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    import afft
    afft._lib.lib()  # fail loudly if libafft.so is missing
    return torch.device("cuda", 0)
