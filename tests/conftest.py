import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not has_gpu():
        pytest.skip("no GPU")
    import torch
    return torch.device("cuda:0")


def gen_inputs(rng, m, n, phi):
    """(U(0,1) - 0.5) * exp(phi * N(0,1)), the input distribution of the reference harness
    (src/programming_examples/basic/matrix_multiplication/test.cpp:219-239)."""
    import numpy as np
    return (rng.random((m, n)) - 0.5) * np.exp(phi * rng.standard_normal((m, n)))
