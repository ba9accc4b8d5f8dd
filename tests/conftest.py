"""Shared test setup: markers, paths to the product package and the oracle."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tee-bench_amd")
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (PKG, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libtbj.so")
    config.addinivalue_line("markers", "slow: full-size workloads")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.Oracle()


@pytest.fixture(scope="session")
def tbj():
    if not gpu_available():
        pytest.fail("GPU test run without a visible HIP device")
    import tbjoin
    tbjoin.init(0)
    return tbjoin
