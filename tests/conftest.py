"""Test configuration.

Markers:
  gpu — needs an AMD GPU and the built libstormbird_amd.so (parity tests proper).
Everything else runs on the CPU: the oracle against closed forms and the
reference's own test cases, host logic, and the C-ABI library's exports.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: requires an AMD Instinct GPU and libstormbird_amd.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle_py

    oracle_py.lib("glibc")
    oracle_py.lib("pm")
    return oracle_py


@pytest.fixture(scope="session", autouse=True)
def _torch_gpu_runtime_first(request):
    """When GPU tests run, torch's HIP runtime is initialised before the library's first
    HIP call in the process (the order bench.py uses); torch initialised after the
    library has created its streams and buffers can report no GPU."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        try:
            import torch

            if torch.cuda.is_available():
                torch.zeros(1, device="cuda")
        except Exception:
            pass
    yield
