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
