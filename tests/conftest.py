"""pytest configuration: the ``gpu`` marker and shared helpers.

CPU tests (``-m "not gpu"``) check the oracle against invariants and closed
forms, the host logic, and that the C-ABI library loads and exports every
entry point.  GPU tests (``-m gpu``) compare the HIP path with the oracle
through the C ABI; they fail (not skip) when the library or the GPU is
missing.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.load_library()
    return oracle
