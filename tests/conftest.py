"""Test configuration: the `gpu` marker and shared fixtures.

GPU tests call the HIP path through the C ABI (libdfx.so); the oracle
(oracle/, a CPU restatement of the reference) is the checker.
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _ensure_built():
    lib = os.path.join(ROOT, "dolfinx_amd", "libdfx.so")
    if not os.path.exists(lib):
        from dolfinx_amd import build

        build.build()
    olib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(olib):
        import oracle

        oracle.build()


_ensure_built()


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(20261202)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
