"""Shared test setup.

Markers: ``gpu`` -- needs a HIP device (MI355X).  Everything else runs on the CPU and uses the
FP64 oracle (oracle/) and the product's host-side scene build.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One HIP context for the whole GPU session.  Fails (does not skip) without a device or
    without the built library: the GPU tests exist to exercise the native path."""
    import tinypt_amd as T
    return T.default_context()
