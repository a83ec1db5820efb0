"""Shared pytest set-up: the ``gpu`` marker, import paths and the seeded inputs."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "fluxoptics.jl_amd")
for p in (ROOT, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def lib():
    """The HIP library handle, for GPU tests only (fails loudly when it is not built)."""
    from fluxoptics import _lib
    return _lib.load()
