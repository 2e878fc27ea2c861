"""Shared pytest set-up: import paths, the `gpu` marker and the GPU context."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("pbr-rust_amd", "oracle", "tests"):
    path = os.path.join(ROOT, sub)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950)")


@pytest.fixture(scope="session")
def ctx():
    import pbrtm

    if pbrtm.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    c = pbrtm.Context(0)
    yield c
    c.close()
