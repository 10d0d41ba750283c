"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on the CPU."""
from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def vctlib():
    from voxel_cone_tracing_amd import _lib

    return _lib.load()


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def gpu_available(vctlib):
    from voxel_cone_tracing_amd import _lib

    n = _lib.device_count()
    if n == 0:
        pytest.fail("GPU test selected but no HIP device is visible")
    return n
