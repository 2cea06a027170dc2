import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP engine)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref (the reference's CPU sources compiled in place)")


def _has_gpu():
    try:
        from gpuvoxels_amd import engine
        return engine.device_count() > 0
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    markexpr = config.getoption("-m") or ""
    if "gpu" in markexpr and "not gpu" not in markexpr:
        return   # explicitly asked for GPU tests: let them fail loudly if the GPU path is broken
    if not _has_gpu():
        skip = pytest.mark.skip(reason="no GPU in this environment")
        for it in items:
            if "gpu" in it.keywords:
                it.add_marker(skip)
