"""Shared test setup: markers, import paths, fixtures, lazy builds."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP kernels")
    config.addinivalue_line("markers", "slow: full-size configuration")


def _ensure_built():
    import oracle
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        oracle.build()
    lib = os.path.join(ROOT, "gpugraphalgo_amd", "lib", "libgga.so")
    if not os.path.exists(lib):
        from gpugraphalgo_amd.build import build
        build()


_ensure_built()


@pytest.fixture(scope="session")
def golden():
    """(index, arrays, meta) of tests/golden/golden_sp.npz."""
    with open(os.path.join(GOLDEN, "golden_sp.json")) as f:
        meta = json.load(f)
    arrays = dict(np.load(os.path.join(GOLDEN, "golden_sp.npz"), allow_pickle=False))
    return meta["index"], arrays, meta


@pytest.fixture(scope="session")
def split_kat():
    with open(os.path.join(GOLDEN, "split_kat.json")) as f:
        return json.load(f)


def golden_graph(arrays, name):
    return arrays[f"{name}.row_ptr"], arrays[f"{name}.dst"], arrays[f"{name}.w"]


@pytest.fixture(scope="session")
def gpu():
    """Skips nothing: a GPU test on a machine without a device must fail loudly."""
    import gpugraphalgo_amd as gga
    n = gga.device_count()
    assert n >= 1, "gpu-marked test but libgga.so sees no HIP device"
    return n
