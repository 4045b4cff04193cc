"""Test configuration.

Markers
  gpu   needs a HIP device (runs on the MI355X box: pytest -m gpu).
Tests without the marker run on any host (pytest -m "not gpu").
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


def load_golden(tag):
    from p3arsec_amd.books import Book
    z = np.load(os.path.join(GOLDEN, f"golden_{tag}.npz"))
    book = Book(z["s"], z["k"], z["r"], z["v"], z["t"], z["otype"])
    return book, z["ref_price"]


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    """The loaded C ABI on a machine with a GPU; fails (never skips) when the
    library or the device is missing, so a GPU run cannot pass silently."""
    from p3arsec_amd import capi
    capi.load()
    n = capi.device_count()
    assert n > 0, "no HIP device visible"
    capi.call("bs_set_device", 0)
    return capi
