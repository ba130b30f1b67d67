"""Shared test set-up: the `gpu` marker, import paths and golden fixtures."""
import gzip
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle"), GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_golden(name: str):
    with gzip.open(os.path.join(GOLDEN, name), "rt") as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def gpu_lib():
    """The product library on a GPU; fails (does not skip) when it is missing."""
    from cusna_amd import lib

    L = lib.load()
    assert L.cusna_has_cuda() == 1, "no HIP device visible to libcusna_amd.so"
    return lib
