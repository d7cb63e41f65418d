"""Shared test plumbing: the gpu marker, and loaders for the oracle (the
CPU restatement the HIP path is checked against) and for oracle/_ref (the
reference's own self-contained sources compiled in the build container)."""

import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _load(path):
    if not os.path.exists(path):
        return None
    return ctypes.CDLL(path)


@pytest.fixture(scope="session")
def oracle():
    lib = _load(os.path.join(ROOT, "oracle", "liboracle.so"))
    if lib is None:
        pytest.fail("oracle/liboracle.so is not built (run __graft_entry__.build())")
    return lib


@pytest.fixture(scope="session")
def refbuild():
    lib = _load(os.path.join(ROOT, "oracle", "_ref", "libsl3ref.so"))
    if lib is None:
        pytest.skip("oracle/_ref not built here (no reference checkout)")
    return lib


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "band_vectors.npz"))


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
