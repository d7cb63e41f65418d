"""The oracle's banded solvers against LAPACK golden vectors and against the
reference's own src/BandSolvers.cpp compiled into oracle/_ref — the checks of
tests/TestBandSolvers.cpp:87-176 (factors, pivots and solutions within 1e-12
relative; LAPACK's failure indices)."""

import ctypes

import numpy as np
import pytest
from conftest import dptr, iptr

TOL = 1e-12


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def chol_names(golden):
    return [n for n in golden["names"] if n.startswith("chol")]


def lu_names(golden):
    return [n for n in golden["names"] if n.startswith("lu")]


@pytest.mark.parametrize("k", range(4))
def test_cholesky_matches_lapack(oracle, golden, k):
    name = f"chol{k}"
    n, kd, nrhs = (int(v) for v in golden[f"{name}_shape"])
    ab = golden[f"{name}_ab"].copy()
    info = oracle.sl3o_band_cholesky_factor(dptr(ab), n, kd)
    assert info == 0
    assert rel(ab, golden[f"{name}_factor"]) < TOL
    b = golden[f"{name}_b"].copy()
    oracle.sl3o_band_cholesky_solve(dptr(ab), n, kd, dptr(b), nrhs)
    assert rel(b, golden[f"{name}_x"]) < TOL


@pytest.mark.parametrize("k", range(4))
def test_lu_matches_lapack(oracle, golden, k):
    name = f"lu{k}"
    n, kl, ku, nrhs = (int(v) for v in golden[f"{name}_shape"])
    ab = golden[f"{name}_ab"].copy()
    piv = np.zeros(n, dtype=np.int32)
    info = oracle.sl3o_band_lu_factor(dptr(ab), n, kl, ku, iptr(piv))
    assert info == 0
    assert np.array_equal(piv, golden[f"{name}_piv"])
    if n > 10:
        assert np.count_nonzero(piv != np.arange(n)) > 0  # pivoting exercised
    assert rel(ab, golden[f"{name}_factor"]) < TOL
    b = golden[f"{name}_b"].copy()
    oracle.sl3o_band_lu_solve(dptr(ab), n, kl, ku, iptr(piv), dptr(b), nrhs)
    assert rel(b, golden[f"{name}_x"]) < TOL


def test_cholesky_reports_the_indefinite_minor(oracle, golden):
    n, kd, _ = (int(v) for v in golden["cholfail_shape"])
    ab = golden["cholfail_ab"].copy()
    assert oracle.sl3o_band_cholesky_factor(dptr(ab), n, kd) == int(golden["cholfail_info"])


def test_lu_reports_the_zero_pivot(oracle, golden):
    n, kl, ku, _ = (int(v) for v in golden["lufail_shape"])
    ab = golden["lufail_ab"].copy()
    piv = np.zeros(n, dtype=np.int32)
    assert oracle.sl3o_band_lu_factor(dptr(ab), n, kl, ku, iptr(piv)) == int(golden["lufail_info"])


@pytest.mark.parametrize("name", ["chol0", "chol1", "chol2", "chol3", "cholfail"])
def test_cholesky_bitwise_equals_reference_build(oracle, refbuild, golden, name):
    shape = [int(v) for v in golden[f"{name}_shape"]]
    n, kd = shape[0], shape[1]
    a1 = golden[f"{name}_ab"].copy()
    a2 = a1.copy()
    i1 = oracle.sl3o_band_cholesky_factor(dptr(a1), n, kd)
    i2 = refbuild.sl3ref_band_cholesky_factor(dptr(a2), n, kd)
    assert i1 == i2
    assert np.array_equal(a1, a2)
    if i1 == 0:
        b1 = golden[f"{name}_b"].copy()
        b2 = b1.copy()
        oracle.sl3o_band_cholesky_solve(dptr(a1), n, kd, dptr(b1), shape[2])
        refbuild.sl3ref_band_cholesky_solve(dptr(a2), n, kd, dptr(b2), shape[2])
        assert np.array_equal(b1, b2)


@pytest.mark.parametrize("name", ["lu0", "lu1", "lu2", "lu3", "lufail"])
def test_lu_bitwise_equals_reference_build(oracle, refbuild, golden, name):
    n, kl, ku, nrhs = (int(v) for v in golden[f"{name}_shape"])
    a1 = golden[f"{name}_ab"].copy()
    a2 = a1.copy()
    p1 = np.zeros(n, dtype=np.int32)
    p2 = np.zeros(n, dtype=np.int32)
    i1 = oracle.sl3o_band_lu_factor(dptr(a1), n, kl, ku, iptr(p1))
    i2 = refbuild.sl3ref_band_lu_factor(dptr(a2), n, kl, ku, iptr(p2))
    assert i1 == i2
    assert np.array_equal(p1, p2)
    assert np.array_equal(a1, a2)
    if i1 == 0:
        b1 = golden[f"{name}_b"].copy()
        b2 = b1.copy()
        oracle.sl3o_band_lu_solve(dptr(a1), n, kl, ku, iptr(p1), dptr(b1), nrhs)
        refbuild.sl3ref_band_lu_solve(dptr(a2), n, kl, ku, iptr(p2), dptr(b2), nrhs)
        assert np.array_equal(b1, b2)
