"""Writes tests/golden/band_vectors.npz: LAPACK (scipy's dpbtrf/dpbtrs,
dgbtrf/dgbtrs) factors, pivots, solutions and info codes for the banded
Cholesky and LU cases the reference's own band-solver tests run
(tests/TestBandSolvers.cpp:133-176), with inputs built the same way: uniform
[-1, 1] values, the symmetric diagonal shifted by 2 kd + 1, the general
diagonal scaled by 0.05 so that pivoting happens, one indefinite minor at
column 17, one vanishing column at 11. numpy's generator replaces the
reference's mt19937 stream (the values differ; the construction and the
LAPACK routines are the same).

Run once in the build container: python tests/golden/make_band_vectors.py
"""

import os

import numpy as np
from scipy.linalg import lapack

HERE = os.path.dirname(os.path.abspath(__file__))


def symmetric_band(n, kd, shift, rng):
    ld = kd + 1
    ab = rng.uniform(-1.0, 1.0, size=(n, ld))  # column j = ab[j]
    for j in range(n):
        ab[j, kd] += shift
        for i in range(j - kd, 0):
            ab[j, kd + i - j] = 0.0
    return ab


def general_band(n, kl, ku, diagonal, rng):
    ld = 2 * kl + ku + 1
    kv = kl + ku
    values = rng.uniform(-1.0, 1.0, size=(n, ld))
    ab = np.zeros((n, ld))
    for j in range(n):
        for i in range(max(0, j - ku), min(n - 1, j + kl) + 1):
            ab[j, kv + i - j] = values[j, kv + i - j] * (diagonal if i == j else 1.0)
    return ab


def chol_case(n, kd, nrhs, seed, out, name, poison=None):
    rng = np.random.default_rng(seed)
    ab = symmetric_band(n, kd, 2.0 * kd + 1.0, rng)
    if poison is not None:
        ab[poison, kd] = -1.0
    # LAPACK takes (ldab, n) Fortran order: ab.T is that array.
    c, info = lapack.dpbtrf(ab.T.copy(order="F"), lower=0)
    b = rng.uniform(-1.0, 1.0, size=(nrhs, n))
    out[f"{name}_ab"] = ab
    out[f"{name}_b"] = b
    out[f"{name}_info"] = np.array(info)
    out[f"{name}_shape"] = np.array([n, kd, nrhs])
    if info == 0:
        x, info2 = lapack.dpbtrs(c, b.T.copy(order="F"), lower=0)
        assert info2 == 0
        out[f"{name}_factor"] = np.asarray(c).T.copy()
        out[f"{name}_x"] = np.asarray(x).T.copy()


def lu_case(n, kl, ku, nrhs, seed, out, name, diagonal=0.05, zero=None):
    rng = np.random.default_rng(seed)
    ab = general_band(n, kl, ku, diagonal, rng)
    if zero is not None:
        for i in range(zero - ku, zero + kl + 1):
            ab[zero, kl + ku + i - zero] = 0.0
    lu, piv, info = lapack.dgbtrf(ab.T.copy(order="F"), kl, ku)
    b = rng.uniform(-1.0, 1.0, size=(nrhs, n))
    out[f"{name}_ab"] = ab
    out[f"{name}_b"] = b
    out[f"{name}_info"] = np.array(info)
    out[f"{name}_shape"] = np.array([n, kl, ku, nrhs])
    if info == 0:
        x, info2 = lapack.dgbtrs(lu, kl, ku, b.T.copy(order="F"), piv)
        assert info2 == 0
        out[f"{name}_factor"] = np.asarray(lu).T.copy()
        out[f"{name}_piv"] = np.asarray(piv, dtype=np.int32)  # 0-based
        out[f"{name}_x"] = np.asarray(x).T.copy()


def main():
    out = {}
    chol_cases = [(200, 14, 4, 1), (60, 5, 1, 2), (4, 7, 2, 3), (1, 0, 1, 4)]
    lu_cases = [(200, 14, 14, 4, 5), (80, 3, 6, 2, 6), (80, 6, 3, 1, 7),
                (6, 9, 9, 2, 8)]
    names = []
    for k, (n, kd, nrhs, seed) in enumerate(chol_cases):
        chol_case(n, kd, nrhs, seed, out, f"chol{k}")
        names.append(f"chol{k}")
    for k, (n, kl, ku, nrhs, seed) in enumerate(lu_cases):
        lu_case(n, kl, ku, nrhs, seed, out, f"lu{k}")
        names.append(f"lu{k}")
    chol_case(40, 4, 1, 9, out, "cholfail", poison=17)
    lu_case(30, 3, 3, 1, 10, out, "lufail", diagonal=1.0, zero=11)
    out["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "band_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "band_vectors.npz"))


if __name__ == "__main__":
    main()
