"""The HIP band solvers against LAPACK (committed golden vectors of the
reference's test shapes, tests/TestBandSolvers.cpp:133-176) and against the
oracle at the 3D step's own shapes: factors and solutions within 1e-12
relative, pivots identical, LAPACK's info codes."""

import ctypes

import numpy as np
import pytest
from conftest import dptr, iptr

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def dev():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    import sl3dpp_amd  # noqa: F401

    return torch.device("cuda:0")


def to_dev(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("k", range(4))
def test_cholesky_matches_lapack(dev, golden, k):
    from sl3dpp_amd import band

    name = f"chol{k}"
    print(f"[band] cholesky {name}", flush=True)
    n, kd, nrhs = (int(v) for v in golden[f"{name}_shape"])
    ab = to_dev(golden[f"{name}_ab"], dev)
    assert band.band_cholesky_factor(ab, n, kd) == 0
    assert rel(ab.cpu().numpy(), golden[f"{name}_factor"]) < TOL
    b = to_dev(golden[f"{name}_b"], dev)
    band.band_cholesky_solve(ab, n, kd, b, nrhs)
    assert rel(b.cpu().numpy(), golden[f"{name}_x"]) < TOL


@pytest.mark.parametrize("k", range(4))
def test_lu_matches_lapack(dev, golden, k):
    import torch
    from sl3dpp_amd import band

    name = f"lu{k}"
    print(f"[band] lu {name}", flush=True)
    n, kl, ku, nrhs = (int(v) for v in golden[f"{name}_shape"])
    ab = to_dev(golden[f"{name}_ab"], dev)
    piv = torch.zeros(n, dtype=torch.int32, device=dev)
    assert band.band_lu_factor(ab, n, kl, ku, piv) == 0
    assert np.array_equal(piv.cpu().numpy(), golden[f"{name}_piv"])
    assert rel(ab.cpu().numpy(), golden[f"{name}_factor"]) < TOL
    b = to_dev(golden[f"{name}_b"], dev)
    band.band_lu_solve(ab, n, kl, ku, piv, b, nrhs)
    assert rel(b.cpu().numpy(), golden[f"{name}_x"]) < TOL


def test_cholesky_reports_the_indefinite_minor(dev, golden):
    from sl3dpp_amd import band

    n, kd, _ = (int(v) for v in golden["cholfail_shape"])
    ab = to_dev(golden["cholfail_ab"], dev)
    assert band.band_cholesky_factor(ab, n, kd) == int(golden["cholfail_info"])


def test_lu_reports_the_zero_pivot(dev, golden):
    import torch
    from sl3dpp_amd import band

    n, kl, ku, _ = (int(v) for v in golden["lufail_shape"])
    ab = to_dev(golden["lufail_ab"], dev)
    piv = torch.zeros(n, dtype=torch.int32, device=dev)
    assert band.band_lu_factor(ab, n, kl, ku, piv) == int(golden["lufail_info"])


def _spd_band(n, kd, rng):
    ab = rng.uniform(-1, 1, size=(n, kd + 1))
    ab[:, kd] += 2 * kd + 1
    for j in range(min(kd, n)):
        ab[j, : kd - j] = 0.0
    return ab


def _general_band(n, kl, ku, rng):
    ld = 2 * kl + ku + 1
    ab = np.zeros((n, ld))
    vals = rng.uniform(-1, 1, size=(n, ld))
    kv = kl + ku
    for j in range(n):
        lo, hi = max(0, j - ku), min(n - 1, j + kl)
        ab[j, kv + lo - j: kv + hi - j + 1] = vals[j, kv + lo - j: kv + hi - j + 1]
        ab[j, kv] *= 0.05
    return ab


def test_batched_degree_shapes_match_oracle(dev, oracle):
    """All degrees of a pass at lmax 128 on the 128-element, 6-node mesh:
    spheroidal kd = 17 (n from 1282 down to 273, 2(l+1) real columns) and
    toroidal kd = 5, in one batched launch each, right-hand sides row-major
    (columns fastest) as the step lays them out."""
    import torch

    import sl3dpp_amd

    lib = sl3dpp_amd.lib
    rng = np.random.default_rng(7)
    degrees = [0, 1, 2, 7, 32, 64, 100, 128]
    sizes = {0: 1282, 1: 1221, 2: 1221, 7: 1100, 32: 828, 64: 483, 100: 330, 128: 273}
    systems = []
    for l in degrees:
        systems.append((sizes[l], 17, 2 * (l + 1)))
        systems.append((max(60, sizes[l] // 3), 5, 2 * (l + 1)))
    count = len(systems)
    n = np.array([s[0] for s in systems], dtype=np.int32)
    kd = np.array([s[1] for s in systems], dtype=np.int32)
    nrhs = np.array([s[2] for s in systems], dtype=np.int32)
    bands = [_spd_band(s[0], s[1], rng) for s in systems]
    rhs = [rng.uniform(-1, 1, size=(s[0], s[2])) for s in systems]  # row-major
    ab_off = np.cumsum([0] + [b.size for b in bands])[:-1].astype(np.int64)
    b_off = np.cumsum([0] + [r.size for r in rhs])[:-1].astype(np.int64)
    row_stride = nrhs.astype(np.int64)
    col_stride = np.ones(count, dtype=np.int64)
    ab_all = to_dev(np.concatenate([b.ravel() for b in bands]), dev)
    b_all = to_dev(np.concatenate([r.ravel() for r in rhs]), dev)
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    P = ctypes.c_void_p
    print("[band] batched cholesky factor", flush=True)
    torch.cuda.synchronize()
    sl3dpp_amd.check(lib.sl3_band_cholesky_factor_batched(
        count, iptr(n), iptr(kd), ab_off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
        P(ab_all.data_ptr()), P(info.data_ptr())))
    assert np.all(info.cpu().numpy() == 0)
    print("[band] batched cholesky solve", flush=True)
    I64 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))  # noqa: E731
    sl3dpp_amd.check(lib.sl3_band_cholesky_solve_batched(
        count, iptr(n), iptr(kd), I64(ab_off), P(ab_all.data_ptr()), iptr(nrhs),
        I64(b_off), I64(row_stride), I64(col_stride), P(b_all.data_ptr())))
    sl3dpp_amd.synchronize()
    got_ab = ab_all.cpu().numpy()
    got_b = b_all.cpu().numpy()
    for s, (ns, kds, nr) in enumerate(systems):
        ref = bands[s].copy()
        assert oracle.sl3o_band_cholesky_factor(dptr(ref), ns, kds) == 0
        mine = got_ab[ab_off[s]: ab_off[s] + ref.size].reshape(ref.shape)
        assert rel(mine, ref) < TOL
        x = np.ascontiguousarray(rhs[s].T)  # column-major for the oracle
        oracle.sl3o_band_cholesky_solve(dptr(ref), ns, kds, dptr(x), nr)
        mineb = got_b[b_off[s]: b_off[s] + rhs[s].size].reshape(ns, nr).T
        assert rel(mineb, x) < TOL


def test_batched_lu_degree_shapes_match_oracle(dev, oracle):
    """The LU fallback at the spheroidal shape kl = ku = 17, batched."""
    import torch

    import sl3dpp_amd

    lib = sl3dpp_amd.lib
    rng = np.random.default_rng(11)
    systems = [(1282, 17, 17, 2), (483, 17, 17, 130), (273, 17, 17, 258), (90, 5, 5, 20)]
    count = len(systems)
    n = np.array([s[0] for s in systems], dtype=np.int32)
    kl = np.array([s[1] for s in systems], dtype=np.int32)
    ku = np.array([s[2] for s in systems], dtype=np.int32)
    nrhs = np.array([s[3] for s in systems], dtype=np.int32)
    bands = [_general_band(s[0], s[1], s[2], rng) for s in systems]
    rhs = [rng.uniform(-1, 1, size=(s[0], s[3])) for s in systems]
    I64 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))  # noqa: E731
    ab_off = np.cumsum([0] + [b.size for b in bands])[:-1].astype(np.int64)
    piv_off = np.cumsum([0] + [s[0] for s in systems])[:-1].astype(np.int64)
    b_off = np.cumsum([0] + [r.size for r in rhs])[:-1].astype(np.int64)
    row_stride = nrhs.astype(np.int64)
    col_stride = np.ones(count, dtype=np.int64)
    ab_all = to_dev(np.concatenate([b.ravel() for b in bands]), dev)
    b_all = to_dev(np.concatenate([r.ravel() for r in rhs]), dev)
    piv = torch.zeros(int(n.sum()), dtype=torch.int32, device=dev)
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    P = ctypes.c_void_p
    print("[band] batched lu factor", flush=True)
    torch.cuda.synchronize()
    sl3dpp_amd.check(lib.sl3_band_lu_factor_batched(
        count, iptr(n), iptr(kl), iptr(ku), I64(ab_off), P(ab_all.data_ptr()),
        I64(piv_off), P(piv.data_ptr()), P(info.data_ptr())))
    assert np.all(info.cpu().numpy() == 0)
    print("[band] batched lu solve", flush=True)
    sl3dpp_amd.check(lib.sl3_band_lu_solve_batched(
        count, iptr(n), iptr(kl), iptr(ku), I64(ab_off), P(ab_all.data_ptr()),
        I64(piv_off), P(piv.data_ptr()), iptr(nrhs), I64(b_off), I64(row_stride),
        I64(col_stride), P(b_all.data_ptr())))
    sl3dpp_amd.synchronize()
    got_ab = ab_all.cpu().numpy()
    got_piv = piv.cpu().numpy()
    got_b = b_all.cpu().numpy()
    for s, (ns, kls, kus, nr) in enumerate(systems):
        ref = bands[s].copy()
        p = np.zeros(ns, dtype=np.int32)
        assert oracle.sl3o_band_lu_factor(dptr(ref), ns, kls, kus, iptr(p)) == 0
        assert np.array_equal(got_piv[piv_off[s]: piv_off[s] + ns], p)
        mine = got_ab[ab_off[s]: ab_off[s] + ref.size].reshape(ref.shape)
        assert rel(mine, ref) < TOL
        x = np.ascontiguousarray(rhs[s].T)
        oracle.sl3o_band_lu_solve(dptr(ref), ns, kls, kus, iptr(p), dptr(x), nr)
        mineb = got_b[b_off[s]: b_off[s] + rhs[s].size].reshape(ns, nr).T
        assert rel(mineb, x) < TOL
