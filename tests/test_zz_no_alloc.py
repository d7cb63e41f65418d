"""The batched LU solve that allocates nothing
(sl3_band_lu_solve_batched_noalloc) against sl3_band_lu_solve_batched, the
entry point it is a variant of: the same kernels, so the same bits, on two
calls in a row, the second with fewer and smaller systems (the shared
scratch keeps the first call's descriptors); sizes beyond the scratch's
limits are refused before any launch."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

I64P = ctypes.POINTER(ctypes.c_int64)
IP = ctypes.POINTER(ctypes.c_int)


def general_band(n, kl, ku, rng):
    """A diagonally weighted random band in LU storage ((2 kl + ku + 1) x n,
    column-major)."""
    ld = 2 * kl + ku + 1
    ab = np.zeros((n, ld))
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            ab[j, kl + ku + i - j] = rng.uniform(-1, 1) + (4.0 if i == j else 0.0)
    return ab


def factor_and_solve(lib, systems, seed):
    import torch

    import sl3dpp_amd

    rng = np.random.default_rng(seed)
    count = len(systems)
    n = np.array([s[0] for s in systems], dtype=np.int32)
    kl = np.array([s[1] for s in systems], dtype=np.int32)
    ku = np.array([s[2] for s in systems], dtype=np.int32)
    nrhs = np.array([s[3] for s in systems], dtype=np.int32)
    bands = [general_band(*s[:3], rng) for s in systems]
    rhs = [rng.uniform(-1, 1, size=(s[0], s[3])) for s in systems]
    ab_off = np.cumsum([0] + [b.size for b in bands])[:-1].astype(np.int64)
    piv_off = np.cumsum([0] + [s[0] for s in systems])[:-1].astype(np.int64)
    b_off = np.cumsum([0] + [r.size for r in rhs])[:-1].astype(np.int64)
    row_stride = nrhs.astype(np.int64)
    col_stride = np.ones(count, dtype=np.int64)
    dev = "cuda"
    ab = torch.from_numpy(np.concatenate([b.ravel() for b in bands])).to(dev)
    piv = torch.zeros(int(n.sum()), dtype=torch.int32, device=dev)
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    P = ctypes.c_void_p
    torch.cuda.synchronize()
    sl3dpp_amd.check(lib.sl3_band_lu_factor_batched(
        count, n.ctypes.data_as(IP), kl.ctypes.data_as(IP), ku.ctypes.data_as(IP),
        ab_off.ctypes.data_as(I64P), P(ab.data_ptr()), piv_off.ctypes.data_as(I64P),
        P(piv.data_ptr()), P(info.data_ptr())))
    assert np.all(info.cpu().numpy() == 0)
    out = []
    for entry in ("sl3_band_lu_solve_batched", "sl3_band_lu_solve_batched_noalloc"):
        b = torch.from_numpy(np.concatenate([r.ravel() for r in rhs])).to(dev)
        torch.cuda.synchronize()
        sl3dpp_amd.check(getattr(lib, entry)(
            count, n.ctypes.data_as(IP), kl.ctypes.data_as(IP), ku.ctypes.data_as(IP),
            ab_off.ctypes.data_as(I64P), P(ab.data_ptr()), piv_off.ctypes.data_as(I64P),
            P(piv.data_ptr()), nrhs.ctypes.data_as(IP), b_off.ctypes.data_as(I64P),
            row_stride.ctypes.data_as(I64P), col_stride.ctypes.data_as(I64P), P(b.data_ptr())))
        sl3dpp_amd.synchronize()
        out.append(b.cpu().numpy())
    return out


def test_noalloc_solve_matches_the_allocating_one():
    import torch

    import sl3dpp_amd

    assert torch.cuda.is_available(), "no HIP device"
    lib = sl3dpp_amd.lib
    print("[noalloc] first call: 5 systems", flush=True)
    first = factor_and_solve(lib, [(1282, 17, 17, 130), (483, 17, 17, 258), (273, 17, 17, 2),
                                   (90, 5, 5, 20), (700, 4, 9, 70)], seed=5)
    assert np.array_equal(first[0], first[1])
    print("[noalloc] second call: 2 smaller systems", flush=True)
    second = factor_and_solve(lib, [(60, 3, 3, 65), (41, 17, 17, 1)], seed=6)
    assert np.array_equal(second[0], second[1])
    assert np.max(np.abs(second[1])) > 0.0


def test_noalloc_refuses_sizes_beyond_its_scratch():
    import sl3dpp_amd

    lib = sl3dpp_amd.lib
    count = 513  # one more system than the scratch holds: refused on the host
    n = np.full(count, 4, dtype=np.int32)
    z = np.zeros(count, dtype=np.int32)
    o = np.zeros(count, dtype=np.int64)
    status = lib.sl3_band_lu_solve_batched_noalloc(
        count, n.ctypes.data_as(IP), z.ctypes.data_as(IP), z.ctypes.data_as(IP),
        o.ctypes.data_as(I64P), ctypes.c_void_p(8), o.ctypes.data_as(I64P), ctypes.c_void_p(8),
        n.ctypes.data_as(IP), o.ctypes.data_as(I64P), None, None, ctypes.c_void_p(8))
    assert status == -1
    assert b"512" in lib.sl3_last_error()
    # too many blocks of 64 right-hand sides
    one = np.ones(1, dtype=np.int32)
    nr = np.array([64 * 4097], dtype=np.int32)
    o1 = np.zeros(1, dtype=np.int64)
    status = lib.sl3_band_lu_solve_batched_noalloc(
        1, one.ctypes.data_as(IP), np.zeros(1, dtype=np.int32).ctypes.data_as(IP),
        np.zeros(1, dtype=np.int32).ctypes.data_as(IP), o1.ctypes.data_as(I64P),
        ctypes.c_void_p(8), o1.ctypes.data_as(I64P), ctypes.c_void_p(8), nr.ctypes.data_as(IP),
        o1.ctypes.data_as(I64P), None, None, ctypes.c_void_p(8))
    assert status == -1
    assert b"4096" in lib.sl3_last_error()
