"""The smoke check: one small invocation of the hot path on cuda:0, checked
against the oracle (the CPU restatement). Used by __graft_entry__.smoke()."""

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _oracle():
    return ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))


def run_smoke():
    import torch

    import sl3dpp_amd
    from sl3dpp_amd import band

    assert torch.cuda.is_available(), "smoke needs the HIP device"
    dev = torch.device("cuda:0")
    oracle = _oracle()
    D = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    rng = np.random.default_rng(3)
    n, kd, nrhs = 300, 17, 8
    ab = rng.uniform(-1, 1, size=(n, kd + 1))
    ab[:, kd] += 2 * kd + 1
    for j in range(kd):
        ab[j, : kd - j] = 0.0
    b = rng.uniform(-1, 1, size=(nrhs, n))
    abd = torch.from_numpy(ab.copy()).to(dev)
    bd = torch.from_numpy(b.copy()).to(dev)
    assert band.band_cholesky_factor(abd, n, kd) == 0
    band.band_cholesky_solve(abd, n, kd, bd, nrhs)
    ref = ab.copy()
    x = b.copy()
    assert oracle.sl3o_band_cholesky_factor(D(ref), n, kd) == 0
    oracle.sl3o_band_cholesky_solve(D(ref), n, kd, D(x), nrhs)
    err = np.max(np.abs(bd.cpu().numpy() - x)) / np.max(np.abs(x))
    assert err < 1e-12, err
    print(f"smoke: band factor+solve n={n} kd={kd} nrhs={nrhs} rel err {err:.2e}",
          flush=True)
    step_smoke = getattr(sl3dpp_amd, "step_smoke", None)
    if step_smoke is not None:
        step_smoke()
