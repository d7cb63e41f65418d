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
    step_check(oracle)


def step_check(oracle):
    """One lateral-viscosity step at lmax 8 on a small Earth (two passes of
    the fixed point: StartStep, transforms, product, force, memory response,
    strain, Anderson start, Finish) against the oracle's step."""
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import sl3dpp_amd.step as st
    from step_problem import OracleStep, amplitudes, small_earth_spec

    lmax = 8
    earth = st.Earth(small_earth_spec())
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    kw = dict(max_iterations=2, rtol=1e-14, anderson_memory=6, anderson_rate=0.5)
    op = st.Viscoelastic3D(earth, lmax, None, st.StepOptions.default(**kw))
    op.set_delta(op.delta_field(st.delta_coefficients(6, op.nodes, seed=5, rms=0.7, lmax=lmax)))
    ref = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, dt, op.delta_grid(),
                     theta=op.theta, wtheta=op.wtheta)
    load, _ = st.zonal_load_rate(lmax, earth)
    state = op.zero_state()
    info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    rinfo = ref.step(dt, st.EXPONENTIAL_TRAPEZOID, amplitudes(load, lmax, dt), **kw)
    c, rc = state.export("coefficients"), ref.coef
    err = np.max(np.abs(c - rc)) / np.max(np.abs(rc))
    assert info.iterations == rinfo.iterations == 2, (info.iterations, rinfo.iterations)
    assert err < 1e-11, err
    print(f"smoke: 3D step lmax {lmax}, {info.iterations} passes, strain coefficients "
          f"rel err {err:.2e} against the oracle", flush=True)
