"""Warm starts along a trajectory (sl3dpp_amd.step.ExtrapolatedStart):
 * CPU: the Lagrange extrapolation weights reproduce every polynomial of
   degree below the point count exactly, and sum to one;
 * GPU: sl3_spectral_combine against numpy; a trajectory stepped from the
   extrapolated starts reaches the plain trajectory's states (the steps'
   own starts) step by step, in fewer passes. A history starts at the
   strain the first step starts from (at the state's time), so only the
   first step takes its own start."""

import numpy as np
import pytest

import sl3dpp_amd.step as st


@pytest.mark.parametrize("times,target", [([0.0, 1.0], 2.0), ([0.0, 1.0, 2.0], 3.0),
                                          ([0.3, 1.1, 1.7, 2.0], 2.45)])
def test_lagrange_weights_are_exact_for_polynomials(times, target):
    w = st.lagrange_extrapolation_weights(times, target)
    assert abs(sum(w) - 1.0) < 1e-12
    for degree in range(len(times)):
        p = np.polynomial.Polynomial(np.arange(1, degree + 2, dtype=float))
        got = sum(wj * p(tj) for wj, tj in zip(w, times))
        assert abs(got - p(target)) <= 1e-11 * max(1.0, abs(p(target)))
    if times == [0.0, 1.0, 2.0]:
        np.testing.assert_allclose(w, [1.0, -3.0, 3.0], atol=1e-14)


@pytest.fixture(scope="module")
def trajectory_problem():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    from step_problem import small_earth_spec

    earth = st.Earth(small_earth_spec())
    _, tscale, _ = earth.scales()
    dt = 100.0 * st.YEAR / tscale
    opts = st.StepOptions.default(max_iterations=200, rtol=1e-10, anderson_memory=6,
                                  anderson_rate=0.5)
    lmax = 16
    op = st.Viscoelastic3D(earth, lmax, None, opts)
    coef = st.delta_coefficients(8, op.nodes, seed=21, rms=0.7, lmax=lmax)
    op.set_delta(op.delta_field(coef))
    load, _ = st.zonal_load_rate(lmax, earth)
    return op, dt, load


@pytest.mark.gpu
def test_combine_against_numpy(trajectory_problem):
    import torch

    op, _, _ = trajectory_problem
    rng = np.random.default_rng(5)
    a = rng.normal(size=op.spectral_doubles)
    b = rng.normal(size=op.spectral_doubles)
    c = rng.normal(size=op.spectral_doubles)
    da, db, dc = (torch.from_numpy(v).cuda() for v in (a, b, c))
    out = torch.empty_like(da)
    torch.cuda.synchronize()
    st.spectral_combine(op, [3.0, -3.0, 1.0], [da.data_ptr(), db.data_ptr(), dc.data_ptr()], out)
    np.testing.assert_allclose(out.cpu().numpy(), 3 * a - 3 * b + c, rtol=0, atol=1e-14)
    # dst aliasing a source
    st.spectral_combine(op, [0.5, 2.0], [da.data_ptr(), db.data_ptr()], da)
    np.testing.assert_allclose(da.cpu().numpy(), 0.5 * a + 2 * b, rtol=0, atol=1e-14)
    # eight terms, the most a combination takes (an order-5 extrapolation
    # uses six)
    w = [6.0, -15.0, 20.0, -15.0, 6.0, -1.0, 0.25, 0.5]
    vs = [rng.normal(size=op.spectral_doubles) for _ in w]
    ds = [torch.from_numpy(v).cuda() for v in vs]
    st.spectral_combine(op, w, [d.data_ptr() for d in ds], out)
    want = sum(c * v for c, v in zip(w, vs))
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=0, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2, 3, 5])
def test_extrapolated_trajectory_matches_the_plain_one(trajectory_problem, order):
    op, dt, load = trajectory_problem
    plain_state = op.zero_state()
    warm_state = op.zero_state()
    ws = st.ExtrapolatedStart(op, order=order)
    plain_passes, warm_passes = [], []
    for k in range(8):
        info = op.step(plain_state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        plain_passes.append(info.iterations)
        info = ws.step(warm_state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        warm_passes.append(info.iterations)
        assert info.converged
        a = plain_state.export("coefficients")
        b = warm_state.export("coefficients")
        assert np.max(np.abs(a - b)) <= 1e-7 * np.max(np.abs(a)), k
        for name in ("m0", "m2"):
            a = plain_state.export(name)
            b = warm_state.export(name)
            assert np.max(np.abs(a - b)) <= 1e-7 * np.max(np.abs(a)), (k, name)
    print(f"order {order}: plain passes {plain_passes}, extrapolated {warm_passes}")
    # a zero-length step keeps one history entry per time: the next
    # extrapolated step is finite and converges to the plain one's state
    info = ws.step(warm_state, 0.0, st.EXPONENTIAL_TRAPEZOID, load)
    op.step(plain_state, 0.0, st.EXPONENTIAL_TRAPEZOID, load)
    assert info.converged
    info = ws.step(warm_state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    op.step(plain_state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    assert info.converged and np.isfinite(info.residual)
    a = plain_state.export("coefficients")
    b = warm_state.export("coefficients")
    assert np.max(np.abs(a - b)) <= 1e-7 * np.max(np.abs(a))
    # a state other than the last one stepped starts a new history: its
    # step takes the step's own start, the plain step's pass count, even
    # when it is at the last one's time (a copy of it)
    other = op.zero_state()
    info = ws.step(other, dt, st.EXPONENTIAL_TRAPEZOID, load)
    assert info.iterations == plain_passes[0]
    ws.step(warm_state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    ws.step(warm_state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    twin = warm_state.clone()
    plain_twin = warm_state.clone()
    own = op.step(plain_twin, dt, st.EXPONENTIAL_TRAPEZOID, load)
    info = ws.step(twin, dt, st.EXPONENTIAL_TRAPEZOID, load)
    assert info.iterations == own.iterations
    # the first step has no history (its start strain is its first entry);
    # afterwards the start is better (on this small Earth the higher orders
    # gain less than at lmax 128; the pass counts are checked to order 2)
    assert warm_passes[0] == plain_passes[0]
    if order <= 2:
        assert sum(warm_passes[3:]) < sum(plain_passes[3:])
