"""The oracle's 3D step on the CPU (small Earth, lmax 8): the zero-lateral
limit and the fixed point's own equation. With delta = 0 the lateral weight
vanishes (beta = beta_ref everywhere), the memory force is fixed and the
iteration closes in two passes (tests/TestViscoelastic3D.cpp:90-196 gates
the same limit against the 1D stepper). With a lateral field the converged
iterate is a fixed point of the pass to the tolerance."""

import numpy as np
import pytest
from step_problem import OracleStep, amplitudes, grid_of, small_earth_spec

import sl3dpp_amd.step as st


@pytest.fixture(scope="module")
def earth():
    return st.Earth(small_earth_spec())


def itau_field(earth, lmax, delta_scale, seed=4):
    from step_problem import oracle_library  # noqa: F401

    elements, ngll = earth.mesh()
    first, count = (int(v) for v in earth.export("memory_elements"))
    np_ = ngll - 1
    nodes = count * np_
    theta, w, nt, nph = grid_of(lmax)
    itau_ref = earth.export("point_itau").reshape(elements, np_)[first: first + count].ravel()
    rng = np.random.default_rng(seed)
    phi = 2 * np.pi * np.arange(nph) / nph
    T, P = np.meshgrid(theta, phi, indexing="ij")
    out = np.zeros((nodes, nt, nph))
    for k in range(nodes):
        a, b = rng.normal(size=2)
        delta = delta_scale * (a * np.sin(T) ** 2 * np.cos(2 * P) + b * np.cos(T))
        out[k] = itau_ref[k] * np.exp(-delta)
    return out.ravel()


def test_zero_lateral_step_closes_in_two_passes(oracle, earth):
    lmax = 8
    length, time, mass = earth.scales()
    dt = 100 * st.YEAR / time
    load, _ = st.zonal_load_rate(lmax, earth)
    p = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, dt,
                   itau_field(earth, lmax, 0.0))
    for k in range(2):
        amp = amplitudes(load, lmax, (k + 1) * dt)
        info = p.step(dt, st.EXPONENTIAL_TRAPEZOID, amp, rtol=1e-12)
        assert info.converged == 1
        assert info.iterations <= 3, info.iterations
    assert np.max(np.abs(p.coef)) > 0.0


def test_lateral_step_converges_to_its_fixed_point(oracle, earth):
    lmax = 8
    length, time, mass = earth.scales()
    dt = 100 * st.YEAR / time
    load, _ = st.zonal_load_rate(lmax, earth)
    itau = itau_field(earth, lmax, 0.7)
    p = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, dt, itau)
    infos = []
    for k in range(2):
        amp = amplitudes(load, lmax, (k + 1) * dt)
        infos.append(p.step(dt, st.EXPONENTIAL_TRAPEZOID, amp, rtol=1e-10))
        assert infos[-1].converged == 1
    # lateral coupling costs passes beyond the zero-lateral two
    assert infos[-1].iterations > 3
