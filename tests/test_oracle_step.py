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


def delta_grid(earth, lmax, delta_scale, seed=4):
    """A lateral field on every memory node's grid: a node-dependent mix of
    sin^2(theta) cos(2 phi) and cos(theta), [node][theta][phi]."""
    elements, ngll = earth.mesh()
    first, count = (int(v) for v in earth.export("memory_elements"))
    nodes = count * (ngll - 1)
    theta, w, nt, nph = grid_of(lmax)
    rng = np.random.default_rng(seed)
    phi = 2 * np.pi * np.arange(nph) / nph
    T, P = np.meshgrid(theta, phi, indexing="ij")
    out = np.zeros((nodes, nt, nph))
    for k in range(nodes):
        a, b = rng.normal(size=2)
        out[k] = delta_scale * (a * np.sin(T) ** 2 * np.cos(2 * P) + b * np.cos(T))
    return out.ravel()


def test_lateral_itau_is_the_reference_scaled_by_exp_minus_delta(oracle, earth):
    """itau = itau_ref(node) exp(-delta) (src/Viscoelastic3D.cpp:250-283):
    delta = 0 leaves the reference's inverse time, delta = ln 2 halves it,
    and a field's factor is exp(-delta) point by point."""
    lmax = 6
    load_dt = 1.0
    zero = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, load_dt, None)
    per = zero.ntheta * zero.nphi
    ref = np.repeat(zero.itau_ref, per)
    assert np.array_equal(zero.arrays["itau"], ref)
    half = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, load_dt,
                      np.full(zero.nodes * per, np.log(2.0)))
    assert np.max(np.abs(half.arrays["itau"] - 0.5 * ref)) <= 1e-15 * np.max(ref)
    d = delta_grid(earth, lmax, 0.7)
    field = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, load_dt, d)
    nz = ref > 0
    ratio = field.arrays["itau"][nz] / ref[nz]
    assert np.max(np.abs(np.log(ratio) + d[nz])) <= 1e-13


def test_zero_lateral_step_closes_in_two_passes(oracle, earth):
    lmax = 8
    length, time, mass = earth.scales()
    dt = 100 * st.YEAR / time
    load, _ = st.zonal_load_rate(lmax, earth)
    p = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, dt, None)
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
    p = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, dt,
                   delta_grid(earth, lmax, 0.7))
    infos = []
    for k in range(2):
        amp = amplitudes(load, lmax, (k + 1) * dt)
        infos.append(p.step(dt, st.EXPONENTIAL_TRAPEZOID, amp, rtol=1e-10))
        assert infos[-1].converged == 1
    # lateral coupling costs passes beyond the zero-lateral two
    assert infos[-1].iterations > 3
