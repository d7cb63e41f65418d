"""The oracle's 3D step pinned by the reference's own gates (on the CPU):

 * the zero-lateral pipeline (tests/TestGridCoupling.cpp:178-332): strain
   recovery, synthesis, the product with a unit weight, analysis, the
   node scaling by mu and the force spread, composed, equal the point-shear
   part of the assembled systems applied to the displacement (1e-11), for
   spheroidal and toroidal degrees; a purely spheroidal displacement exerts
   no toroidal force and the reverse;
 * the zero-lateral step (tests/TestViscoelastic3D.cpp:90-196): with
   delta = 0 the oracle's 3D step (transforms, grid weights, Converge,
   Finish) gives the degree-basis stepper's solutions over two steps under a
   load with every order growing from zero (1e-10), exponential trapezoid and backward Euler.

The point-shear systems are themselves pinned: the elastic block by the
traction numbers against pyslfp (test_traction_numbers.py) and the point
shear by its identity with a raised shear modulus at the GLL nodes."""

import numpy as np
import pytest
from reference_gates import (DegreeBasisStepper, forces_gap, general_load, oracle_pipeline_force,
                             oracle_solutions, point_shear_force, random_displacement, tasks_of,
                             zero_lateral_problem)
from step_problem import amplitudes, small_earth_spec

import sl3dpp_amd.step as st


@pytest.fixture(scope="module")
def earth():
    return st.Earth(small_earth_spec())


@pytest.mark.parametrize("lmax", [6, 9])
def test_zero_lateral_pipeline_is_the_point_shear_operator(oracle, earth, lmax):
    L, T, M = earth.scales()
    ref = zero_lateral_problem(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, 100 * st.YEAR / T)
    x = random_displacement(ref.systems, lmax)
    want = point_shear_force(earth, lmax, x)
    got, _ = oracle_pipeline_force(oracle, ref, x)
    pinned = [s["pinned"] for s in ref.systems]
    gap, scale = forces_gap(got, want, pinned)
    print(f"pipeline lmax {lmax}: force gap {gap:.2e} of {scale:.2e}")
    assert scale > 0.0
    assert gap < 1e-11


def test_spheroidal_and_toroidal_forces_decouple(oracle, earth):
    lmax = 7
    L, T, M = earth.scales()
    ref = zero_lateral_problem(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, 100 * st.YEAR / T)
    x = random_displacement(ref.systems, lmax)
    kinds = [kind for (l, kind) in tasks_of(lmax)]
    full, _ = oracle_pipeline_force(oracle, ref, x)
    scale = max(np.max(np.abs(f)) for f in full)
    for keep in (0, 1):
        only = [z if k == keep else np.zeros_like(z) for z, k in zip(x, kinds)]
        got, _ = oracle_pipeline_force(oracle, ref, only)
        other = max(np.max(np.abs(f)) for f, k in zip(got, kinds) if k != keep)
        assert other < 1e-11 * scale, (keep, other / scale)


@pytest.mark.parametrize("scheme", [st.EXPONENTIAL_TRAPEZOID, st.BACKWARD_EULER])
def test_zero_lateral_step_matches_the_degree_basis_stepper(oracle, earth, scheme):
    lmax = 8
    L, T, M = earth.scales()
    tau = 1.0 / np.max(earth.export("point_itau"))
    dt = 0.7 * tau
    ref = zero_lateral_problem(oracle, earth, lmax, scheme, dt)
    one = DegreeBasisStepper(oracle, ref)
    load = general_load(lmax, earth)
    for k in range(2):
        amp = amplitudes(load, lmax, (k + 1) * dt)
        info = ref.step(dt, scheme, amp, max_iterations=10, rtol=1e-13)
        assert info.converged == 1 and info.iterations <= 3  # no lateral term
        x1 = one.step(dt, scheme, amp)
        x3 = oracle_solutions(oracle, 2 * lmax + 1)
        scale = max(np.max(np.abs(x)) for x in x1)
        gap = max(np.max(np.abs(a - b)) for a, b in zip(x3[: lmax + 1], x1))
        tor = max(np.max(np.abs(x)) for x in x3[lmax + 1:])
        print(f"scheme {scheme} step {k}: solution gap {gap / scale:.2e}, toroidal {tor:.1e}")
        assert scale > 0.0
        assert gap < 1e-10 * scale
        assert tor < 1e-12 * scale
