"""The step at the benchmark's full size (lmax 128, the 194 x 388 grid, 128
elements of 6 GLL nodes, 280 Gauss memory nodes, the (1 + l/8)^-3 lateral
field of rms 0.7 with AR(1) 0.9 across the nodes, seed 2026, exponential
trapezoid over 100 yr, the zonal load rate).

Against the oracle, on the same inputs: two steps each capped at six
passes (StartStep, Picard passes until the adaptive Anderson mixer engages,
mixed passes, Finish), the strain coefficients, memory and strain grids
compared after each step.

Converged steps, where the oracle would take minutes: two independent
closures of the same linear system, the within-step fixed point (with adaptive Anderson mixing)
and the preconditioned conjugate gradients, must land on the same state.
Both converge; their strain coefficients agree to 1e-9 (measured 5.7e-11
with rtol 1e-11 and solve_rtol 1e-13) and their memory on the grid to 1e-7
(measured 1.7e-9 to 2.0e-8 by block: the grid values relative to the
largest of a block)."""

import numpy as np
import pytest
from step_problem import OracleStep, amplitudes

import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu

LMAX = 128


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def earth():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return st.Earth(st.three_region_earth(70, 56, 2, 6))


def full_step(earth, solver):
    L, T, M = earth.scales()
    dt = 100.0 * st.YEAR / T
    opts = st.StepOptions.default(max_iterations=200, rtol=1e-11, anderson_memory=6,
                                  anderson_rate=0.5, solver=solver, solve_rtol=1e-13)
    op = st.Viscoelastic3D(earth, LMAX, None, opts)
    assert (op.nodes, op.ntheta, op.nphi) == (280, 194, 388)
    coef = st.delta_coefficients(64, op.nodes, seed=2026, rms=0.7, ar=0.9, lmax=LMAX)
    op.set_delta(op.delta_field(coef))
    load, _ = st.zonal_load_rate(LMAX, earth)
    state = op.zero_state()
    infos = []
    for k in range(2):
        print(f"[full] solver {solver} step {k}", flush=True)
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        infos.append(info)
        assert info.converged == 1
    return infos, state.export("coefficients"), [state.export(f"m{b}") for b in range(3)]


def test_full_size_fixed_point_and_conjugate_gradients_agree(earth):
    fp_info, fp_c, fp_m = full_step(earth, st.FIXED_POINT)
    cg_info, cg_c, cg_m = full_step(earth, st.CONJUGATE_GRADIENTS)
    print(f"[full] fixed point {[i.iterations for i in fp_info]} passes, "
          f"conjugate gradients {[i.solve_iterations for i in cg_info]} iterations", flush=True)
    assert all(i.mixed_passes > 0 for i in fp_info)  # Anderson engaged
    assert all(i.mixed_passes == 0 and i.solve_iterations > 0 for i in cg_info)
    gap = rel(cg_c, fp_c)
    print(f"[full] coefficients gap {gap:.2e}", flush=True)
    assert gap < 1e-9
    for b in range(3):
        gm = rel(cg_m[b], fp_m[b])
        print(f"[full] memory block {b} gap {gm:.2e}", flush=True)
        assert gm < 1e-7


def test_full_size_passes_match_the_oracle(earth, oracle):
    L, T, M = earth.scales()
    dt = 100.0 * st.YEAR / T
    kw = dict(max_iterations=6, rtol=1e-14, anderson_memory=6, anderson_rate=0.5)
    op = st.Viscoelastic3D(earth, LMAX, None, st.StepOptions.default(**kw))
    coef = st.delta_coefficients(64, op.nodes, seed=2026, rms=0.7, ar=0.9, lmax=LMAX)
    op.set_delta(op.delta_field(coef))
    print("[full] building the oracle problem", flush=True)
    ref = OracleStep(oracle, earth, LMAX, st.EXPONENTIAL_TRAPEZOID, dt, op.delta_grid(),
                     theta=op.theta, wtheta=op.wtheta)
    load, _ = st.zonal_load_rate(LMAX, earth)
    state = op.zero_state()
    for k in range(2):
        print(f"[full] oracle parity step {k}", flush=True)
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        rinfo = ref.step(dt, st.EXPONENTIAL_TRAPEZOID, amplitudes(load, LMAX, (k + 1) * dt), **kw)
        assert info.iterations == rinfo.iterations == 6
        assert info.mixed_passes == rinfo.mixed_passes > 0
        assert abs(info.residual - rinfo.residual) <= 1e-6 * rinfo.residual
        gap = rel(state.export("coefficients"), ref.coef)
        print(f"[full] step {k}: residual {info.residual:.3e} / {rinfo.residual:.3e}, "
              f"mixed {info.mixed_passes}, coefficients {gap:.2e}", flush=True)
        assert gap < 1e-10
        for b in range(3):
            gm = rel(state.export(f"m{b}"), ref.m[b])
            gs = rel(state.export(f"strain{b}"), ref.strain[b])
            print(f"[full] step {k} block {b}: memory {gm:.2e}, strain {gs:.2e}", flush=True)
            assert gm < 1e-10 and gs < 1e-10
