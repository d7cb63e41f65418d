"""The step closed by preconditioned conjugate gradients
(Options::StepSolver::ConjugateGradients: OpenStep::Forces and Solve,
src/Viscoelastic3D.cpp:1448-1496 and 2008-2015, through
detail::LateralConjugateGradients, include/sl3dpp/detail/LateralSolve.hpp:
110-183), after the reference's ConjugateGradientStepMatchesTheFixedPoint
(tests/TestViscoelastic3D.cpp:664-726):
 * at a 10x contrast (delta = 0.5 ln 10 sin(theta) cos(phi)) over two steps
   of 2 tau_min with solve_rtol 1e-12, the solutions of every degree,
   spheroidal and toroidal, are the fixed point's to 1e-9 and no pass is
   mixed; the same with the tide channel beside the load, and for Sdirk2
   (one solve per stage);
 * the state after conjugate-gradient steps is the oracle's converged fixed
   point's (1e-9);
 * the zero-lateral step needs no iteration;
 * the solve's scale <x0, f>, x0 = K^-1 f (the relative residual's
   denominator), is the pairing of the exported x0 with the assembled band
   systems applied to it (a solve capped at zero iterations returns x0)."""

import math

import numpy as np
import pytest
from step_problem import OracleStep, amplitudes, small_earth_spec, tidal_rate

import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu

LMAX = 8


def gap(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def earth():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return st.Earth(small_earth_spec())


def load_columns(lmax, seed=5):
    """A fixed load: seeded columns per degree, order 0 real (the
    reference's LoadColumns is likewise fixed noise per (l, m))."""
    rng = np.random.default_rng(seed)
    cols = {}
    for l in range(lmax + 1):
        c = rng.uniform(-0.5, 0.5, 2 * (l + 1))
        c[1] = 0.0
        cols[l] = c

    def load(l, t, columns, user):
        for k in range(2 * (l + 1)):
            columns[k] = cols[l][k]
        return 1

    return st.LOAD_FN(load)


def contrast_operator(earth, solver, lmax=LMAX, **overrides):
    import torch

    opts = st.StepOptions.default(max_iterations=200, rtol=1e-12, solve_rtol=1e-12,
                                  anderson_memory=6, anderson_rate=0.5, solver=solver)
    for k, v in overrides.items():
        setattr(opts, k, v)
    op = st.Viscoelastic3D(earth, lmax, None, opts)
    theta = torch.tensor(op.theta, dtype=torch.float64, device="cuda")
    phi = torch.arange(op.nphi, dtype=torch.float64, device="cuda") * (2 * math.pi / op.nphi)
    one = 0.5 * math.log(10.0) * torch.sin(theta)[:, None] * torch.cos(phi)[None, :]
    op.set_delta(one.expand(op.nodes, -1, -1).contiguous().ravel())
    return op


def tau_min(earth):
    return 1.0 / np.max(earth.export("point_itau"))


def all_tasks(lmax):
    return list(range(2 * lmax + 1))  # spheroidal 0..L, toroidal 1..L


def solutions(op, lmax):
    return [op.export(f"solution:{t}") for t in all_tasks(lmax)]


def solutions_gap(a, b):
    """SolutionsGap (tests/TestViscoelastic3D.cpp:643-661): the largest
    difference over every degree's columns against the largest value."""
    scale = max(np.max(np.abs(y)) for y in b)
    return max(np.max(np.abs(x - y)) for x, y in zip(a, b)) / scale


@pytest.mark.parametrize("tide", [False, True])
def test_conjugate_gradient_step_matches_the_fixed_point(earth, tide):
    dt = 2.0 * tau_min(earth)
    load = load_columns(LMAX)
    tide_fn = tidal_rate(LMAX, earth) if tide else None
    runs = {}
    for solver in (st.FIXED_POINT, st.CONJUGATE_GRADIENTS):
        op = contrast_operator(earth, solver)
        state = op.zero_state()
        out = []
        for k in range(2):
            print(f"[cg] solver {solver} tide {tide} step {k}", flush=True)
            info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load, tide=tide_fn)
            assert info.converged == 1
            out.append((info, solutions(op, LMAX), state.export("coefficients")))
        runs[solver] = out
    for k in range(2):
        fp, cg = runs[st.FIXED_POINT][k], runs[st.CONJUGATE_GRADIENTS][k]
        print(f"[cg] step {k}: {fp[0].iterations} fixed-point passes "
              f"({fp[0].mixed_passes} mixed), {cg[0].solve_iterations} conjugate-gradient "
              f"iterations", flush=True)
        assert cg[0].mixed_passes == 0
        assert cg[0].solve_iterations == cg[0].iterations > 0
        assert cg[0].residual <= 1e-12
        assert solutions_gap(cg[1], fp[1]) < 1e-9
        assert gap(cg[2], fp[2]) < 1e-9


def test_conjugate_gradient_sdirk2_matches_the_fixed_point(earth):
    dt = 2.0 * tau_min(earth)
    load = load_columns(LMAX)
    got = {}
    for solver in (st.FIXED_POINT, st.CONJUGATE_GRADIENTS):
        op = contrast_operator(earth, solver)
        state = op.zero_state()
        print(f"[cg] sdirk2 solver {solver}", flush=True)
        info = op.step(state, dt, st.SDIRK2, load)
        assert info.converged == 1
        got[solver] = (info, solutions(op, LMAX), [state.export(f"m{b}") for b in range(3)])
    cg = got[st.CONJUGATE_GRADIENTS]
    assert cg[0].solve_iterations == cg[0].iterations > 0  # both stages' solves
    assert cg[0].mixed_passes == 0
    assert solutions_gap(cg[1], got[st.FIXED_POINT][1]) < 1e-9
    for b in range(3):
        assert gap(cg[2][b], got[st.FIXED_POINT][2][b]) < 1e-9


def test_conjugate_gradient_steps_match_the_oracle_fixed_point(earth, oracle):
    lmax = 12
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    op = st.Viscoelastic3D(earth, lmax, None, st.StepOptions.default(
        solver=st.CONJUGATE_GRADIENTS, solve_rtol=1e-13))
    coef = st.delta_coefficients(6, op.nodes, seed=11, rms=1.5, lmax=lmax)
    op.set_delta(op.delta_field(coef))
    ref = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, dt, op.delta_grid(),
                     theta=op.theta, wtheta=op.wtheta)
    load, _ = st.zonal_load_rate(lmax, earth)
    state = op.zero_state()
    for k in range(2):
        print(f"[cg] oracle lmax {lmax} step {k}", flush=True)
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        assert info.converged == 1 and info.mixed_passes == 0
        rinfo = ref.step(dt, st.EXPONENTIAL_TRAPEZOID, amplitudes(load, lmax, (k + 1) * dt),
                         max_iterations=200, rtol=1e-13, anderson_memory=6, anderson_rate=0.5)
        assert rinfo.converged == 1
        assert gap(state.export("coefficients"), ref.coef) < 1e-9
        for b in range(3):
            assert gap(state.export(f"m{b}"), ref.m[b]) < 1e-9
            assert gap(state.export(f"strain{b}"), ref.strain[b]) < 1e-9


def test_zero_lateral_step_needs_no_iteration(earth):
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    load, _ = st.zonal_load_rate(LMAX, earth)
    op = st.Viscoelastic3D(earth, LMAX, None,
                           st.StepOptions.default(solver=st.CONJUGATE_GRADIENTS))
    state = op.zero_state()
    info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    assert info.converged == 1
    assert info.iterations == info.solve_iterations == 0
    assert info.residual == 0.0


def dense(band):
    n, w = band.shape
    kd = w - 1
    a = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - kd), j + 1):
            a[i, j] = a[j, i] = band[j, kd + i - j]
    return a


def test_solve_scale_is_the_energy_of_the_reference_response(earth):
    """<x0, f> with f = K x0 over every dof, the pairing weighting m > 0
    twice (LateralPair, LateralSolve.hpp:66-91): pins the memory-dof
    decomposition of the scale, the channels' unit terms included (the tide
    moves it by more than a per cent)."""
    dt = 2.0 * tau_min(earth)
    load = load_columns(LMAX)
    scales = []
    for tide_fn in (None, tidal_rate(LMAX, earth)):
        op = contrast_operator(earth, st.CONJUGATE_GRADIENTS, solve_max_iterations=0)
        state = op.zero_state()
        op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load, tide=tide_fn)
        # the second step: a nonzero memory force beside the channels
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load, tide=tide_fn)
        assert info.solve_iterations == 0
        energy = op.export("solve_energy")[0]
        tasks = op.export("tasks").reshape(-1, 5)
        total = 0.0
        for task, row in enumerate(tasks):
            l, kind = int(row[0]), int(row[1])
            s = earth.degree_system(l, kind, st.EXPONENTIAL_TRAPEZOID, dt)
            k = dense(s["band"])
            x = op.export(f"solution:{task}").reshape(l + 1, s["ndof"], 2)
            xc = x[..., 0] + 1j * x[..., 1]
            f = xc @ k.T  # K x per column (K real symmetric)
            w = np.where(np.arange(l + 1) == 0, 1.0, 2.0)
            total += float(np.sum(w[:, None] * (xc * np.conj(f)).real))
        print(f"[cg] tide {tide_fn is not None} scale {energy:.15e} direct {total:.15e}",
              flush=True)
        assert abs(energy - total) <= 1e-9 * abs(total)
        scales.append(energy)
    assert abs(scales[1] - scales[0]) > 1e-2 * abs(scales[0])
