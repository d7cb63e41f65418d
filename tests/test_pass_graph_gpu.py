"""Passes replayed from captured hipGraphs (StepOptions graph_passes) against
passes launched kernel by kernel: the same kernels with the same arguments,
so the steps agree bit for bit. The trajectory covers what keys a captured
pass: both tensor pairs of the fixed point (Picard passes swap them, mixed
passes do not), a second flavour (another dt), the tide channel switched on
after steps without it, Sdirk2's stage flavour, and a warm start."""

import numpy as np
import pytest
from step_problem import small_earth_spec, tidal_rate

import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def earth():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return st.Earth(small_earth_spec())


def make(earth, lmax, graph):
    opts = st.StepOptions.default(max_iterations=60, rtol=1e-10, anderson_memory=6,
                                  anderson_rate=0.5, graph_passes=int(graph))
    op = st.Viscoelastic3D(earth, lmax, None, opts)
    coef = st.delta_coefficients(min(lmax, 6), op.nodes, seed=5, rms=1.0, lmax=lmax)
    op.set_delta(op.delta_field(coef))
    return op


def exports(state):
    out = [state.export("coefficients")]
    out += [state.export(f"m{b}") for b in range(3)]
    out += [state.export(f"strain{b}") for b in range(3)]
    return out


# lmax 8: the matrix-core ring kernels (nphi = 28 = 4 x 7); lmax 12: the
# generic ring kernel (nphi = 40, P = 10 even)
@pytest.mark.parametrize("lmax", [8, 12])
def test_graph_passes_equal_launched_passes(earth, lmax):
    import torch

    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    load, _ = st.zonal_load_rate(lmax, earth)
    tide = tidal_rate(lmax, earth)
    plan = [  # (dt, scheme, tide)
        (dt, st.EXPONENTIAL_TRAPEZOID, None),
        (dt, st.EXPONENTIAL_TRAPEZOID, None),
        (0.5 * dt, st.EXPONENTIAL_TRAPEZOID, None),
        (dt, st.EXPONENTIAL_TRAPEZOID, tide),
        (dt, st.SDIRK2, tide),
        (dt, st.BACKWARD_EULER, None),
    ]
    runs = []
    for graph in (False, True):
        print(f"[graph] lmax {lmax} graph_passes {int(graph)}", flush=True)
        op = make(earth, lmax, graph)
        state = op.zero_state()
        infos, snaps = [], []
        for k, (h, scheme, td) in enumerate(plan):
            infos.append(op.step(state, h, scheme, load, tide=td))
            snaps.append(exports(state))
        # a warm start: the last converged strain as the next step's start
        start = torch.from_numpy(state.export("coefficients_device")).cuda()
        infos.append(op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load, start=start))
        snaps.append(exports(state))
        runs.append((infos, snaps))
    (i0, s0), (i1, s1) = runs
    mixed = 0
    for a, b in zip(i0, i1):
        assert a.converged == b.converged == 1
        assert a.iterations == b.iterations
        assert a.mixed_passes == b.mixed_passes
        assert a.residual == b.residual
        mixed += a.mixed_passes
    assert mixed > 0  # the trajectory takes mixed passes as well as Picard ones
    for x, y in zip(s0, s1):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)
