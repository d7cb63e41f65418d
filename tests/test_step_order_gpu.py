"""Temporal order of the HIP step with lateral viscosity, after the
reference's LateralStepSelfConvergesAtSecondOrder
(tests/TestViscoelastic3D.cpp:323-349): delta = 0.6 sin(theta) sin(phi), a
constant load from the zero state, T = 2 tau_min advanced in 4, 8 and 16
steps; the observable is the surface U of degree 2, order 2 (real part) in
the last step's solutions. The ratio (x1 - x2) / (x2 - x4) is 4 +- 1.5 for
the exponential trapezoid (order two). The same run pins the other schemes'
orders: Sdirk2 (the recipe's two backward Euler stages) is also of order
two, backward Euler of order one (ratio 2)."""

import ctypes

import numpy as np
import pytest
from step_problem import small_earth_spec

import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu

LMAX = 8


@pytest.fixture(scope="module")
def setup():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    earth = st.Earth(small_earth_spec())
    opts = st.StepOptions.default(max_iterations=80, rtol=1e-12, anderson_memory=6,
                                  anderson_rate=0.5)
    op = st.Viscoelastic3D(earth, LMAX, None, opts)
    theta = torch.tensor(op.theta, dtype=torch.float64, device="cuda")
    phi = torch.arange(op.nphi, dtype=torch.float64, device="cuda") * (2 * np.pi / op.nphi)
    one = 0.6 * torch.sin(theta)[:, None] * torch.sin(phi)[None, :]
    delta = one.expand(op.nodes, -1, -1).contiguous().ravel()
    op.set_delta(delta)
    rng = np.random.default_rng(7)
    cols = {l: rng.uniform(-0.5, 0.5, 2 * (l + 1)) for l in range(LMAX + 1)}
    for l in cols:
        cols[l][1] = 0.0  # order 0 is real

    def load(l, t, columns, user):
        for k in range(2 * (l + 1)):
            columns[k] = cols[l][k]
        return 1

    load_fn = st.LOAD_FN(load)
    tau = 1.0 / np.max(earth.export("point_itau"))
    surface_u = int(earth.degree_system(2, 0, dt=0.0)["dofs"][-1][0])
    yield op, load_fn, tau, surface_u
    del delta


def advance(op, load_fn, scheme, T, steps, surface_u):
    state = op.zero_state()
    for _ in range(steps):
        info = op.step(state, T / steps, scheme, load_fn)
        assert info.converged == 1
    x = op.export("solution:2")  # spheroidal degree 2: ndof x 3 complex, column-major
    ndof = x.size // 6
    return x[2 * (2 * ndof + surface_u)]  # order 2, real part


@pytest.mark.parametrize("scheme,order", [(st.EXPONENTIAL_TRAPEZOID, 2), (st.SDIRK2, 2),
                                          (st.BACKWARD_EULER, 1)])
def test_lateral_step_self_converges_at_its_order(setup, scheme, order):
    op, load_fn, tau, surface_u = setup
    T = 2.0 * tau
    x1, x2, x4 = (advance(op, load_fn, scheme, T, n, surface_u) for n in (4, 8, 16))
    ratio = (x1 - x2) / (x2 - x4)
    print(f"[order] scheme {scheme} x {x1:.12e} {x2:.12e} {x4:.12e} ratio {ratio:.4f}",
          flush=True)
    expected = 2.0 ** order
    assert abs(ratio - expected) <= 0.375 * expected
