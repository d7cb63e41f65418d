"""The time stepping against an external code: the oracle's degree-basis
Maxwell stepper (sl3o_step_degree_basis, the path the 3D step reproduces at
delta = 0: test_oracle_gates.py, test_reference_gates_gpu.py) against
pyslfp's stepper history in the reference's fixture
(tests/data/pyslfp_stepper.txt, copied to tests/golden/), as
tests/TestFixtures.cpp:255-368 checks it: the prem8 deck (not-a-knot cubic
joins) with a 1e21 Pa s Maxwell mantle on the solid layers inside
[3480, 6151] km, the lmax-10 mesh rule with 5 GLL nodes and Gauss memory
points, a unit surface load (1 kg m^-2 at every degree) held from t = 0,
50 steps of 3.15576e9 s with the exponential trapezoid and then with
backward Euler, each from rest with its elastic start; surface U and phi of
degrees 2, 5 and 10 at every step, to 1e-8 relative. The minimum
relaxation time is checked to 1e-10. The fixture's effective Love numbers
are not compared."""

import os

import numpy as np
import pytest
from reference_gates import DegreeBasisStepper
from step_problem import OracleStep
from test_traction_numbers import GOLDEN

import sl3dpp_amd.step as st
from sl3dpp_amd import radial

LMAX = 10
DEGREES = (2, 5, 10)
TOL = 1e-8


def viscous_mantle(layer, model):
    """1e21 Pa s on the solid layers fully inside [3480, 6151] km."""
    lo, hi = model.boundaries[layer], model.boundaries[layer + 1]
    if not model.is_fluid(layer) and lo >= 3.4799e6 and hi <= 6.1511e6:
        return 1e21
    return None


def read_fixture():
    tau, runs = None, []
    with open(os.path.join(GOLDEN, "pyslfp_stepper.txt")) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            w = line.split()
            if w[0] == "min_relaxation_time":
                tau = float(w[1])
            elif w[0] == "scheme":
                runs.append((w[1], []))
            elif w[0] == "effective_love":
                break
            else:
                runs[-1][1].append([float(x) for x in w])
    return tau, runs


@pytest.fixture(scope="module")
def prem8_maxwell():
    return radial.Earth.from_deck(os.path.join(GOLDEN, "prem8.elastic.deck"), lmax=LMAX, ngll=5,
                                  join="cubic", viscosity=viscous_mantle)


def test_minimum_relaxation_time(prem8_maxwell):
    tau, _ = read_fixture()
    itau = prem8_maxwell.branch_itau
    mine = 1.0 / np.max(itau[itau > 0])
    assert abs(mine - tau) <= 1e-10 * tau


@pytest.mark.parametrize("run", [0, 1])
def test_stepper_history_matches_pyslfp(oracle, prem8_maxwell, run):
    earth = prem8_maxwell
    _, runs = read_fixture()
    name, rows = runs[run]
    scheme = st.BACKWARD_EULER if name == "backward_euler" else st.EXPONENTIAL_TRAPEZOID
    length, time, mass = earth.scales()
    dt = 3.15576e9 / time
    potential = (length / time) ** 2
    elastic = OracleStep(oracle, earth, LMAX, st.BACKWARD_EULER, 0.0, None)
    stepping = OracleStep(oracle, earth, LMAX, scheme, dt, None)
    one = DegreeBasisStepper(oracle, stepping)
    # a unit surface mass at every degree, the m = 0 column
    amp = np.zeros(2 * (LMAX + 1) * (LMAX + 2) // 2)
    for l in range(LMAX + 1):
        amp[2 * (l * (l + 1) // 2)] = 1.0 / (mass / length ** 2)
    elements, ngll = earth.mesh()
    top = elements * (ngll - 1)
    surf = {l: (stepping.systems[l]["dofs"][top, 0], stepping.systems[l]["dofs"][top, 2])
            for l in DEGREES}
    # the elastic start: a zero-length step under the load at t = 0
    x = one.step(0.0, st.BACKWARD_EULER, amp, problem=elastic)
    worst = 0.0
    for k, row in enumerate(rows):
        t = row[0]
        if k > 0:
            x = one.step(dt, scheme, amp)
        assert abs(one.state.t * time - t) <= 1e-9 * max(t, 1.0)
        for j, l in enumerate(DEGREES):
            iu, ip = surf[l]
            U = x[l][2 * iu] * length
            phi = x[l][2 * ip] * potential
            u_ref, phi_ref = row[1 + 2 * j], row[2 + 2 * j]
            worst = max(worst, abs(U - u_ref) / abs(u_ref), abs(phi - phi_ref) / abs(phi_ref))
            assert abs(U - u_ref) <= TOL * abs(u_ref), (name, t, l, "U")
            assert abs(phi - phi_ref) <= TOL * abs(phi_ref), (name, t, l, "phi")
    print(f"{name}: {len(rows)} rows, worst relative difference {worst:.2e}")
