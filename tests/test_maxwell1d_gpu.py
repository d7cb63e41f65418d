"""The batched radially symmetric Maxwell evolution on the GPU
(sl3_maxwell1d_*, ViscoelasticOperator's Step for every spheroidal degree
at once, include/sl3dpp/Viscoelastic.hpp:163-301):

 * against pyslfp's stepper history (the reference's fixture tests/data/
   pyslfp_stepper.txt, as tests/TestFixtures.cpp:255-368 checks it): prem8
   with a 1e21 Pa s mantle, lmax 10, 5 GLL nodes, a unit surface load held
   from t = 0, 50 steps of 3.15576e9 s under each scheme from rest with the
   elastic start; surface U and phi of degrees 2, 5 and 10 at every step
   to 1e-8 relative, read through sl3_maxwell1d_surface, each step one
   replay of the captured graph;
 * against the oracle's degree-basis stepper (sl3o_step_degree_basis) on
   the three-region Earth with the 1D configuration's viscosity (5e20 Pa s
   above 5701 km, 1e22 below, elastic lithosphere) at 26 / 21 / 1 elements
   x 6 GLL nodes, lmax 24, a seeded load pattern on every order ramped
   linearly to zero, 8 steps of 100 yr per scheme: memory and strain of
   every degree to 1e-11 of each degree's scale, the solutions to 1e-10
   (degree 1's system, the worst conditioned, sits near 1e-11 once the
   load has ramped to zero);
 * graph replays and stream launches give the same bits, and a reset
   trajectory repeats itself bit for bit; every band panel the tuner
   chooses from gives the same bits;
 * the argument checks (scheme, ramp length, pattern size)."""

import ctypes

import numpy as np
import pytest
from reference_gates import DegreeBasisStepper
from step_problem import DP, OracleStep
from test_stepper_pins import DEGREES, LMAX, TOL, prem8_maxwell, read_fixture  # noqa: F401

import sl3dpp_amd
import sl3dpp_amd.step as st
from sl3dpp_amd import radial

pytestmark = pytest.mark.gpu


def unit_load_pattern(lmax, earth):
    """1 kg m^-2 at every degree, the m = 0 column."""
    length, time, mass = earth.scales()
    p = np.zeros((lmax + 1) * (lmax + 2))
    for l in range(lmax + 1):
        p[2 * (l * (l + 1) // 2)] = 1.0 / (mass / length ** 2)
    return p


@pytest.mark.parametrize("run", [0, 1])
def test_surface_history_matches_pyslfp(prem8_maxwell, run):  # noqa: F811
    import torch

    earth = prem8_maxwell
    _, runs = read_fixture()
    name, rows = runs[run]
    scheme = st.BACKWARD_EULER if name == "backward_euler" else st.EXPONENTIAL_TRAPEZOID
    length, time, mass = earth.scales()
    dt = 3.15576e9 / time
    potential = (length / time) ** 2
    op = radial.ViscoelasticOperator(earth, LMAX)
    op.set_load(unit_load_pattern(LMAX, earth), np.ones(len(rows) + 1))
    out_u = torch.empty((LMAX + 1) * (LMAX + 2), dtype=torch.float64, device="cuda")
    out_p = torch.empty_like(out_u)
    worst = 0.0
    for k, row in enumerate(rows):
        # the first call takes the elastic start at t = 0 (zero steps)
        op.evolve(dt, scheme, 0 if k == 0 else 1, graph=True)
        assert op.steps_taken() == k
        assert abs(k * dt * time - row[0]) <= 1e-9 * max(row[0], 1.0)
        u = op.surface(radial.U, out_u).cpu().numpy()
        p = op.surface(radial.PHI, out_p).cpu().numpy()
        for j, l in enumerate(DEGREES):
            U = u[l * (l + 1)] * length
            phi = p[l * (l + 1)] * potential
            u_ref, phi_ref = row[1 + 2 * j], row[2 + 2 * j]
            worst = max(worst, abs(U - u_ref) / abs(u_ref), abs(phi - phi_ref) / abs(phi_ref))
            assert abs(U - u_ref) <= TOL * abs(u_ref), (name, k, l, "U")
            assert abs(phi - phi_ref) <= TOL * abs(phi_ref), (name, k, l, "phi")
    print(f"{name}: {len(rows)} rows on the GPU, worst relative difference {worst:.2e}")


C1_SMALL = dict(core=26, mantle=21, lithosphere=1, ngll=6)
C1_LMAX = 24
C1_STEPS = 8


@pytest.fixture(scope="module")
def c1_small():
    return st.Earth(st.three_region_earth(**C1_SMALL))


def seeded_pattern(lmax, earth, seed=41):
    """Seeded complex coefficients on every (l, m), m = 0 real, of about a
    kilometre of ice at low degrees: (Re, Im) doubles."""
    length, time, mass = earth.scales()
    sigma = mass / length ** 2
    rng = np.random.default_rng(seed)
    out = np.zeros((lmax + 1) * (lmax + 2))
    for l in range(lmax + 1):
        for m in range(l + 1):
            at = 2 * (l * (l + 1) // 2 + m)
            scale = 917.0 * 1000.0 / sigma * (1.0 + l) ** -1.5
            out[at] = rng.normal() * scale
            out[at + 1] = 0.0 if m == 0 else rng.normal() * scale
    return out


def oracle_layouts(op, l, nodes, m_flat, sol_flat):
    """The oracle's per-degree memory ((node, order, scalar, part) doubles)
    and solutions ((order, dof, part)) in the GPU's layouts: memory (nodes,
    3, 2 (l+1)), solutions (ndof, 2 (l+1))."""
    cols = l + 1
    m = m_flat.reshape(nodes, cols, 3, 2)
    mem = np.concatenate([m[..., 0].transpose(0, 2, 1), m[..., 1].transpose(0, 2, 1)], axis=2)
    nd = op.degrees[l]["ndof"]
    x = sol_flat.reshape(cols, nd, 2)
    sol = np.concatenate([x[..., 0].T, x[..., 1].T], axis=1)
    return mem, sol


@pytest.mark.parametrize("scheme", [st.EXPONENTIAL_TRAPEZOID, st.BACKWARD_EULER])
def test_matches_the_oracle_stepper(oracle, c1_small, scheme):
    earth = c1_small
    length, time, mass = earth.scales()
    dt = 100.0 * st.YEAR / time
    pattern = seeded_pattern(C1_LMAX, earth)
    ramp = np.linspace(1.0, 0.0, C1_STEPS + 1)
    op = radial.ViscoelasticOperator(earth, C1_LMAX)
    op.set_load(pattern, ramp)
    elastic = OracleStep(oracle, earth, C1_LMAX, st.BACKWARD_EULER, 0.0, None)
    stepping = OracleStep(oracle, earth, C1_LMAX, scheme, dt, None)
    ref = DegreeBasisStepper(oracle, stepping)
    assert ref.ref.problem.nodes == op.nodes
    x = ref.step(0.0, st.BACKWARD_EULER, pattern * ramp[0], problem=elastic)
    op.evolve(dt, scheme, 0)
    checkpoints = {0, 1, C1_STEPS}
    worst = {"memory": 0.0, "strain": 0.0, "solutions": 0.0}
    for k in range(C1_STEPS + 1):
        if k > 0:
            x = ref.step(dt, scheme, pattern * ramp[k])
            op.evolve(dt, scheme, 1, graph=True)
        if k not in checkpoints:
            continue
        mem, strain, sol = (op.export(n) for n in ("memory", "strain", "solutions"))
        off = 0
        for l in range(C1_LMAX + 1):
            n = op.nodes * (l + 1) * 6
            want_m, want_x = oracle_layouts(op, l, op.nodes, ref.m[off: off + n], x[l])
            want_d, _ = oracle_layouts(op, l, op.nodes, ref.d[off: off + n], x[l])
            off += n
            got_x = op.solution_of(sol, l)
            sx = np.max(np.abs(want_x))
            gap = np.max(np.abs(got_x - want_x)) / sx
            worst["solutions"] = max(worst["solutions"], gap)
            assert gap <= 1e-10, (k, l, "solutions", gap)
            for name, got, want in (("memory", op.memory_of(mem, l), want_m),
                                    ("strain", op.memory_of(strain, l), want_d)):
                s = np.max(np.abs(want))
                if s == 0.0:
                    assert np.max(np.abs(got)) == 0.0
                    continue
                gap = np.max(np.abs(got - want)) / s
                worst[name] = max(worst[name], gap)
                assert gap <= 1e-11, (k, l, name, gap)
    print(f"scheme {scheme}: GPU vs oracle over {C1_STEPS} steps, lmax {C1_LMAX}: " +
          ", ".join(f"{k} {v:.2e}" for k, v in worst.items()))


def test_graph_replays_equal_stream_launches(c1_small):
    earth = c1_small
    length, time, mass = earth.scales()
    dt = 100.0 * st.YEAR / time
    pattern = seeded_pattern(C1_LMAX, earth)
    ramp = np.linspace(1.0, 0.0, C1_STEPS + 1)
    out = []
    op = radial.ViscoelasticOperator(earth, C1_LMAX)
    op.set_load(pattern, ramp)
    for graph in (True, False, True):
        op.reset()
        op.evolve(dt, st.EXPONENTIAL_TRAPEZOID, C1_STEPS, graph=graph)
        out.append((op.export("memory"), op.export("strain"), op.export("solutions")))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    for a, b in zip(out[0], out[2]):
        assert np.array_equal(a, b)
    # the trajectory in pieces equals it in one call
    op.reset()
    op.evolve(dt, st.EXPONENTIAL_TRAPEZOID, 3, graph=True)
    op.evolve(dt, st.EXPONENTIAL_TRAPEZOID, C1_STEPS - 3, graph=True)
    assert np.array_equal(op.export("memory"), out[0][0])
    assert op.steps_taken() == C1_STEPS


def test_every_band_panel_gives_the_same_bits(c1_small):
    """The tuner's band panels (4, 8, 16 rows staged per panel of the
    batched Cholesky solve) give the same trajectory bit for bit; an
    evolution reports the panel it used and the timings that chose it."""
    earth = c1_small
    length, time, mass = earth.scales()
    dt = 100.0 * st.YEAR / time
    pattern = seeded_pattern(C1_LMAX, earth)
    ramp = np.linspace(1.0, 0.0, C1_STEPS + 1)
    out = []
    for panel in (4, 8, 16, 0):
        sl3dpp_amd.tuning_set(band_panel_rows=panel)
        try:
            op = radial.ViscoelasticOperator(earth, C1_LMAX)
            op.set_load(pattern, ramp)
            op.evolve(dt, st.EXPONENTIAL_TRAPEZOID, C1_STEPS, graph=True)
        finally:
            sl3dpp_amd.tuning_set()
        t = op.export("tuning")
        if panel:
            assert int(t[0]) == panel
        else:  # tuned: the fastest panel timed
            assert int(t[0]) in (4, 8, 16)
            assert t[1 + (4, 8, 16).index(int(t[0]))] == np.min(t[1:])
        assert np.all(t[1:] > 0)
        out.append((op.export("memory"), op.export("strain"), op.export("solutions")))
    for run in out[1:]:
        for a, b in zip(out[0], run):
            assert np.array_equal(a, b)


def test_surface_equals_the_solutions_surface_rows(c1_small):
    earth = c1_small
    length, time, mass = earth.scales()
    dt = 100.0 * st.YEAR / time
    op = radial.ViscoelasticOperator(earth, C1_LMAX)
    op.set_load(seeded_pattern(C1_LMAX, earth), np.linspace(1.0, 0.0, C1_STEPS + 1))
    op.evolve(dt, st.BACKWARD_EULER, 2)
    sol = op.export("solutions")
    for comp in (radial.U, radial.V, radial.PHI):
        s = op.surface(comp).cpu().numpy()
        for l in range(C1_LMAX + 1):
            deg = radial.DegreeOperator(earth, l)
            dof = deg.dof(earth.mesh()[0] * (earth.mesh()[1] - 1), comp)
            want = op.solution_of(sol, l)[dof] if dof >= 0 else np.zeros(2 * (l + 1))
            got = s[l * (l + 1): (l + 1) * (l + 2)]
            # the device fuses scale x + amp unit into one rounding, the
            # host export rounds twice; where the two terms nearly cancel
            # that is a few ulps of the terms, not of the sum
            assert np.max(np.abs(got - want)) <= 1e-12 * max(np.max(np.abs(want)), 1e-300)


def test_argument_errors(c1_small):
    earth = c1_small
    op = radial.ViscoelasticOperator(earth, 4)
    with pytest.raises(sl3dpp_amd.InvalidArgument):
        op.evolve(1e-3, st.EXPONENTIAL_TRAPEZOID, 1)  # no load yet
    with pytest.raises(sl3dpp_amd.InvalidArgument):
        op.set_load(np.zeros(7), np.ones(3))  # (lmax+1)(lmax+2) = 30 doubles
    op.set_load(np.zeros(30), np.ones(3))
    with pytest.raises(sl3dpp_amd.InvalidArgument):
        op.evolve(1e-3, st.SDIRK2, 1)
    with pytest.raises(sl3dpp_amd.InvalidArgument):
        op.evolve(1e-3, st.EXPONENTIAL_TRAPEZOID, 3)  # the ramp holds 2 steps
    with pytest.raises(sl3dpp_amd.InvalidArgument):
        op.evolve(0.0, st.EXPONENTIAL_TRAPEZOID, 1)
    op.evolve(1e-3, st.EXPONENTIAL_TRAPEZOID, 2)
    assert op.steps_taken() == 2
    assert not np.any(op.export("memory"))  # a zero load leaves no memory
