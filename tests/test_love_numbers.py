"""The gravitating degree systems on a realistic Earth against an external
code, and the static elastic solve of the 1D configuration:

 * the generalised Love numbers of the prem8 deck (fluid outer core, solid
   inner core, twelve layers; not-a-knot cubic joins of the deck's columns)
   for l = 0..32 on the reference's mesh rule (lmax 32, ngll 5) against
   pyslfp's values in the reference's fixture (tests/data/
   pyslfp_love_numbers.txt with tests/data/prem8.elastic.deck, copied to
   tests/golden/), to 1e-9 of each degree's scale as
   tests/TestFixtures.cpp:370-443 checks them: the load's force and
   potential parts, the tangential load and the tide, U, V and phi at the
   surface. The fluid-solid boundary couplings, the fluid's stratification,
   the exterior closure and the tide are pinned here (the traction numbers
   of test_traction_numbers.py hold only solid decks). The axial moments
   (h_c, k_c, m_u, m_phi, m_c) are not compared.
 * the static elastic solve of the 1D configuration (the three-region
   Earth, 48 elements x 6 GLL nodes, l = 2..64, one surface load with
   seeded complex coefficients N(0, 1) (l + 1)^-2 kPa): displacement and
   potential coefficients and the load Love numbers h_l, l_l, k_l on the
   HIP band kernels against the oracle's band solves (1e-12).

CPU tests solve with the oracle's band solvers; -m gpu tests with the HIP
band kernels (sl3_degree_operator_solve)."""

import os

import numpy as np
import pytest
from test_traction_numbers import GOLDEN, oracle_solve

from sl3dpp_amd import radial

LMAX = 32
TOL = 1e-9
COLUMNS = [k + "_" + c for c in radial.LOVE_CHANNELS for k in ("h", "l", "k")]


def read_fixture():
    header, rows = {}, []
    with open(os.path.join(GOLDEN, "pyslfp_love_numbers.txt")) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            w = line.split()
            if w[0] in ("radius", "surface_gravity", "fluid_tidal_2", "axial"):
                header[w[0]] = [float(x) for x in w[1:]]
            elif w[0] != "table":
                rows.append([float(x) for x in w])
    return header, np.array(rows)


def prem8():
    return radial.Earth.from_deck(os.path.join(GOLDEN, "prem8.elastic.deck"), lmax=LMAX, ngll=5,
                                  join="cubic")


def expect_matches(table, rows):
    assert rows.shape == (LMAX + 1, 13)
    for l in range(LMAX + 1):
        assert int(rows[l, 0]) == l
        mine = np.array([table[l][c] for c in COLUMNS])
        ref = rows[l, 1:]
        scale = np.max(np.abs(ref))
        err = np.max(np.abs(mine - ref))
        assert err <= TOL * scale, f"l = {l}: {err / scale:.2e} of the degree's scale"


def test_prem8_model_radius_and_surface_gravity():
    header, _ = read_fixture()
    earth = prem8()
    g_s = earth._keep[2].ravel()[-1]
    assert abs(earth.length - header["radius"][0]) <= TOL * header["radius"][0]
    assert abs(g_s - header["surface_gravity"][0]) <= TOL * header["surface_gravity"][0]


def test_love_numbers_match_pyslfp(oracle):
    _, rows = read_fixture()
    earth = prem8()
    table = []
    for l in range(LMAX + 1):
        op = radial.DegreeOperator(earth, l)
        x = oracle_solve(oracle, op, radial.love_columns(op), 4).reshape(4, op.ndof)
        table.append(radial.love_numbers_from(earth, l, op, x))
    expect_matches(table, rows)


@pytest.mark.gpu
def test_love_numbers_on_the_gpu_match_pyslfp():
    _, rows = read_fixture()
    expect_matches(radial.love_numbers(prem8(), LMAX), rows)


def c0_earth():
    """The 1D configuration's Earth: the three regions (fluid core to 3480
    km, mantle to 6271 km, lithosphere) in 48 elements (26 / 21 / 1, the
    3D configuration's 70 / 56 / 2 split at this count) x 6 GLL nodes."""
    import sl3dpp_amd.step as st

    return st.Earth(st.three_region_earth(26, 21, 1, 6))


def c0_load(lmax, seed=2024):
    """Seeded complex m >= 0 coefficients N(0, 1) (l + 1)^-2 kPa of a real
    field (m = 0 real), degrees 2..lmax."""
    rng = np.random.default_rng(seed)
    out = {}
    for l in range(2, lmax + 1):
        z = (rng.normal(size=l + 1) + 1j * rng.normal(size=l + 1)) * (l + 1.0) ** -2 * 1e3
        z[0] = z[0].real
        out[l] = z
    return out


def c0_columns(earth, l, op, pressure):
    """The load's right-hand sides: a surface pressure p is a surface mass
    p / g_s; [Re | Im] columns of the unit load times the coefficients."""
    length, time, mass = earth.scales()
    g_s = earth.export("surface_gravity")[0] * length / time ** 2
    sigma = pressure / g_s / (mass / length ** 2)  # non-dimensional surface mass
    unit = op.export("load")
    return np.concatenate([np.outer(sigma.real, unit), np.outer(sigma.imag, unit)]).ravel()


C0_LMAX = 64


def c0_dofs(earth, op):
    elements, ngll = earth.mesh()
    top = elements * (ngll - 1)
    return op.dof(top, radial.U), op.dof(top, radial.V), op.dof(top, radial.PHI)


@pytest.mark.gpu
def test_static_elastic_solve_on_the_gpu_matches_the_oracle(oracle):
    import torch

    earth = c0_earth()
    load = c0_load(C0_LMAX)
    worst = 0.0
    for l in range(2, C0_LMAX + 1):
        op = radial.DegreeOperator(earth, l)
        rhs = c0_columns(earth, l, op, load[l])
        ncol = 2 * (l + 1)
        b = torch.tensor(rhs, dtype=torch.float64, device="cuda")
        op.solve(b, ncol)
        got = b.cpu().numpy().reshape(ncol, op.ndof)
        want = oracle_solve(oracle, op, rhs, ncol).reshape(ncol, op.ndof)
        worst = max(worst, np.max(np.abs(got - want)) / np.max(np.abs(want)))
        # the load Love numbers: the unit load's surface U, V, phi
        unit = torch.tensor(op.export("load"), dtype=torch.float64, device="cuda")
        op.solve(unit, 1)
        iu, iv, ip = c0_dofs(earth, op)
        h, ll, k = (unit.cpu().numpy()[i] for i in (iu, iv, ip))
        ref = oracle_solve(oracle, op, op.export("load"), 1)
        # one scale per degree (the numbers' largest), as the reference's
        # Love-number gates compare them
        scale = max(abs(ref[iu]), abs(ref[iv]), abs(ref[ip]))
        for mine, i in ((h, iu), (ll, iv), (k, ip)):
            assert abs(mine - ref[i]) <= 1e-12 * scale
        # the surface displacement coefficients are the load's (as surface
        # mass) times h
        cols = rhs.reshape(ncol, -1)[:, iu] / op.export("load")[iu]
        sigma = cols[: l + 1] + 1j * cols[l + 1:]
        surface = got[: l + 1, iu] + 1j * got[l + 1:, iu]
        assert np.max(np.abs(surface - sigma * h)) <= 1e-12 * np.max(np.abs(surface))
    print(f"C0 static solve l = 2..{C0_LMAX}: HIP vs oracle {worst:.2e}")
    assert worst < 1e-12
