"""Pins of the unit tide vector (DegreeOperator::Tide,
src/DegreeOperator.cpp:999-1038), the force of the external potential
psi = (r/a)^p on a degree system, which the oracle takes as an input of the
step problem like the unit load:

 * a homogeneous, nearly incompressible, self-gravitating solid sphere: the
   degree-l tidal Love numbers h = -g U(a) and k = phi(a) of the solved
   system approach the closed forms h = (2l+1)/(2(l-1)) / (1 + f) and
   k = 3/(2(l-1)) / (1 + f), f = (2l^2+4l+3) mu / (l rho g a) (Love's
   homogeneous sphere), with an error of order mu/kappa;
 * on the layered Earth (fluid core, solid mantle), the tide vector is
   minus the assembled operator applied to the potential-only vector
   phi = psi on every U and V row (the fluid-solid interface's point
   coupling included) and on every phi row but the surface one (the fluid
   core's stratification; the Poisson part cancels row by row since psi is
   harmonic and the quadrature is exact for p <= 4), while the surface phi
   row carries the exterior field's (2l+1) a / (4 pi G).

CPU only: the host assembly through the C ABI, no device calls."""

import math

import numpy as np
import pytest

import sl3dpp_amd.step as st

G_SI = 6.6743e-11


def dense(band):
    n, w = band.shape
    kd = w - 1
    a = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - kd), j + 1):
            a[i, j] = a[j, i] = band[j, kd + i - j]
    return a


def homogeneous_sphere(rho, mu, kappa, elements=12, ngll=6):
    s = st.EarthSpec()
    s.layers = 1
    s.r0[0], s.r1[0] = 0.0, 6371e3
    s.rho0[0] = s.rho1[0] = rho
    s.kappa0[0] = s.kappa1[0] = kappa
    s.mu0[0] = s.mu1[0] = mu
    s.elements[0] = elements
    s.ngll = ngll
    s.viscous_layer = 0
    s.viscosity_boundary = 3000e3
    s.viscosity_upper = 1e21
    s.viscosity_lower = 1e21
    return s


@pytest.mark.parametrize("l", [2, 3, 4])
def test_tidal_love_numbers_of_a_homogeneous_sphere(l):
    rho, mu, a = 5500.0, 50e9, 6371e3
    earth = st.Earth(homogeneous_sphere(rho, mu, 1e5 * mu))
    g_nd = earth.export("surface_gravity")[0]
    s = earth.degree_system(l, 0, dt=0.0)  # the elastic (unrelaxed) operator
    x = np.linalg.solve(dense(s["band"]), s["tide"])
    top = s["dofs"][-1]
    h, k = -g_nd * x[top[0]], x[top[2]]
    g = 4.0 / 3.0 * math.pi * G_SI * rho * a
    f = (2 * l * l + 4 * l + 3) * mu / (l * rho * g * a)
    h_closed = (2 * l + 1) / (2 * (l - 1)) / (1 + f)
    k_closed = 3 / (2 * (l - 1)) / (1 + f)
    # the compressible sphere departs by O(mu / kappa) = 1e-5
    assert abs(h - h_closed) < 1e-5 * h_closed
    assert abs(k - k_closed) < 2e-6 * k_closed


def test_degree_zero_tide_vanishes_in_the_incompressible_limit():
    """At l = 0 the potential is (r/a)^2 (the centrifugal mean); an
    incompressible sphere cannot change its radius, so U(a) ~ 1/kappa."""
    rho, mu = 5500.0, 50e9
    u = []
    for ratio in (1e3, 1e4, 1e5):
        earth = st.Earth(homogeneous_sphere(rho, mu, ratio * mu))
        s = earth.degree_system(0, 0, dt=0.0)
        x = np.linalg.solve(dense(s["band"]), s["tide"])
        u.append(x[s["dofs"][-1][0]] * ratio)
    assert u[0] != 0.0
    assert abs(u[2] - u[1]) < 1e-3 * abs(u[1])
    assert abs(u[1] - u[0]) < 1e-2 * abs(u[0])


@pytest.fixture(scope="module")
def layered():
    return st.Earth(st.three_region_earth(core=6, mantle=5, lithosphere=1, ngll=6))


def node_radii(earth):
    elements, ngll = earth.mesh()
    rad = earth.export("radius")
    r = np.zeros(elements * (ngll - 1) + 1)
    for e in range(elements):
        for i in range(ngll):
            r[e * (ngll - 1) + i] = rad[e * ngll + i]
    return r


@pytest.mark.parametrize("l", [2, 3, 4])
def test_tide_is_the_operator_on_the_potential(layered, l):
    earth = layered
    fluid = earth.export("fluid")
    assert fluid.any() and not fluid.all()  # a fluid-solid interface is present
    s = earth.degree_system(l, 0, dt=0.0)
    dofs = s["dofs"]
    r = node_radii(earth)
    a = r[-1]
    psi = np.zeros(s["ndof"])
    for gn in range(len(r)):
        if dofs[gn, 2] >= 0:
            psi[dofs[gn, 2]] = (r[gn] / a) ** l
    y = dense(s["band"]) @ psi
    tide = s["tide"]
    top_phi = dofs[-1, 2]
    rows = np.ones(s["ndof"], dtype=bool)
    rows[top_phi] = False
    scale = np.max(np.abs(tide))
    assert np.max(np.abs(tide[rows] + y[rows])) < 1e-12 * scale
    # the fluid rows are exercised: the core's phi rows carry the force
    elements, ngll = earth.mesh()
    core_nodes = [e * (ngll - 1) + i for e in range(elements) if fluid[e]
                  for i in range(ngll - 1)]
    core_phi = [dofs[gn, 2] for gn in core_nodes if dofs[gn, 2] >= 0]
    assert np.max(np.abs(tide[core_phi])) > 1e-6 * scale
    length, time, mass = earth.scales()
    g_nd = G_SI * mass * time ** 2 / length ** 3
    assert tide[top_phi] == 0.0
    exterior = (2 * l + 1) * a / (4 * math.pi * g_nd)
    assert abs(y[top_phi] - exterior) < 1e-12 * exterior

