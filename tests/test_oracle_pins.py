"""Pins of the oracle's non-band functions against independent checks:
 * step weights and spin algebra: bitwise against the reference's own
   StepWeights.hpp and SpinComponents.hpp compiled into oracle/_ref, and the
   weights against mpmath closed forms (exp, expm1) at 50 digits;
 * Wigner d: sympy's Rotation.d (exact) at small l for N = 0, +-1, +-2 in the
   standard convention; the N = 0 harmonics against scipy's sph_harm_y
   (Condon-Shortley);
 * the spin basis: on the grid, the canonical components of grad Y_lm in the
   Dahlen-Tromp basis e_- = (theta^ - i phi^)/sqrt2 equal
   (zeta/sqrt2) Y^{-1}_lm (the identity behind the strain components of
   include/sl3dpp/SpinComponents.hpp:53-72); derivatives by sympy;
 * the transforms: synthesis -> analysis round trips (1e-12) for N = -2..0
   (tests/TestSpinMachinery.cpp:24-45), the real-field relation;
 * the Anderson least squares: scipy's dgelss (rcond -1) on Gram systems,
   rank-deficient ones included."""

import ctypes
import math

import numpy as np
import pytest
from conftest import dptr

D = ctypes.c_double


@pytest.fixture(scope="module")
def o(oracle):
    oracle.sl3o_phi1.restype = ctypes.c_double
    oracle.sl3o_backward_euler_weight.restype = ctypes.c_double
    return oracle


def test_step_weights_bitwise_equal_reference_build(o, refbuild):
    refbuild.sl3ref_backward_euler_weight.restype = ctypes.c_double
    for h in [0.0, 1e-9, 3e-4, 9.99e-4, 1e-3, 1.01e-3, 0.37, 1.0, 7.5, 50.0, 800.0]:
        a = [D(), D(), D()]
        b = [D(), D(), D()]
        o.sl3o_trapezoid_weights(D(h), *[ctypes.byref(x) for x in a])
        refbuild.sl3ref_trapezoid_weights(D(h), *[ctypes.byref(x) for x in b])
        assert [x.value for x in a] == [x.value for x in b]
        assert o.sl3o_backward_euler_weight(D(h)) == refbuild.sl3ref_backward_euler_weight(D(h))


def test_step_weights_closed_form(o):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 50
    for h in [0.0, 1e-7, 5e-4, 2e-3, 0.5, 3.0, 40.0]:
        e, a, b = D(), D(), D()
        o.sl3o_trapezoid_weights(D(h), ctypes.byref(e), ctypes.byref(a), ctypes.byref(b))
        H = mpmath.mpf(h)
        E = mpmath.exp(-H)
        P = mpmath.mpf(1) if h == 0 else -mpmath.expm1(-H) / H
        assert abs(e.value - float(E)) <= 2e-16 * max(1.0, float(E))
        # the direct quotient (1 - e)/h above the series switch cancels to
        # about eps/h absolute (StepWeights.hpp:98-108)
        tol = 1e-15 + (4e-16 / h if h > 0 else 0.0)
        assert abs(a.value - float(P - E)) <= tol
        assert abs(b.value - float(1 - P)) <= tol
    # h = 0: e = 1, alpha = beta = 0 (StepWeights.hpp:29-31)
    e, a, b = D(), D(), D()
    o.sl3o_trapezoid_weights(D(0.0), ctypes.byref(e), ctypes.byref(a), ctypes.byref(b))
    assert (e.value, a.value, b.value) == (1.0, 0.0, 0.0)


def test_spin_algebra_bitwise_equals_reference_build(o, refbuild):
    rng = np.random.default_rng(5)
    for l in [0, 1, 2, 3, 17, 128]:
        for _ in range(5):
            v = rng.normal(size=12)
            r = rng.uniform(0.5, 1.0)
            a1 = np.zeros(10)
            a2 = np.zeros(10)
            o.sl3o_deviatoric_components(l, dptr(v), D(r), dptr(a1))
            refbuild.sl3ref_deviatoric_components(l, dptr(v), D(r), dptr(a2))
            assert np.array_equal(a1, a2)
            A = rng.normal(size=10)
            r1 = np.zeros(12)
            r2 = np.zeros(12)
            o.sl3o_force_rows(l, dptr(A), D(r), dptr(r1))
            refbuild.sl3ref_force_rows(l, dptr(A), D(r), dptr(r2))
            assert np.array_equal(r1, r2)


def wigner_column(o, lmax, N, m, theta):
    lmin = max(abs(m), abs(N))
    out = np.zeros(lmax - lmin + 1)
    o.sl3o_wigner_d_column(lmax, N, m, D(theta), dptr(out))
    return out


def test_wigner_d_matches_sympy(o):
    sympy = pytest.importorskip("sympy")
    from sympy.physics.quantum.spin import Rotation

    beta = sympy.Symbol("beta")
    theta = 0.83
    for N in [-2, -1, 0, 1, 2]:
        for m in range(-4, 5):
            col = wigner_column(o, 6, N, m, theta)
            lmin = max(abs(m), abs(N))
            for l in range(lmin, 7):
                # sympy's Rotation.d(j, m, mp, beta) = <j m| e^{-i beta J_y} |j mp>,
                # the standard small-d with rows m (first index) and columns mp
                exact = float(Rotation.d(l, m, N, beta).doit().subs(beta, theta).evalf(30))
                assert abs(col[l - lmin] - exact) < 1e-13, (N, m, l)


def test_spin_zero_harmonics_are_condon_shortley(o):
    from scipy.special import sph_harm_y

    theta = 1.1
    for m in range(0, 6):
        col = wigner_column(o, 8, 0, m, theta)
        for l in range(m, 9):
            mine = math.sqrt((2 * l + 1) / (4 * math.pi)) * col[l - m]
            ref = sph_harm_y(l, m, theta, 0.0).real
            assert abs(mine - ref) < 1e-13, (l, m)


def test_gradient_of_a_spin_zero_harmonic_is_the_spin_minus_one_harmonic(o):
    """u = grad_1 Y_lm; its canonical component u^- = -u . e_+ =
    (u_theta + i u_phi)/sqrt2 equals (zeta/sqrt2) Y^{-1}_lm and u^+ =
    -u . e_- = -(u_theta - i u_phi)/sqrt2 equals (zeta/sqrt2) Y^{+1}_lm."""
    for l in range(1, 7):
        zeta = math.sqrt(l * (l + 1))
        for m in range(-l, l + 1):
            for theta in [0.3, 1.2, 2.5]:
                h = 1e-5
                y = lambda t: math.sqrt((2 * l + 1) / (4 * math.pi)) * wigner_column(o, l, 0, m, t)[l - max(abs(m), 0)]  # noqa: E731
                dth = (y(theta + h) - y(theta - h)) / (2 * h)
                # Y = y(theta) e^{i m phi}: u_theta = y' e^{imphi}, u_phi = i m y / sin
                u_th = dth
                u_ph = 1j * m * y(theta) / math.sin(theta)
                uminus = (u_th + 1j * u_ph) / math.sqrt(2)
                uplus = -(u_th - 1j * u_ph) / math.sqrt(2)
                ym1 = math.sqrt((2 * l + 1) / (4 * math.pi)) * wigner_column(o, l, -1, m, theta)[l - max(abs(m), 1)]
                yp1 = math.sqrt((2 * l + 1) / (4 * math.pi)) * wigner_column(o, l, 1, m, theta)[l - max(abs(m), 1)]
                assert abs(uminus - zeta / math.sqrt(2) * ym1) < 1e-8
                assert abs(uplus - zeta / math.sqrt(2) * yp1) < 1e-8


def grid(o, lmax, oversampling=1.5):
    nt = math.ceil(oversampling * (lmax + 1) - 1e-9)
    x = np.zeros(nt)
    w = np.zeros(nt)
    o.sl3o_gauss_legendre(nt, dptr(x), dptr(w))
    return np.arccos(x), w, nt, 2 * nt


@pytest.mark.parametrize("N", [-2, -1, 0])
def test_transform_round_trip(o, N):
    lmax = 12
    theta, w, nt, nph = grid(o, lmax)
    rng = np.random.default_rng(abs(N) + 1)
    real = N == 0
    size = (lmax + 1) * (lmax + 2) // 2 if real else (lmax + 1) ** 2 - N * N
    c = rng.normal(size=(size, 2))
    if real:
        for l in range(lmax + 1):
            c[l * (l + 1) // 2, 1] = 0.0
    c = np.ascontiguousarray(c)
    f = np.zeros(nt * nph * (1 if real else 2))
    o.sl3o_synthesise(lmax, N, int(real), nt, nph, dptr(theta), dptr(c), dptr(f))
    back = np.zeros_like(c)
    o.sl3o_analyse(lmax, N, int(real), nt, nph, dptr(theta), dptr(w), dptr(f), dptr(back))
    assert np.max(np.abs(back - c)) < 1e-12 * np.max(np.abs(c))


def test_gram_least_squares_matches_dgelss(o):
    from scipy.linalg import lapack

    rng = np.random.default_rng(9)
    for m in range(1, 7):
        for deficient in [False, True]:
            F = rng.normal(size=(50, m))
            if deficient and m > 1:
                F[:, -1] = F[:, 0]
            G = F.T @ F
            rhs = F.T @ rng.normal(size=50)
            v, x, s, rank, work, info = lapack.dgelss(G.copy(), rhs.copy(), cond=-1)
            gamma = np.zeros(m)
            g = np.ascontiguousarray(G)
            rk = o.sl3o_gram_least_squares(m, dptr(g), dptr(np.ascontiguousarray(rhs)), dptr(gamma))
            assert rk == rank
            assert np.allclose(gamma, x[:m], rtol=1e-8, atol=1e-10 * np.max(np.abs(x[:m])))
