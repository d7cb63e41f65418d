"""The degree systems' assembly against an external code: the traction numbers
of two solid decks, gravitating and not, against pyslfp's values in the
reference's fixture (tests/data/e1_traction_numbers.txt, copied to
tests/golden/ with its two decks), to 1e-9 of each degree's scale as
tests/TestTractionNumbers.cpp:91-131 checks them; the degree-0 closed form
(:133-150) and reciprocity h_v = l(l+1) l_u.

The systems are assembled by the library (sl3_degree_operator_*) on the
reference's mesh rule (drmax = 0.1 a / (lmax + 1), ngll 5, lmax 8) and
solved by the oracle's band solvers on the CPU, and by the HIP band kernels
on the GPU (the -m gpu tests)."""

import math
import os

import numpy as np
import pytest
from conftest import dptr, iptr

from sl3dpp_amd import radial

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LMAX = 8
TOL = 1e-9


def read_fixture():
    """{(model, gravitating): {"radius": a, "rows": (9 x 4) h_u l_u h_v l_v}}."""
    out, current = {}, None
    with open(os.path.join(GOLDEN, "e1_traction_numbers.txt")) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            words = line.split()
            if words[0] == "model":
                current = {"radius": float(words[5]), "rows": []}
                out[(words[1], int(words[3]))] = current
            elif words[0] != "table":
                current["rows"].append([float(w) for w in words[1:5]])
    for v in out.values():
        v["rows"] = np.array(v["rows"])
    return out


def earth_of(name):
    return radial.Earth.from_deck(os.path.join(GOLDEN, f"e1_{name}.deck"), lmax=LMAX, ngll=5)


def oracle_solve(oracle, op, rhs, nrhs):
    """DegreeOperator::Solve with the oracle's band solvers: the symmetric
    scaling, Cholesky of the scaled band (LU on the full band if that
    fails), unscaling (src/DegreeOperator.cpp:775-901)."""
    n, kd = op.ndof, op.kd
    band = op.export("band").reshape(n, kd + 1)
    s = op.export("scale")
    x = (rhs.reshape(nrhs, n) * s).copy()
    scaled = band.copy()
    for j in range(n):
        for r in range(kd + 1):
            i = j - kd + r
            if i >= 0:
                scaled[j, r] *= s[i] * s[j]
    ab = np.ascontiguousarray(scaled)
    if oracle.sl3o_band_cholesky_factor(dptr(ab), n, kd) == 0:
        oracle.sl3o_band_cholesky_solve(dptr(ab), n, kd, dptr(x), nrhs)
    else:
        lu = np.zeros((n, 3 * kd + 1))
        for j in range(n):
            for i in range(max(0, j - kd), min(n, j + kd + 1)):
                lo, hi = min(i, j), max(i, j)
                lu[j, 2 * kd + i - j] = scaled[hi, kd + lo - hi]
        piv = np.zeros(n, dtype=np.int32)
        assert oracle.sl3o_band_lu_factor(dptr(lu), n, kd, kd, iptr(piv)) == 0
        oracle.sl3o_band_lu_solve(dptr(lu), n, kd, kd, iptr(piv), dptr(x), nrhs)
    return (x * s).reshape(-1)


def host_traction_numbers(oracle, earth, gravitating):
    h_factor = earth.length ** 2 * earth.time ** 2 / earth.mass
    out = {k: np.full(LMAX + 1, np.nan) for k in ("h_u", "l_u", "h_v", "l_v")}
    top = (len(earth.edges) - 1) * (earth.ngll - 1)
    for l in range(LMAX + 1):
        if l == 1 and not gravitating:
            continue
        op = radial.DegreeOperator(earth, l, gravitating=gravitating)
        rhs = np.concatenate([op.export("traction"), op.export("traction_tangential")])
        x = oracle_solve(oracle, op, rhs, 2).reshape(2, op.ndof)
        iu, iv = op.dof(top, radial.U), op.dof(top, radial.V)
        val = lambda d, c: x[c, d] if d >= 0 else 0.0  # noqa: E731
        out["h_u"][l] = val(iu, 0) * h_factor
        out["l_u"][l] = val(iv, 0) * h_factor
        out["h_v"][l] = val(iu, 1) * h_factor
        out["l_v"][l] = val(iv, 1) * h_factor
    return out


def expect_matches(ours, theirs):
    rows = theirs["rows"]
    assert rows.shape == (LMAX + 1, 4)
    for l in range(LMAX + 1):
        mine = np.array([ours["h_u"][l], ours["l_u"][l], ours["h_v"][l], ours["l_v"][l]])
        finite = ~np.isnan(rows[l])
        scale = np.max(np.abs(rows[l][finite])) if finite.any() else 0.0
        assert np.array_equal(np.isnan(mine), ~finite), f"l = {l}: excluded degrees differ"
        err = np.max(np.abs(mine[finite] - rows[l][finite])) if finite.any() else 0.0
        assert err <= TOL * scale, f"l = {l}: {err:.3e} > {TOL} x {scale:.3e}"


def test_mesh_follows_the_reference_rule():
    # drmax = 0.1 a / 9: 90 elements in one span; 71 + 20 with the 5000 km
    # boundary pinned
    assert len(earth_of("homogeneous").edges) - 1 == 90
    assert len(earth_of("two_solid").edges) - 1 == 71 + 20


@pytest.mark.parametrize("name", ["homogeneous", "two_solid"])
@pytest.mark.parametrize("gravitating", [0, 1])
def test_traction_numbers_match_pyslfp(oracle, name, gravitating):
    fixture = read_fixture()
    earth = earth_of(name)
    theirs = fixture[(name, gravitating)]
    assert abs(earth.length - theirs["radius"]) <= TOL * theirs["radius"]
    expect_matches(host_traction_numbers(oracle, earth, bool(gravitating)), theirs)


def test_degree_zero_is_uniform_compression(oracle):
    # a unit pressure compresses the homogeneous sphere uniformly:
    # U(a) = -a / (3 kappa), kappa = rho (vp^2 - 4 vs^2 / 3)
    tn = host_traction_numbers(oracle, earth_of("homogeneous"), False)
    rho, vp, vs, a = 5500.0, 10000.0, 5500.0, 6371.0e3
    kappa = rho * (vp * vp - 4.0 * vs * vs / 3.0)
    assert abs(tn["h_u"][0] - (-a / (3.0 * kappa))) <= 1e-11 * a / (3.0 * kappa)


@pytest.mark.parametrize("gravitating", [False, True])
def test_reciprocity(oracle, gravitating):
    # the systems are symmetric: the tangential traction's U equals
    # l(l + 1) times the normal traction's V (Betti)
    tn = host_traction_numbers(oracle, earth_of("two_solid"), gravitating)
    for l in range(1 if gravitating else 2, LMAX + 1):
        k2 = l * (l + 1.0)
        assert abs(tn["h_v"][l] - k2 * tn["l_u"][l]) <= 1e-10 * abs(tn["h_v"][l])


def test_point_shear_at_the_nodes_is_the_raised_elastic_shear():
    """The point shear's quadratic form s (l(l+1)(l(l+1) - 2) M^2 + 4/3 R^2
    + l(l+1) S^2), taken at the GLL nodes with their weights, is exactly
    the elastic block with the shear modulus raised by s (A, C += 4/3 s,
    F -= 2/3 s, L, N += s): both assemblies of the same energy."""
    base = earth_of("two_solid")
    E, n = len(base.edges) - 1, base.ngll
    rho, drho, g, A, C, F, L, N = (x.reshape(E, n) for x in base._keep)
    rng = np.random.default_rng(5)
    s = rng.uniform(1e9, 3e10, size=(E, n))
    s[:, 0] = s[np.arange(E) - 1, n - 1]  # continuous at shared nodes
    s[0, 0] = s[0, 1]
    raised = radial.Earth(base.edges, n, rho, drho, g, A + 4 * s / 3, C + 4 * s / 3,
                          F - 2 * s / 3, L + s, N + s, base._fluid, length=base.length)
    for l in (0, 2, 5):
        for gravitating in (False, True):
            a = radial.DegreeOperator(base, l, gravitating=gravitating, point_shear=s,
                                      point_set=radial.POINTS_GLL)
            b = radial.DegreeOperator(raised, l, gravitating=gravitating)
            ba, bb = a.export("band"), b.export("band")
            assert np.max(np.abs(ba - bb)) <= 1e-12 * np.max(np.abs(bb))
        for_t = radial.DegreeOperator(base, l or 1, kind=radial.TOROIDAL, point_shear=s,
                                      point_set=radial.POINTS_GLL)
        rt = radial.DegreeOperator(raised, l or 1, kind=radial.TOROIDAL)
        ta, tb = for_t.export("band"), rt.export("band")
        assert np.max(np.abs(ta - tb)) <= 1e-12 * np.max(np.abs(tb))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["homogeneous", "two_solid"])
@pytest.mark.parametrize("gravitating", [0, 1])
def test_traction_numbers_on_the_gpu_match_pyslfp(name, gravitating):
    fixture = read_fixture()
    tn = radial.traction_numbers(earth_of(name), LMAX, bool(gravitating))
    expect_matches(tn, fixture[(name, gravitating)])
