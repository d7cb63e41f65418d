"""Shared set-up of the reference's two value gates for the 3D step, used by
the oracle's CPU tests and by the HIP path's GPU tests. Test infrastructure.

 * The zero-lateral pipeline (tests/TestGridCoupling.cpp:178-332): the strain
   of a displacement, synthesised, multiplied by one on the grid, analysed,
   scaled by the memory points' shear modulus and turned into a memory
   force, equals the point-shear part of the assembled degree systems
   (the band with the elastic point shear minus the band without) applied
   to the displacement.
 * The zero-lateral step (tests/TestViscoelastic3D.cpp:90-196): with
   delta = 0 the 3D step reproduces the degree-basis (1D) stepper."""

import ctypes

import numpy as np
from step_problem import DP, OracleStep

import sl3dpp_amd.step as st
from sl3dpp_amd import radial

I64 = ctypes.c_int64


def band_apply(band, x):
    """A x for a symmetric band in the rows-per-column layout of
    degree_system ((ndof, kd + 1): row j holds column j's upper entries
    A(j - kd + r, j) at r), x (ndof,) real or complex."""
    n, w = band.shape
    kd = w - 1
    y = np.zeros_like(x)
    for j in range(n):
        for r in range(w):
            i = j - kd + r
            if i < 0:
                continue
            a = band[j, r]
            y[i] += a * x[j]
            if i != j:
                y[j] += a * x[i]
    return y


def tasks_of(lmax):
    return [(l, 0) for l in range(lmax + 1)] + [(l, 1) for l in range(1, lmax + 1)]


def split(flat, systems, lmax):
    """End-to-end complex columns -> per task (ndof, l+1) complex arrays
    (column-major in the flat form)."""
    out, off = [], 0
    for (l, kind), s in zip(tasks_of(lmax), systems):
        n = 2 * s["ndof"] * (l + 1)
        z = flat[off: off + n].reshape(l + 1, s["ndof"], 2)
        out.append(z[..., 0] + 1j * z[..., 1])
        off += n
    return out


def join(per_task):
    parts = []
    for z in per_task:
        a = np.zeros(z.shape + (2,))
        a[..., 0], a[..., 1] = z.real, z.imag
        parts.append(a.ravel())
    return np.concatenate(parts)


def random_displacement(systems, lmax, seed=7):
    """Random complex solutions on every dof of every task, m = 0 real (a
    real field), the toroidal degree-1 pinned dof zero (the step's
    solutions are zero there)."""
    rng = np.random.default_rng(seed)
    out = []
    for (l, kind), s in zip(tasks_of(lmax), systems):
        z = rng.normal(size=(l + 1, s["ndof"])) + 1j * rng.normal(size=(l + 1, s["ndof"]))
        z[0] = z[0].real
        if s["pinned"] >= 0:
            z[:, s["pinned"]] = 0.0
        out.append(z)
    return out


def point_shear_force(earth, lmax, x, eps=1e-8):
    """The point-shear part of the step's systems applied to x, per task:
    the memory points' unrelaxed modulus mu as point shear on the long-term
    material's systems, formed as K(2 mu) - K(mu) (the point shear is
    linear in its modulus; the long-term systems without any point shear
    are singular where the whole solid relaxes)."""
    length, time, mass = earth.scales()
    mu = earth.export("point_mu") * mass / (length * time ** 2)  # Pa
    out = []
    for (l, kind), z in zip(tasks_of(lmax), x):
        k = radial.TOROIDAL if kind else radial.SPHEROIDAL
        ops = [radial.DegreeOperator(earth, l, kind=k, material=radial.MATERIAL_LONG_TERM,
                                     eps=eps, point_shear=c * mu) for c in (1.0, 2.0)]
        cols = np.ascontiguousarray(np.concatenate([z.real, z.imag]))
        y = ops[1].apply(cols, cols.shape[0]) - ops[0].apply(cols, cols.shape[0])
        n = z.shape[0]
        out.append(y[:n] + 1j * y[n:])
    return out


def scale_by_node(coef, mu, lmax, nodes):
    """GridCoupling::ScaleByNode on a reference-layout tensor: each node's
    coefficients times mu[node]."""
    out = coef.copy()
    off = 0
    sizes = [(lmax + 1) ** 2 - 4, (lmax + 1) ** 2 - 1, (lmax + 1) * (lmax + 2) // 2]
    for size in sizes:
        block = out[off: off + 2 * size * nodes].reshape(nodes, size, 2)
        block *= mu[:, None, None]
        off += 2 * size * nodes
    return out


def oracle_pipeline_force(oracle, ref, x):
    """The oracle's pipeline: strain, synthesis, x 1, analysis, x mu,
    force."""
    coef = np.zeros(ref.spectral_doubles)
    flat = np.ascontiguousarray(join(x))
    oracle.sl3o_strain_of(ctypes.byref(ref.problem), flat.ctypes.data_as(DP),
                          coef.ctypes.data_as(DP))
    G = ref.nodes * ref.ntheta * ref.nphi
    grid = [np.zeros(2 * G), np.zeros(2 * G), np.zeros(G)]
    gp = (DP * 3)(*[g.ctypes.data_as(DP) for g in grid])
    oracle.sl3o_synthesise_tensor(ctypes.byref(ref.problem), coef.ctypes.data_as(DP), gp)
    for g in grid:
        g *= 1.0  # Multiply by the unit weight
    back = np.zeros_like(coef)
    oracle.sl3o_analyse_tensor(ctypes.byref(ref.problem), gp, back.ctypes.data_as(DP))
    weighted = scale_by_node(back, ref.mu_node, ref.lmax, ref.nodes)
    oracle.sl3o_solution_doubles.restype = I64
    forces = np.zeros(oracle.sl3o_solution_doubles(ctypes.byref(ref.problem)))
    oracle.sl3o_memory_force(ctypes.byref(ref.problem), weighted.ctypes.data_as(DP),
                             forces.ctypes.data_as(DP))
    return split(forces, ref.systems, ref.lmax), coef


def forces_gap(got, want, pinned_of):
    scale = max(np.max(np.abs(w)) for w in want)
    worst = 0.0
    for g, w, p in zip(got, want, pinned_of):
        d = np.abs(g - w)
        if p >= 0:
            d[:, p] = 0.0
        worst = max(worst, np.max(d))
    return worst / scale, scale


def general_load(lmax, earth, seed=97):
    """A load with every order growing linearly from zero at t = 0: seeded
    m >= 0 columns of a real field (m = 0 real) times t / (100 yr), of the
    zonal load's size. Zero at t = 0, so the zero state is its own elastic
    start (the step's start strain from rest is then zero on every path).
    Returns the LOAD_FN."""
    _, rate = st.zonal_load_rate(lmax, earth)
    length, time, mass = earth.scales()
    sigma = mass / length ** 2
    rng = np.random.default_rng(seed)
    amp = {}
    for l in range(1, lmax + 1):
        size = rate[l] * 100.0 * st.YEAR / sigma
        cols = size * (rng.normal(size=l + 1) + 1j * rng.normal(size=l + 1))
        cols[0] = cols[0].real
        amp[l] = cols

    unit = 100.0 * st.YEAR / time

    def load(l, t, columns, user):
        if l not in amp:
            return 0
        f = t / unit
        for m in range(l + 1):
            columns[2 * m] = amp[l][m].real * f
            columns[2 * m + 1] = amp[l][m].imag * f
        return 1

    return st.LOAD_FN(load)


class DegreeBasisStepper:
    """The oracle's 1D stepper (sl3o_step_degree_basis) on an OracleStep's
    problem (its bands are the (scheme, dt) flavour's)."""

    class State(ctypes.Structure):
        _fields_ = [("t", ctypes.c_double), ("m", DP), ("d", DP), ("has_strain", ctypes.c_int)]

    def __init__(self, oracle, ref):
        self.oracle, self.ref = oracle, ref
        oracle.sl3o_scalar_doubles.restype = I64
        n = oracle.sl3o_scalar_doubles(ctypes.byref(ref.problem))
        self.m, self.d = np.zeros(n), np.zeros(n)
        self.state = self.State(0.0, self.m.ctypes.data_as(DP), self.d.ctypes.data_as(DP), 0)
        self.nsol = sum(2 * s["ndof"] * (l + 1)
                        for (l, kind), s in zip(tasks_of(ref.lmax), ref.systems) if kind == 0)

    def step(self, dt, scheme, amp, problem=None):
        """One step; problem: another OracleStep of the same Earth and lmax
        (another (scheme, dt) flavour's bands), or None for this one's."""
        ref = problem if problem is not None else self.ref
        sol = np.zeros(self.nsol)
        status = self.oracle.sl3o_step_degree_basis(
            ctypes.byref(ref.problem), ctypes.byref(self.state), ctypes.c_double(dt),
            scheme, amp.ctypes.data_as(DP), sol.ctypes.data_as(DP))
        assert status == 0
        sph = [s for (l, kind), s in zip(tasks_of(self.ref.lmax), self.ref.systems) if kind == 0]
        out, off = [], 0
        for l, s in enumerate(sph):
            n = 2 * s["ndof"] * (l + 1)
            out.append(sol[off: off + n].copy())
            off += n
        return out


def oracle_solutions(oracle, ntasks):
    """The last sl3o_step's solutions of every task (flat complex pairs)."""
    oracle.sl3o_last_solution.restype = I64
    out = []
    for task in range(ntasks):
        n = oracle.sl3o_last_solution(task, None, I64(0))
        assert n > 0
        x = np.zeros(n)
        oracle.sl3o_last_solution(task, x.ctypes.data_as(DP), I64(n))
        out.append(x)
    return out


def zero_lateral_problem(oracle, earth, lmax, scheme, dt, theta=None, wtheta=None):
    return OracleStep(oracle, earth, lmax, scheme, dt, None, theta=theta, wtheta=wtheta)
