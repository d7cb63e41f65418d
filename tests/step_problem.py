"""Builds the oracle's step problem (oracle/step.h) from the host set-up of
sl3dpp_amd (the Earth, its assembled degree bands and dof maps, the memory
points, the grid and the lateral field), so that the oracle and the HIP path
step from identical inputs. Test infrastructure."""

import ctypes
import math
import os

import numpy as np

from conftest import ROOT

D = ctypes.c_double
DP = ctypes.POINTER(ctypes.c_double)
IP = ctypes.POINTER(ctypes.c_int)
I64P = ctypes.POINTER(ctypes.c_int64)


class Problem(ctypes.Structure):
    _fields_ = [
        ("lmax", ctypes.c_int), ("ntheta", ctypes.c_int), ("nphi", ctypes.c_int),
        ("theta", DP), ("wtheta", DP),
        ("nodes", ctypes.c_int), ("np", ctypes.c_int), ("ngll", ctypes.c_int),
        ("elem0", ctypes.c_int), ("elements", ctypes.c_int),
        ("radius", DP), ("jac", DP), ("pweight", DP), ("pvalue", DP), ("pderiv", DP),
        ("mu_node", DP), ("itau_ref_node", DP), ("itau", DP),
        ("ndof", IP), ("kd", IP), ("first_element", IP), ("pinned", IP),
        ("band_off", I64P), ("bands", DP), ("dof_off", I64P), ("dofs", IP),
        ("load_off", I64P), ("loads", DP), ("tide_off", I64P), ("tides", DP),
        ("block_count", ctypes.c_int), ("block_first_node", IP), ("block_lmax", IP),
        ("block_ntheta", IP), ("block_nphi", IP), ("block_theta", ctypes.POINTER(DP)),
        ("block_wtheta", ctypes.POINTER(DP)),
    ]


class Options(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int), ("rtol", D), ("damping", D),
                ("anderson_memory", ctypes.c_int), ("anderson_rate", D)]


class Info(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("residual", D), ("converged", ctypes.c_int),
                ("mixed_passes", ctypes.c_int), ("setup_ms", D), ("start_ms", D),
                ("passes_ms", D), ("finish_ms", D)]


class OracleState(ctypes.Structure):
    _fields_ = [("t", D), ("m", DP * 3), ("strain", DP * 3), ("has_strain", ctypes.c_int),
                ("coef", DP), ("has_coef", ctypes.c_int)]


def small_earth_spec(core=6, mantle=5, lithosphere=1, ngll=6):
    from sl3dpp_amd.step import three_region_earth

    return three_region_earth(core=core, mantle=mantle, lithosphere=lithosphere, ngll=ngll)


def grid_of(lmax, oversampling=1.5):
    nt = math.ceil(oversampling * (lmax + 1) - 1e-9)
    x, w = np.polynomial.legendre.leggauss(nt)
    # descending x: ascending colatitude
    return np.arccos(x[::-1]).copy(), w[::-1].copy(), nt, 2 * nt


class OracleStep:
    """The oracle's problem for one (scheme, dt) on an Earth; keeps every
    array alive while the ctypes structure points at it."""

    def __init__(self, oracle, earth, lmax, scheme, dt, delta, theta=None, wtheta=None,
                 eps=1e-8, blocks=None):
        """delta: the lateral field on every memory node's grid, host array
        [node][theta][phi] (the one the operator was given), or None for
        zero; the oracle forms itau = itau_ref(node) exp(-delta) itself.
        blocks: the operator's node blocks (Viscoelastic3D.blocks) when its
        transforms are cut with depth: each node on its block's grid, the
        grids end to end."""
        self.oracle = oracle
        self.keep = []
        elements, ngll = earth.mesh()
        first, count = (int(v) for v in earth.export("memory_elements"))
        np_ = ngll - 1
        nodes = count * np_
        if theta is None:
            theta, wtheta, nt, nph = grid_of(lmax)
        else:
            nt, nph = len(theta), 2 * len(theta)
        prad = earth.export("point_radius").reshape(elements, np_)
        pmu = earth.export("point_mu").reshape(elements, np_)
        pitau = earth.export("point_itau").reshape(elements, np_)
        self.radius = prad[first: first + count].ravel().copy()
        self.mu_node = pmu[first: first + count].ravel().copy()
        self.itau_ref = pitau[first: first + count].ravel().copy()
        jac = earth.export("jac")[first: first + count].copy()
        tasks = [(l, 0) for l in range(lmax + 1)] + [(l, 1) for l in range(1, lmax + 1)]
        ndof, kd, fe, pinned = [], [], [], []
        bands, band_off, dofs, dof_off, loads, load_off, tides = [], [], [], [], [], [], []
        bo = do = lo = 0
        self.systems = []
        for (l, kind) in tasks:
            s = earth.degree_system(l, kind, scheme, dt, eps)
            self.systems.append(s)
            ndof.append(s["ndof"])
            kd.append(s["kd"])
            fe.append(s["first_element"])
            pinned.append(s["pinned"])
            bands.append(s["band"].ravel())
            band_off.append(bo)
            bo += s["band"].size
            table = []
            for e in range(first, first + count):
                for i in range(ngll):
                    gn = e * (ngll - 1) + i
                    if e < s["first_element"]:
                        table += [-1, -1]
                    elif kind == 0:
                        table += [int(s["dofs"][gn, 0]), int(s["dofs"][gn, 1])]
                    else:
                        table += [int(s["dofs"][gn]), -1]
            dofs.append(np.array(table, dtype=np.int32))
            dof_off.append(do)
            do += len(table)
            if kind == 0:
                loads.append(s["load"])
                tides.append(s["tide"])
                load_off.append(lo)
                lo += s["ndof"]
        if blocks is not None and len(blocks) > 1:
            sizes = [k["ntheta"] * k["nphi"] for k in blocks]
            per_node = np.concatenate([np.full(k["nodes"], n) for k, n in zip(blocks, sizes)])
        else:
            blocks = None
            per_node = np.full(nodes, nt * nph)
        self.arrays = dict(
            theta=np.ascontiguousarray(theta), wtheta=np.ascontiguousarray(wtheta),
            radius=self.radius, jac=jac,
            pweight=earth.export("point_weight"), pvalue=earth.export("point_value"),
            pderiv=earth.export("point_deriv"), mu=self.mu_node, itau_ref=self.itau_ref,
            itau=self._itau(oracle, delta, per_node),
            ndof=np.array(ndof, dtype=np.int32), kd=np.array(kd, dtype=np.int32),
            fe=np.array(fe, dtype=np.int32), pinned=np.array(pinned, dtype=np.int32),
            band_off=np.array(band_off, dtype=np.int64), bands=np.concatenate(bands),
            dof_off=np.array(dof_off, dtype=np.int64), dofs=np.concatenate(dofs),
            load_off=np.array(load_off, dtype=np.int64), loads=np.concatenate(loads),
            tides=np.concatenate(tides),
        )
        a = self.arrays
        p = Problem()
        p.lmax, p.ntheta, p.nphi = lmax, nt, nph
        p.theta, p.wtheta = a["theta"].ctypes.data_as(DP), a["wtheta"].ctypes.data_as(DP)
        p.nodes, p.np, p.ngll, p.elem0, p.elements = nodes, np_, ngll, first, count
        for name, key in [("radius", "radius"), ("jac", "jac"), ("pweight", "pweight"),
                          ("pvalue", "pvalue"), ("pderiv", "pderiv"), ("mu_node", "mu"),
                          ("itau_ref_node", "itau_ref"), ("itau", "itau"),
                          ("bands", "bands"), ("loads", "loads"), ("tides", "tides")]:
            setattr(p, name, a[key].ctypes.data_as(DP))
        for name, key in [("ndof", "ndof"), ("kd", "kd"), ("first_element", "fe"),
                          ("pinned", "pinned"), ("dofs", "dofs")]:
            setattr(p, name, a[key].ctypes.data_as(IP))
        for name, key in [("band_off", "band_off"), ("dof_off", "dof_off"),
                          ("load_off", "load_off"), ("tide_off", "load_off")]:
            setattr(p, name, a[key].ctypes.data_as(I64P))
        p.block_count = 0
        if blocks is not None:
            n = len(blocks)
            for key in ("first_node", "lmax", "ntheta", "nphi"):
                a["block_" + key] = np.array([k[key] for k in blocks], dtype=np.int32)
                setattr(p, "block_" + key, a["block_" + key].ctypes.data_as(IP))
            a["block_theta_arrays"] = [np.ascontiguousarray(k["theta"]) for k in blocks]
            a["block_wtheta_arrays"] = [np.ascontiguousarray(k["wtheta"]) for k in blocks]
            a["block_theta"] = (DP * n)(*[x.ctypes.data_as(DP) for x in a["block_theta_arrays"]])
            a["block_wtheta"] = (DP * n)(*[x.ctypes.data_as(DP)
                                           for x in a["block_wtheta_arrays"]])
            p.block_theta = ctypes.cast(a["block_theta"], ctypes.POINTER(DP))
            p.block_wtheta = ctypes.cast(a["block_wtheta"], ctypes.POINTER(DP))
            p.block_count = n
        self.problem = p
        self.blocks = blocks
        self.lmax, self.nodes, self.ntheta, self.nphi = lmax, nodes, nt, nph
        oracle.sl3o_spectral_doubles.restype = ctypes.c_int64
        self.spectral_doubles = oracle.sl3o_spectral_doubles(lmax, nodes)
        G = int(np.sum(per_node))
        self.grid_values = G
        self.m = [np.zeros(2 * G), np.zeros(2 * G), np.zeros(G)]
        self.strain = [np.zeros(2 * G), np.zeros(2 * G), np.zeros(G)]
        self.coef = np.zeros(self.spectral_doubles)
        self.state = OracleState()
        self.state.t = 0.0
        for b in range(3):
            self.state.m[b] = self.m[b].ctypes.data_as(DP)
            self.state.strain[b] = self.strain[b].ctypes.data_as(DP)
        self.state.coef = self.coef.ctypes.data_as(DP)
        self.state.has_strain = 0
        self.state.has_coef = 0

    def _itau(self, oracle, delta, per_node):
        """itau on every node's grid (per_node values each), node runs of
        equal grids at a time."""
        itau = np.zeros(int(np.sum(per_node)))
        d = None
        if delta is not None:
            d = np.ascontiguousarray(delta, dtype=np.float64).ravel()
            assert d.size == itau.size, "delta holds a value per point of every node's grid"
        node, at = 0, 0
        while node < len(per_node):
            end = node
            while end < len(per_node) and per_node[end] == per_node[node]:
                end += 1
            n, size = end - node, int(per_node[node])
            oracle.sl3o_lateral_itau(
                ctypes.c_int64(n), ctypes.c_int64(size),
                self.itau_ref[node:end].ctypes.data_as(DP),
                d[at:at + n * size].ctypes.data_as(DP) if d is not None else None,
                itau[at:].ctypes.data_as(DP))
            node, at = end, at + n * size
        self.delta = d
        return itau

    def step(self, dt, scheme, amp, max_iterations=40, rtol=1e-11, damping=1.0,
             anderson_memory=5, anderson_rate=0.5, tamp=None):
        o = Options(max_iterations, rtol, damping, anderson_memory, anderson_rate)
        info = Info()
        amp_ptr = amp.ctypes.data_as(DP) if amp is not None else None
        tamp_ptr = tamp.ctypes.data_as(DP) if tamp is not None else None
        status = self.oracle.sl3o_step(ctypes.byref(self.problem), ctypes.byref(self.state),
                                       D(dt), scheme, amp_ptr, tamp_ptr, ctypes.byref(o),
                                       ctypes.byref(info))
        assert status == 0, "oracle factorisation failed"
        return info


def amplitudes(load_fn, lmax, t):
    """A channel's columns at t, complex per (l, m) at l(l+1)/2 + m."""
    amp = np.zeros(2 * (lmax + 1) * (lmax + 2) // 2)
    cols = (D * (2 * (lmax + 1)))()
    for l in range(lmax + 1):
        if load_fn(l, D(t), ctypes.cast(cols, DP), None):
            for m in range(l + 1):
                at = l * (l + 1) // 2 + m
                amp[2 * at] = cols[2 * m]
                amp[2 * at + 1] = cols[2 * m + 1]
    return amp


def oracle_library():
    return ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))


def tidal_rate(lmax, earth, scale=1.0):
    """A tide channel growing linearly from zero at t = 0, so the zero
    state is its elastic start: psi_lm(t) = rate_lm t, with degree 0 and
    degrees 2..lmax, orders 0 and 2 (the centrifugal and semidiurnal
    shapes), about 3 m^2 s^-2 per 30 years at degree 2 falling by e per
    three degrees; degree 1 stays zero (a tide carries none). In the
    non-dimensional units; returns a LOAD_FN for the C ABI."""
    from sl3dpp_amd.step import LOAD_FN, YEAR

    length, time, _ = earth.scales()
    potential = (length / time) ** 2
    rate = {0: (0.4, 0.0, None)}
    for l in range(2, lmax + 1):
        r = 3.0 * math.exp(-(l - 2) / 3.0) / (30.0 * YEAR) * scale
        rate[l] = (r, 0.0, (0.6 * r, -0.3 * r))

    def tide(l, t, columns, user):
        if l not in rate:
            return 0
        for k in range(2 * (l + 1)):
            columns[k] = 0.0
        f = (t * time) / potential
        r0, _, r2 = rate[l]
        columns[0] = r0 * f
        if r2 is not None and l >= 2:
            columns[4] = r2[0] * f
            columns[5] = r2[1] * f
        return 1

    return LOAD_FN(tide)
