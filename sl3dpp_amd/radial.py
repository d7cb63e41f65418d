"""Python mirror of the reference's radial set-up over the C ABI: an Earth from
an elastic deck (EarthModel::FromDeck, RadialMesh::Create, Material::Create,
src/EarthModel.cpp:86-200, src/RadialMesh.cpp:13-200), one degree's radial
system (DegreeOperator / ToroidalOperator, include/sl3dpp/DegreeOperator.hpp:
94-384) factorised and solved on the GPU, the batched radially symmetric
Maxwell evolution (ViscoelasticOperator, include/sl3dpp/Viscoelastic.hpp:
163-301), and the static responses built on the degree systems: the traction numbers (ComputeTractionNumbers, src/LoveNumbers.cpp:
463-512) and the load Love numbers of a static elastic solve.

Deck values between rows are joined linearly within a layer; the reference
joins them by planetmodel's splines, which agree with that on piecewise-
constant and piecewise-linear decks (the only ones used here)."""

import ctypes
import math

import numpy as np

from . import check, lib
from .step import Earth as _StepEarth

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_int64_p = ctypes.POINTER(ctypes.c_int64)

G_SI = 6.6743e-11
SPHEROIDAL, TOROIDAL = 0, 1
MATERIAL_ELASTIC, MATERIAL_LONG_TERM = 0, 1
POINTS_GAUSS, POINTS_GLL = 0, 1
U, V, PHI = 0, 1, 2


class NodalSpec(ctypes.Structure):
    _fields_ = [
        ("elements", ctypes.c_int),
        ("ngll", ctypes.c_int),
        ("edges", c_double_p),
        ("rho", c_double_p),
        ("drho", c_double_p),
        ("gravity", c_double_p),
        ("A", c_double_p),
        ("C", c_double_p),
        ("F", c_double_p),
        ("L", c_double_p),
        ("N", c_double_p),
        ("fluid", c_int_p),
        ("branches", ctypes.c_int),
        ("branch_mu", c_double_p),
        ("branch_itau", c_double_p),
        ("length", ctypes.c_double),
        ("density", ctypes.c_double),
    ]


class DegreeOptions(ctypes.Structure):
    _fields_ = [
        ("eps", ctypes.c_double),
        ("gravitating", ctypes.c_int),
        ("material", ctypes.c_int),
        ("point_shear", c_double_p),
        ("point_set", ctypes.c_int),
    ]


lib.sl3_degree_options_default.restype = None


def _dp(a):
    return a.ctypes.data_as(c_double_p)


# ---- decks ------------------------------------------------------------------

def read_deck(path):
    """An elastic deck's rows (radius m, rho, vpv, vsv, vph, vsh, eta), blank
    lines and '#' comments skipped, split into layers where a radius
    repeats (a discontinuity). Returns a list of (rows x 7) arrays."""
    rows = []
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if line:
                rows.append([float(x) for x in line.split()])
    data = np.asarray(rows, dtype=np.float64)
    if data.ndim != 2 or data.shape[1] != 7:
        raise ValueError(f"{path}: an elastic deck has 7 columns")
    layers, start = [], 0
    for i in range(1, len(data)):
        if data[i, 0] == data[i - 1, 0]:
            layers.append(data[start:i])
            start = i
    layers.append(data[start:])
    return layers


class DeckModel:
    """EarthModel::FromDeck for an elastic deck: per layer the columns joined
    between the deck's rows (join "linear", or "cubic": not-a-knot cubic
    splines, planetmodel's join of a deck's columns, a parabola through
    three rows and a line through two), moduli A = rho vph^2,
    C = rho vpv^2, L = rho vsv^2, N = rho vsh^2, F = eta (A - 2L), gravity
    from the density's mass integral taken on the join's segments (SI)."""

    def __init__(self, layers, join="linear"):
        if join not in ("linear", "cubic"):
            raise ValueError("join: linear or cubic")
        self.layers = layers
        self.join = join
        self.boundaries = [layers[0][0, 0]] + [ly[-1, 0] for ly in layers]
        self.radius = self.boundaries[-1]
        self.splines = []
        if join == "cubic":
            from scipy.interpolate import CubicSpline

            for ly in layers:
                if len(ly) < 2:
                    self.splines.append(None)
                    continue
                self.splines.append(CubicSpline(ly[:, 0], ly[:, 1:], axis=0,
                                                bc_type="not-a-knot"))
        # mass below each layer's bottom
        self.mass_below = [0.0]
        for i, ly in enumerate(layers):
            self.mass_below.append(self.mass_below[-1] + self._mass(i, ly[0, 0], ly[-1, 0]))

    def _rho(self, layer, r):
        return self.values(layer, r)[0]

    def _mass(self, layer, r0, r1):
        """4 pi int_r0^r1 rho r^2 dr over one layer: per segment between rows
        by 4-point Gauss-Legendre, exact for a density up to cubic."""
        ly = self.layers[layer]
        x, w = np.polynomial.legendre.leggauss(4)
        total = 0.0
        for k in range(max(len(ly) - 1, 1)):
            lo = ly[k, 0]
            hi = ly[k + 1, 0] if len(ly) > 1 else ly[k, 0]
            a, b = max(lo, r0), min(hi, r1)
            if b <= a:
                continue
            for xi, wi in zip(x, w):
                r = 0.5 * (a + b) + 0.5 * (b - a) * xi
                total += 0.5 * (b - a) * wi * self._rho(layer, r) * r * r
        return 4.0 * math.pi * total

    def layer_of(self, r, above):
        """The layer holding r; at a boundary the one above or below."""
        for i in range(len(self.layers)):
            lo, hi = self.boundaries[i], self.boundaries[i + 1]
            if (lo <= r < hi) if above else (lo < r <= hi):
                return i
        return 0 if r <= self.boundaries[0] else len(self.layers) - 1

    def values(self, layer, r):
        """(rho, drho/dr, vpv, vsv, vph, vsh, eta) in `layer` at r."""
        ly = self.layers[layer]
        if self.join == "cubic" and self.splines[layer] is not None:
            sp = self.splines[layer]
            row = sp(r)
            drho = float(sp(r, 1)[0])
            return (float(row[0]), drho, *[float(v) for v in row[1:]])
        k = int(np.clip(np.searchsorted(ly[:, 0], r, side="right") - 1, 0, len(ly) - 2)) \
            if len(ly) > 1 else 0
        if len(ly) == 1:
            row = ly[0]
            return (row[1], 0.0, *row[2:])
        x0, x1 = ly[k, 0], ly[k + 1, 0]
        t = (r - x0) / (x1 - x0)
        row = ly[k] + t * (ly[k + 1] - ly[k])
        drho = (ly[k + 1, 1] - ly[k, 1]) / (x1 - x0)
        return (row[1], drho, *row[2:])

    def gravity(self, layer, r):
        if r <= 0.0:
            return 0.0
        ly = self.layers[layer]
        m = self.mass_below[layer] + self._mass(layer, ly[0, 0], r)
        return G_SI * m / (r * r)

    def is_fluid(self, layer):
        ly = self.layers[layer]
        return bool(np.all(ly[:, 3] == 0.0) and np.all(ly[:, 5] == 0.0))


def read_legacy_deck(path):
    """Deck::ReadLegacySl3d (src/Deck.cpp:103-170): the legacy gia1d/sl3d
    model (a title line; `nknot ntau`, ntau = 1; rows `r rho kappa[GPa]
    mu[GPa] log10(viscosity)`, 0 meaning no viscosity), split into layers
    where a radius repeats. Returns (title, layers): per layer a (rows x 5)
    array r (m), rho, kappa (Pa), mu (Pa), viscosity (Pa s, NaN where
    absent)."""
    with open(path) as f:
        title = f.readline().strip()
        head = f.readline().split()
        nknot, ntau = int(head[0]), int(head[1])
        if ntau != 1:
            raise ValueError(f"{path}: only ntau = 1 legacy files are supported")
        values = [float(x) for x in f.read().split()]
    data = np.asarray(values[: 5 * nknot], dtype=np.float64).reshape(nknot, 5)
    data[:, 2:4] *= 1e9
    with np.errstate(over="ignore"):
        data[:, 4] = np.where(data[:, 4] == 0.0, np.nan, 10.0 ** data[:, 4])
    layers, start = [], 0
    for i in range(1, nknot):
        if data[i, 0] == data[i - 1, 0]:
            layers.append(data[start:i])
            start = i
    layers.append(data[start:])
    for ly in layers:
        absent = np.isnan(ly[:, 4])
        if absent.any() and not absent.all():
            raise ValueError(f"{path}: a layer mixes zero and non-zero viscosity")
    return title, layers


class LegacyDeckModel(DeckModel):
    """EarthModel::FromDeck with Options::spline = Natural on a legacy deck
    (src/EarthModel.cpp:86-260): rho, kappa and mu joined per layer by
    natural cubic splines (a line through two rows), the viscosity in its
    logarithm by straight lines (planetmodel's `log-linear` kind for a
    viscosity column); isotropic moduli A = C = kappa + 4 mu / 3, F = kappa
    - 2 mu / 3, L = N = mu (NodalModuli::Isotropic, src/Material.cpp:9-24);
    gravity from the density's mass integral on the spline's segments."""

    def __init__(self, layers):
        from scipy.interpolate import CubicSpline

        self.layers = layers
        self.join = "natural"
        self.boundaries = [layers[0][0, 0]] + [ly[-1, 0] for ly in layers]
        self.radius = self.boundaries[-1]
        self.splines = [CubicSpline(ly[:, 0], ly[:, 1:4], axis=0, bc_type="natural")
                        if len(ly) > 1 else None for ly in layers]
        self.mass_below = [0.0]
        for i, ly in enumerate(layers):
            self.mass_below.append(self.mass_below[-1] + self._mass(i, ly[0, 0], ly[-1, 0]))

    def _rho(self, layer, r):
        return self.isotropic(layer, r)[0]

    def isotropic(self, layer, r):
        """(rho, drho/dr, kappa, mu) in `layer` at r."""
        sp = self.splines[layer]
        if sp is None:
            row = self.layers[layer][0]
            return row[1], 0.0, row[2], row[3]
        v = sp(r)
        return float(v[0]), float(sp(r, 1)[0]), float(v[1]), float(v[2])

    def viscosity(self, layer, r):
        """The viscosity (Pa s) at r, None where the layer has none: the
        logarithm joined by straight lines between the rows."""
        ly = self.layers[layer]
        if np.isnan(ly[0, 4]):
            return None
        if len(ly) == 1:
            return float(ly[0, 4])
        return float(np.exp(np.interp(r, ly[:, 0], np.log(ly[:, 4]))))

    def is_fluid(self, layer):
        return bool(np.all(self.layers[layer][:, 3] == 0.0))


def graded_mesh_edges(model, lmax, eps=1e-8):
    """RadialMesh::CreateGraded (src/RadialMesh.cpp:59-111): from the
    surface down, elements of width max(0.1 a / (lmax + 1), 0.1 a
    ln(1 - d/a) / ln(eps)) at depth d (non-dimensional radii, the surface
    one), every model boundary an element boundary, the last one or two
    elements above a boundary spanning what is left (split in two where it
    exceeds the width). Returns the edges in metres and each element's
    layer."""
    a_si = model.radius
    b = [x / a_si for x in model.boundaries]
    a = b[-1]
    floor = 0.1 * a / (lmax + 1)
    log_eps = math.log(eps)

    def width(r):
        depth = a - r
        if depth <= 0.0:
            return floor
        return max(floor, 0.1 * a * math.log1p(-depth / a) / log_eps)

    edges = [a]
    for s in range(len(b) - 2, -1, -1):
        lo = b[s]
        r = edges[-1]
        while True:
            w = width(r)
            if r - lo <= 2.0 * w:
                n = 1 if r - lo <= w else 2
                for k in range(1, n + 1):
                    edges.append(r + (lo - r) * (k / n))
                edges[-1] = lo
                break
            r -= w
            edges.append(r)
    edges.reverse()
    layer = [model.layer_of(0.5 * (edges[e] + edges[e + 1]) * a_si, True)
             for e in range(len(edges) - 1)]
    return [x * a_si for x in edges], layer


def mesh_edges(model, lmax=None, drmax=None):
    """RadialMesh::Create(model, {lmax} or {drmax}): every model boundary an
    element boundary, each span cut uniformly into elements no wider than
    drmax = 0.1 a / (lmax + 1), in the model's non-dimensional radii (the
    surface radius one), returned in metres with each element's layer."""
    if (lmax is None) == (drmax is None):
        raise ValueError("give exactly one of lmax and drmax")
    a = model.radius
    b = [x / a for x in model.boundaries]
    step = 0.1 * b[-1] / (lmax + 1) if lmax is not None else drmax / a
    edges, layer = [b[0]], []
    for s in range(len(b) - 1):
        lo, hi = b[s], b[s + 1]
        nel = max(1, math.ceil((hi - lo) / step))
        for e in range(1, nel + 1):
            t = e / nel
            edges.append(hi if e == nel else (1.0 - t) * lo + t * hi)
            layer.append(s)
    return [x * a for x in edges], layer


def gll_nodes(n):
    """GLL nodes on [-1, 1], ascending (numpy's Legendre derivative roots)."""
    N = n - 1
    inner = np.polynomial.legendre.Legendre.basis(N).deriv().roots()
    return np.concatenate([[-1.0], np.sort(inner.real), [1.0]])


class Earth(_StepEarth):
    """An Earth handle from nodal arrays (sl3_earth_create_nodal), with the
    layered Earth's host queries (scales, mesh, export, degree_system)."""

    def __init__(self, edges, ngll, rho, drho, gravity, A, C, F, L, N, fluid,
                 branch_mu=None, branch_itau=None, length=0.0, density=0.0):
        self.ngll = ngll
        self.edges = np.ascontiguousarray(edges, dtype=np.float64)
        self._keep = [np.ascontiguousarray(x, dtype=np.float64)
                      for x in (rho, drho, gravity, A, C, F, L, N)]
        self._fluid = np.ascontiguousarray(fluid, dtype=np.int32)
        spec = NodalSpec()
        spec.elements = len(self.edges) - 1
        spec.ngll = ngll
        spec.edges = _dp(self.edges)
        (spec.rho, spec.drho, spec.gravity, spec.A, spec.C, spec.F, spec.L,
         spec.N) = [_dp(x) for x in self._keep]
        spec.fluid = self._fluid.ctypes.data_as(c_int_p)
        spec.branches = 0
        if branch_mu is not None:
            self._branch = [np.ascontiguousarray(branch_mu, dtype=np.float64),
                            np.ascontiguousarray(branch_itau, dtype=np.float64)]
            spec.branches = 1
            spec.branch_mu, spec.branch_itau = [_dp(x) for x in self._branch]
        spec.length = length
        spec.density = density
        h = ctypes.c_void_p()
        check(lib.sl3_earth_create_nodal(ctypes.byref(spec), ctypes.byref(h)))
        self.handle = h
        t, L_, m = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib.sl3_earth_scales(self.handle, ctypes.byref(L_), ctypes.byref(t), ctypes.byref(m)))
        self.length, self.time, self.mass = L_.value, t.value, m.value

    @classmethod
    def from_deck(cls, path, lmax=None, drmax=None, ngll=5, join="linear", viscosity=None):
        """EarthModel::FromDeck + RadialMesh::Create + Material::Create, and
        with viscosity(layer, model) -> eta (Pa s) or None per layer,
        Rheology::Maxwell's one branch: mu the Voigt shear, itau = mu / eta
        (include/sl3dpp/Rheology.hpp:100-190)."""
        model = DeckModel(read_deck(path), join=join)
        edges, layer = mesh_edges(model, lmax=lmax, drmax=drmax)
        xi = gll_nodes(ngll)
        E = len(edges) - 1
        arrays = {k: np.zeros((E, ngll)) for k in ("rho", "drho", "g", "A", "C", "F", "L", "N")}
        fluid = np.zeros(E, dtype=np.int32)
        for e in range(E):
            k = layer[e]
            fluid[e] = model.is_fluid(k)
            lo, hi = edges[e], edges[e + 1]
            for i in range(ngll):
                r = lo + (xi[i] + 1.0) * 0.5 * (hi - lo)
                if i == 0:
                    r = lo
                if i == ngll - 1:
                    r = hi
                rho, drho, vpv, vsv, vph, vsh, eta = model.values(k, r)
                A, C = rho * vph ** 2, rho * vpv ** 2
                L, N = rho * vsv ** 2, rho * vsh ** 2
                arrays["rho"][e, i], arrays["drho"][e, i] = rho, drho
                arrays["g"][e, i] = model.gravity(k, r)
                arrays["A"][e, i], arrays["C"][e, i] = A, C
                arrays["L"][e, i], arrays["N"][e, i] = L, N
                arrays["F"][e, i] = eta * (A - 2.0 * L)
        branch_mu = branch_itau = None
        if viscosity is not None:
            mu_v = (arrays["C"] + arrays["A"] - 2.0 * arrays["F"] + 5.0 * arrays["N"]
                    + 6.0 * arrays["L"]) / 15.0
            branch_mu = np.zeros((E, ngll))
            branch_itau = np.zeros((E, ngll))
            for e in range(E):
                eta = viscosity(layer[e], model)
                if eta is not None and not fluid[e]:
                    branch_mu[e] = mu_v[e]
                    branch_itau[e] = mu_v[e] / eta
        earth = cls(edges, ngll, arrays["rho"], arrays["drho"], arrays["g"], arrays["A"],
                    arrays["C"], arrays["F"], arrays["L"], arrays["N"], fluid,
                    branch_mu=branch_mu, branch_itau=branch_itau, length=model.radius)
        earth.model = model
        earth.layer = layer
        earth.branch_itau = branch_itau
        return earth


def legacy_earth(path, lmax, ngll=5, eps=1e-8):
    """The legacy sl3d model as the Fortran gate builds it
    (benchmarks/fortran_gate.cpp:52-57): Deck::ReadLegacySl3d,
    EarthModel::FromDeck with natural splines, RadialMesh::CreateGraded
    (lmax, ngll 5), Material::Create and Rheology::FromModel (one Maxwell
    branch where a layer holds a viscosity: mu the shear modulus, itau =
    mu / eta at the nodes; src/Rheology.cpp:129-163, 288-326)."""
    _, layers = read_legacy_deck(path)
    model = LegacyDeckModel(layers)
    edges, layer = graded_mesh_edges(model, lmax, eps)
    xi = gll_nodes(ngll)
    E = len(edges) - 1
    arrays = {k: np.zeros((E, ngll)) for k in ("rho", "drho", "g", "A", "C", "F", "L", "N")}
    branch_mu, branch_itau = np.zeros((E, ngll)), np.zeros((E, ngll))
    fluid = np.zeros(E, dtype=np.int32)
    for e in range(E):
        k = layer[e]
        fluid[e] = model.is_fluid(k)
        lo, hi = edges[e], edges[e + 1]
        for i in range(ngll):
            r = lo if i == 0 else hi if i == ngll - 1 else lo + (xi[i] + 1.0) * 0.5 * (hi - lo)
            rho, drho, kappa, mu = model.isotropic(k, r)
            arrays["rho"][e, i], arrays["drho"][e, i] = rho, drho
            arrays["g"][e, i] = model.gravity(k, r)
            arrays["A"][e, i] = arrays["C"][e, i] = kappa + 4.0 * mu / 3.0
            arrays["F"][e, i] = kappa - 2.0 * mu / 3.0
            arrays["L"][e, i] = arrays["N"][e, i] = mu
            eta = model.viscosity(k, r)
            if eta is not None and not fluid[e]:
                branch_mu[e, i] = mu
                branch_itau[e, i] = mu / eta
    viscous = bool(np.any(branch_itau > 0.0))
    earth = Earth(edges, ngll, arrays["rho"], arrays["drho"], arrays["g"], arrays["A"],
                  arrays["C"], arrays["F"], arrays["L"], arrays["N"], fluid,
                  branch_mu=branch_mu if viscous else None,
                  branch_itau=branch_itau if viscous else None, length=model.radius)
    earth.model = model
    earth.layer = layer
    earth.branch_itau = branch_itau
    return earth


class DegreeOperator:
    """One degree's radial system (sl3_degree_operator_*): DegreeOperator or
    ToroidalOperator of the reference, solved on the GPU."""

    def __init__(self, earth, l, kind=SPHEROIDAL, gravitating=True, material=MATERIAL_ELASTIC,
                 eps=1e-8, point_shear=None, point_set=POINTS_GAUSS):
        self.earth = earth
        self.l = l
        self.kind = kind
        o = DegreeOptions()
        lib.sl3_degree_options_default(ctypes.byref(o))
        o.eps = eps
        o.gravitating = int(gravitating)
        o.material = material
        o.point_set = point_set
        if point_shear is not None:
            self._shear = np.ascontiguousarray(point_shear, dtype=np.float64)
            o.point_shear = _dp(self._shear)
        h = ctypes.c_void_p()
        check(lib.sl3_degree_operator_create(earth.handle, l, kind, ctypes.byref(o),
                                             ctypes.byref(h)))
        self.handle = h
        n, kd, fe, pin = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib.sl3_degree_operator_shape(self.handle, ctypes.byref(n), ctypes.byref(kd),
                                            ctypes.byref(fe), ctypes.byref(pin)))
        self.ndof, self.kd, self.first_element, self.pinned = n.value, kd.value, fe.value, pin.value

    def __del__(self, _lib=lib):
        if getattr(self, "handle", None):
            _lib.sl3_degree_operator_destroy(self.handle)
            self.handle = None

    def export(self, name):
        count = ctypes.c_int64()
        check(lib.sl3_degree_operator_export(self.handle, name.encode(), None, ctypes.byref(count)))
        out = np.empty(count.value)
        check(lib.sl3_degree_operator_export(self.handle, name.encode(), _dp(out),
                                             ctypes.byref(count)))
        return out

    def dof(self, node, component=0):
        dofs = self.export("dofs")
        per = 3 if self.kind == SPHEROIDAL else 1
        return int(dofs[node * per + component])

    def solve(self, b_dev, nrhs):
        """DegreeOperator::Solve on a device tensor (ndof x nrhs column-major,
        i.e. a contiguous (nrhs, ndof) float64 tensor), in place."""
        assert b_dev.is_cuda and b_dev.dtype.is_floating_point and b_dev.is_contiguous()
        assert b_dev.numel() == self.ndof * nrhs
        used = ctypes.c_int()
        check(lib.sl3_degree_operator_solve(self.handle, ctypes.c_void_p(b_dev.data_ptr()), nrhs,
                                            ctypes.byref(used)))
        return bool(used.value)

    def apply(self, x, nrhs):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        check(lib.sl3_degree_operator_apply(self.handle, _dp(x), nrhs, _dp(y)))
        return y


class ViscoelasticOperator:
    """The radially symmetric Maxwell evolution of every spheroidal degree
    0..lmax on the GPU (sl3_maxwell1d_*): ViscoelasticOperator's ZeroState,
    Step under one load channel and Surface (include/sl3dpp/Viscoelastic.hpp:
    163-301). Each degree carries 2 (l + 1) real columns (Re of m = 0..l,
    then Im); the load is a pattern of surface-mass coefficients (complex
    per (l, m >= 0) at l(l+1)/2 + m, non-dimensional) times a ramp sampled
    at the step times."""

    def __init__(self, earth, lmax, eps=1e-8):
        self.earth = earth
        self.lmax = lmax
        h = ctypes.c_void_p()
        check(lib.sl3_maxwell1d_create(earth.handle, lmax, ctypes.c_double(eps), ctypes.byref(h)))
        self.handle = h
        mem, sol, nodes = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        check(lib.sl3_maxwell1d_shape(self.handle, ctypes.byref(mem), ctypes.byref(sol),
                                      ctypes.byref(nodes), None))
        self.memory_doubles, self.solution_doubles = mem.value, sol.value
        self.nodes = nodes.value
        self.degrees = []
        for l in range(lmax + 1):
            nd, fn = ctypes.c_int(), ctypes.c_int()
            mo, so = ctypes.c_int64(), ctypes.c_int64()
            check(lib.sl3_maxwell1d_degree(self.handle, l, ctypes.byref(nd), ctypes.byref(fn),
                                           ctypes.byref(mo), ctypes.byref(so)))
            self.degrees.append({"ndof": nd.value, "first_node": fn.value,
                                 "memory_offset": mo.value, "solution_offset": so.value})

    def __del__(self, _lib=lib):
        if getattr(self, "handle", None):
            _lib.sl3_maxwell1d_destroy(self.handle)
            self.handle = None

    def reset(self):
        check(lib.sl3_maxwell1d_reset(self.handle))

    def set_load(self, pattern, ramp):
        """pattern: complex (l, m) coefficients, a ((lmax+1)(lmax+2)/2,)
        complex array or its (Re, Im) doubles; ramp: the factor at every
        step time."""
        p = np.asarray(pattern)
        if np.iscomplexobj(p):
            p = np.stack([p.real, p.imag], axis=-1)
        self._pattern = np.ascontiguousarray(p, dtype=np.float64).ravel()
        self._ramp = np.ascontiguousarray(ramp, dtype=np.float64).ravel()
        check(lib.sl3_maxwell1d_set_load(self.handle, _dp(self._pattern),
                                         ctypes.c_int64(self._pattern.size), _dp(self._ramp),
                                         ctypes.c_int64(self._ramp.size)))

    def evolve(self, dt, scheme, steps, graph=True):
        check(lib.sl3_maxwell1d_evolve(self.handle, ctypes.c_double(dt), scheme, steps,
                                       int(graph)))

    def steps_taken(self):
        k = ctypes.c_int()
        check(lib.sl3_maxwell1d_shape(self.handle, None, None, None, ctypes.byref(k)))
        return k.value

    def export(self, name):
        n = {"solutions": self.solution_doubles, "tuning": 4}.get(name, self.memory_doubles)
        out = np.empty(n)
        check(lib.sl3_maxwell1d_export(self.handle, name.encode(), _dp(out), ctypes.c_int64(n)))
        return out

    def memory_of(self, flat, l):
        """Degree l's block of a memory or strain export as (nodes, 3, cols)."""
        d = self.degrees[l]
        cols = 2 * (l + 1)
        n = self.nodes * 3 * cols
        return flat[d["memory_offset"]: d["memory_offset"] + n].reshape(self.nodes, 3, cols)

    def solution_of(self, flat, l):
        """Degree l's block of the solutions as (ndof, cols)."""
        d = self.degrees[l]
        cols = 2 * (l + 1)
        n = d["ndof"] * cols
        return flat[d["solution_offset"]: d["solution_offset"] + n].reshape(d["ndof"], cols)

    def surface(self, component, out_dev=None):
        """Surface U (0), V (1) or phi (2) of the last solutions per degree
        and column, a ((lmax+1)(lmax+2),) device tensor (degree l's columns
        at l(l+1) onwards)."""
        import torch

        if out_dev is None:
            out_dev = torch.empty((self.lmax + 1) * (self.lmax + 2), dtype=torch.float64,
                                  device="cuda")
        check(lib.sl3_maxwell1d_surface(self.handle, component,
                                        ctypes.c_void_p(out_dev.data_ptr())))
        return out_dev


LOVE_CHANNELS = ("u", "phi", "v", "t")


def love_columns(op):
    """ComputeLoveNumbers' right-hand sides (src/LoveNumbers.cpp:116-128):
    the load's force and potential parts, the tangential load and the
    tide (r/a)^l."""
    return np.concatenate([op.export("load_force"), op.export("load_potential"),
                           op.export("tangential"), op.export("tide:-1")])


def love_numbers_from(earth, l, op, x):
    """The twelve generalised Love numbers of degree l from the solutions x
    (4 x ndof) of love_columns: h, l, k (surface U, V, phi) of each channel
    in SI (src/LoveNumbers.cpp:138-156)."""
    L, T, M = earth.length, earth.time, earth.mass
    h_load, k_load = L ** 3 / M, L ** 4 / (M * T * T)
    h_pot, k_pot = T * T / L, 1.0
    top = (len(earth.edges) - 1) * (earth.ngll - 1)
    iu, iv, ip = op.dof(top, U), op.dof(top, V), op.dof(top, PHI)
    out = {}
    for c, name in enumerate(LOVE_CHANNELS):
        hf, kf = (h_pot, k_pot) if name == "t" else (h_load, k_load)
        val = lambda d: x[c, d] if d >= 0 else 0.0  # noqa: E731
        out["h_" + name] = val(iu) * hf
        out["l_" + name] = val(iv) * hf
        out["k_" + name] = val(ip) * kf
    return out


def love_numbers(earth, lmax, device="cuda"):
    """ComputeLoveNumbers (src/LoveNumbers.cpp:87-175) without the axial
    moments: per degree the four channels solved together on the GPU."""
    import torch

    table = []
    for l in range(lmax + 1):
        op = DegreeOperator(earth, l)
        b = torch.tensor(love_columns(op), dtype=torch.float64, device=device)
        op.solve(b, 4)
        table.append(love_numbers_from(earth, l, op, b.cpu().numpy().reshape(4, op.ndof)))
    return table


def traction_numbers(earth, lmax, gravitating, device="cuda"):
    """ComputeTractionNumbers (src/LoveNumbers.cpp:463-512): per degree the
    surface U and V (SI, metres per pascal) of the responses to a unit
    normal traction (h_u, l_u) and a unit tangential traction (h_v, l_v),
    each degree's pair solved on the GPU; NaN where the degree is excluded
    (the non-gravitating degree 1)."""
    import torch

    h_factor = earth.length ** 2 * earth.time ** 2 / earth.mass  # M^-1 L^2 T^2
    out = {k: np.full(lmax + 1, np.nan) for k in ("h_u", "l_u", "h_v", "l_v")}
    nodes = (len(earth.edges) - 1) * (earth.ngll - 1) + 1
    for l in range(lmax + 1):
        if l == 1 and not gravitating:
            continue
        op = DegreeOperator(earth, l, gravitating=gravitating)
        rhs = np.concatenate([op.export("traction"), op.export("traction_tangential")])
        b = torch.tensor(rhs, dtype=torch.float64, device=device)
        op.solve(b, 2)
        x = b.cpu().numpy().reshape(2, op.ndof)
        iu, iv = op.dof(nodes - 1, U), op.dof(nodes - 1, V)
        val = lambda dof, col: x[col, dof] if dof >= 0 else 0.0  # noqa: E731
        out["h_u"][l] = val(iu, 0) * h_factor
        out["l_u"][l] = val(iv, 0) * h_factor
        out["h_v"][l] = val(iu, 1) * h_factor
        out["l_v"][l] = val(iv, 1) * h_factor
    return out
