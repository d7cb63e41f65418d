"""Python mirror of the 3D step's interface over the C ABI: the Earth set-up,
Viscoelastic3D (Step, ZeroState, the exports) and the synthetic inputs of
the benchmark configurations (the three-region Earth, the zonal load rate of
the reference's step benchmark, the lateral viscosity field).

Names and argument meaning follow include/sl3dpp/Viscoelastic3D.hpp; device
arrays are torch tensors on the HIP device (memory and stream plumbing)."""

import ctypes
import math

import numpy as np

from . import check, lib

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_int64_p = ctypes.POINTER(ctypes.c_int64)

YEAR = 365.25 * 24.0 * 3600.0
EXPONENTIAL_TRAPEZOID = 0
BACKWARD_EULER = 1
SDIRK2 = 2
SDIRK2_GAMMA = 0.29289321881345247560  # 1 - 1/sqrt(2)
FIXED_POINT = 0  # Options::StepSolver
CONJUGATE_GRADIENTS = 1

MAX_LAYERS = 8


class EarthSpec(ctypes.Structure):
    _fields_ = [
        ("layers", ctypes.c_int),
        ("r0", ctypes.c_double * MAX_LAYERS),
        ("r1", ctypes.c_double * MAX_LAYERS),
        ("rho0", ctypes.c_double * MAX_LAYERS),
        ("rho1", ctypes.c_double * MAX_LAYERS),
        ("kappa0", ctypes.c_double * MAX_LAYERS),
        ("kappa1", ctypes.c_double * MAX_LAYERS),
        ("mu0", ctypes.c_double * MAX_LAYERS),
        ("mu1", ctypes.c_double * MAX_LAYERS),
        ("elements", ctypes.c_int * MAX_LAYERS),
        ("ngll", ctypes.c_int),
        ("viscous_layer", ctypes.c_int),
        ("viscosity_boundary", ctypes.c_double),
        ("viscosity_upper", ctypes.c_double),
        ("viscosity_lower", ctypes.c_double),
    ]


MAX_TRANSFORM_DEGREES = 32


class StepOptions(ctypes.Structure):
    _fields_ = [
        ("eps", ctypes.c_double),
        ("oversampling", ctypes.c_double),
        ("max_iterations", ctypes.c_int),
        ("rtol", ctypes.c_double),
        ("damping", ctypes.c_double),
        ("anderson_memory", ctypes.c_int),
        ("anderson_rate", ctypes.c_double),
        ("solver", ctypes.c_int),
        ("solve_rtol", ctypes.c_double),
        ("solve_max_iterations", ctypes.c_int),
        ("graph_passes", ctypes.c_int),
        ("transform_degree_count", ctypes.c_int),
        ("transform_degrees_", ctypes.c_int * MAX_TRANSFORM_DEGREES),
        ("realisations", ctypes.c_int),
    ]

    @classmethod
    def default(cls, **overrides):
        o = cls()
        lib.sl3_step_options_default(ctypes.byref(o))
        for k, v in overrides.items():
            setattr(o, k, v)
        return o

    @property
    def transform_degrees(self):
        return list(self.transform_degrees_[:self.transform_degree_count])

    @transform_degrees.setter
    def transform_degrees(self, degrees):
        degrees = list(degrees)
        if len(degrees) > MAX_TRANSFORM_DEGREES:
            raise ValueError(f"at most {MAX_TRANSFORM_DEGREES} transform degrees")
        self.transform_degree_count = len(degrees)
        for i, d in enumerate(degrees):
            self.transform_degrees_[i] = int(d)


class StepInfo(ctypes.Structure):
    """Viscoelastic3D::StepInfo plus the device time of the step's phases."""

    _fields_ = [
        ("iterations", ctypes.c_int),
        ("residual", ctypes.c_double),
        ("converged", ctypes.c_int),
        ("mixed_passes", ctypes.c_int),
        ("start_ms", ctypes.c_float),
        ("passes_ms", ctypes.c_float),
        ("finish_ms", ctypes.c_float),
        ("solve_iterations", ctypes.c_int),
        ("warm_fallback", ctypes.c_int),
    ]


LOAD_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.c_double,
                           c_double_p, ctypes.c_void_p)
def device_copy(dst, src, nbytes):
    """bytes from one device address to another on the library's stream
    (sl3_copy), synchronised."""
    check(lib.sl3_copy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_int64(nbytes)))
    check(lib.sl3_synchronize())


# sl3_pass_fn: (s_dev, next_dev, user) -> 0 or a failure
PASS_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)

lib.sl3_step_options_default.restype = None


def three_region_earth(core=70, mantle=56, lithosphere=2, ngll=6):
    """The benchmark configurations' Earth: radius 6371 km; fluid core to
    3480 km (rho 12000 -> 9900, kappa 1300 -> 650 GPa, mu 0); mantle to
    6271 km (rho 5500 -> 3400, kappa 650 -> 130 GPa, mu 290 -> 65 GPa);
    lithosphere (rho 3300, kappa 130 GPa, mu 65 GPa), all linear in r; a
    Maxwell mantle of 5e20 Pa s above 5701 km and 1e22 below, the
    lithosphere elastic. Elements per region and GLL nodes per element as
    given (128 elements x 6 nodes for the 3D step)."""
    s = EarthSpec()
    s.layers = 3
    G = 1e9
    rows = [
        (0.0, 3480e3, 12000.0, 9900.0, 1300 * G, 650 * G, 0.0, 0.0, core),
        (3480e3, 6271e3, 5500.0, 3400.0, 650 * G, 130 * G, 290 * G, 65 * G, mantle),
        (6271e3, 6371e3, 3300.0, 3300.0, 130 * G, 130 * G, 65 * G, 65 * G, lithosphere),
    ]
    for i, (r0, r1, d0, d1, k0, k1, m0, m1, n) in enumerate(rows):
        s.r0[i], s.r1[i], s.rho0[i], s.rho1[i] = r0, r1, d0, d1
        s.kappa0[i], s.kappa1[i], s.mu0[i], s.mu1[i] = k0, k1, m0, m1
        s.elements[i] = n
    s.ngll = ngll
    s.viscous_layer = 1
    s.viscosity_boundary = 5701e3
    s.viscosity_upper = 5e20
    s.viscosity_lower = 1e22
    return s


class Earth:
    """The layered Earth, mesh, material, Maxwell rheology (host)."""

    def __init__(self, spec=None):
        self.spec = spec if spec is not None else three_region_earth()
        self.handle = ctypes.c_void_p()
        check(lib.sl3_earth_create(ctypes.byref(self.spec), ctypes.byref(self.handle)))

    def __del__(self, _lib=lib):  # the default keeps the library through shutdown
        if getattr(self, "handle", None):
            _lib.sl3_earth_destroy(self.handle)
            self.handle = None

    def scales(self):
        length, time, mass = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib.sl3_earth_scales(self.handle, ctypes.byref(length),
                                   ctypes.byref(time), ctypes.byref(mass)))
        return length.value, time.value, mass.value

    def mesh(self):
        e, n = ctypes.c_int(), ctypes.c_int()
        check(lib.sl3_earth_mesh(self.handle, ctypes.byref(e), ctypes.byref(n)))
        return e.value, n.value

    def export(self, name):
        count = ctypes.c_int64()
        check(lib.sl3_earth_export(self.handle, name.encode(), None, ctypes.byref(count)))
        out = np.zeros(count.value)
        check(lib.sl3_earth_export(self.handle, name.encode(),
                                   out.ctypes.data_as(c_double_p), ctypes.byref(count)))
        return out

    def degree_system(self, l, kind, scheme=EXPONENTIAL_TRAPEZOID, dt=-1.0, eps=1e-8):
        """The assembled degree-l system: ndof, kd, first element, pinned
        dof, unscaled upper band ((ndof, kd+1): row j is column j of the
        band), dof map, unit load and unit tide (the potential (r/a)^p,
        p = 2 at l = 0, else l)."""
        ndof, kd, first, pinned = (ctypes.c_int() for _ in range(4))
        check(lib.sl3_earth_degree_system(self.handle, l, kind, scheme, ctypes.c_double(dt),
                                          ctypes.c_double(eps), ctypes.byref(ndof),
                                          ctypes.byref(kd), ctypes.byref(first),
                                          ctypes.byref(pinned), None, None, None, None))
        elements, ngll = self.mesh()
        nodes = elements * (ngll - 1) + 1
        band = np.zeros((ndof.value, kd.value + 1))
        dofs = np.zeros(nodes * (3 if kind == 0 else 1), dtype=np.int32)
        load = np.zeros(ndof.value)
        tide = np.zeros(ndof.value)
        check(lib.sl3_earth_degree_system(
            self.handle, l, kind, scheme, ctypes.c_double(dt), ctypes.c_double(eps),
            ctypes.byref(ndof), ctypes.byref(kd), ctypes.byref(first), ctypes.byref(pinned),
            band.ctypes.data_as(c_double_p), dofs.ctypes.data_as(c_int_p),
            load.ctypes.data_as(c_double_p) if kind == 0 else None,
            tide.ctypes.data_as(c_double_p) if kind == 0 else None))
        return dict(ndof=ndof.value, kd=kd.value, first_element=first.value,
                    pinned=pinned.value, band=band,
                    dofs=dofs.reshape(nodes, 3) if kind == 0 else dofs,
                    load=load if kind == 0 else None,
                    tide=tide if kind == 0 else None)


def zonal_load_rate(lmax, earth):
    """The reference step benchmark's load (benchmarks/step_timing.cpp:
    185-199): zonal columns sigma_l(t) = rate_l t with rate_l =
    exp(-2 pi (l+1)/10.5) sqrt((2l+1)/4 pi) 5515 kg m^-2 / 30 yr for l >= 1,
    in the non-dimensional units. Returns a LoadCallable3D for the C ABI."""
    length, time, mass = earth.scales()
    sigma = mass / length ** 2
    rate = np.zeros(lmax + 1)
    for l in range(1, lmax + 1):
        rate[l] = (math.exp(-2.0 * math.pi * (l + 1) / 10.5)
                   * math.sqrt((2.0 * l + 1) / (4.0 * math.pi)) * 5515.0 / (30.0 * YEAR))

    def load(l, t, columns, user):
        if l < 1 or l > lmax:
            return 0
        # the 2 (l + 1) columns zeroed in one call (an element-wise loop
        # through the pointer costs about 1 ms per step at lmax 128, with
        # the device idle before the step's start)
        ctypes.memset(columns, 0, 16 * (l + 1))
        columns[0] = rate[l] * (t * time) / sigma
        return 1

    return LOAD_FN(load), rate


def delta_coefficients(lmax_field, nodes, seed=2026, rms=0.7, ar=0.9, lmax=None):
    """The lateral log-viscosity field of the 3D step configurations: at each
    memory node an isotropic Gaussian random field on the sphere with angular
    power (1 + l/8)^-3 for l <= lmax_field (zero above) and rms `rms`,
    successive nodes (in radius order) AR(1) with coefficient `ar`:
    c_k = ar c_{k-1} + sqrt(1 - ar^2) xi_k. Spin-0 coefficients in the real
    m >= 0 layout l(l+1)/2 + m up to `lmax`, complex, per node end to end."""
    lmax = lmax_field if lmax is None else lmax
    rng = np.random.default_rng(seed)
    size = (lmax + 1) * (lmax + 2) // 2
    power = np.array([(1.0 + l / 8.0) ** -3 if l <= lmax_field else 0.0
                      for l in range(lmax + 1)])
    variance = sum((2 * l + 1) * power[l] for l in range(lmax + 1)) / (4 * math.pi)
    power *= rms ** 2 / variance
    out = np.zeros((nodes, size), dtype=np.complex128)
    prev = None
    for k in range(nodes):
        xi = np.zeros(size, dtype=np.complex128)
        for l in range(lmax_field + 1):
            base = l * (l + 1) // 2
            sd = math.sqrt(power[l])
            xi[base] = rng.normal(0.0, sd)
            m = l
            if m > 0:
                re = rng.normal(0.0, sd / math.sqrt(2.0), size=l)
                im = rng.normal(0.0, sd / math.sqrt(2.0), size=l)
                xi[base + 1: base + l + 1] = re + 1j * im
        cur = xi if prev is None else ar * prev + math.sqrt(1.0 - ar * ar) * xi
        out[k] = cur
        prev = cur
    return out


class State:
    """Viscoelastic3D::State on the device (one Maxwell branch)."""

    def __init__(self, op, _from=None):
        self.op = op
        self.handle = ctypes.c_void_p()
        if _from is None:
            check(lib.sl3_state_zero(op.handle, ctypes.byref(self.handle)))
        else:
            check(lib.sl3_state_clone(op.handle, _from.handle, ctypes.byref(self.handle)))

    def clone(self):
        """A device copy of the state (sl3_state_clone)."""
        return State(self.op, _from=self)

    def __del__(self, _lib=lib):  # the default keeps the library through shutdown
        if getattr(self, "handle", None):
            _lib.sl3_state_destroy(self.handle)
            self.handle = None

    @property
    def t(self):
        t = ctypes.c_double()
        check(lib.sl3_state_time(self.handle, ctypes.byref(t)))
        return t.value

    def export(self, name):
        count = ctypes.c_int64()
        check(lib.sl3_state_export(self.op.handle, self.handle, name.encode(), None,
                                   ctypes.byref(count)))
        out = np.zeros(count.value)
        check(lib.sl3_state_export(self.op.handle, self.handle, name.encode(),
                                   out.ctypes.data_as(c_double_p), ctypes.byref(count)))
        return out


class Viscoelastic3D:
    """The 3D lateral-viscosity step on the GPU (Viscoelastic3D of
    include/sl3dpp/Viscoelastic3D.hpp). delta: a torch tensor on the device
    holding delta on every memory node's grid, or None."""

    def __init__(self, earth, lmax, delta=None, options=None):
        import torch

        self.earth = earth
        self.lmax = lmax
        self.delta = delta
        self.options = options if options is not None else StepOptions.default()
        self.handle = ctypes.c_void_p()
        torch.cuda.synchronize()
        ptr = ctypes.c_void_p(delta.data_ptr()) if delta is not None else None
        check(lib.sl3_operator_create(earth.handle, lmax, ptr, ctypes.byref(self.options),
                                      ctypes.byref(self.handle)))
        nodes, nt, nph = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        sd = ctypes.c_int64()
        check(lib.sl3_operator_shape(self.handle, ctypes.byref(nodes), ctypes.byref(nt),
                                     ctypes.byref(nph), ctypes.byref(sd)))
        self.nodes, self.ntheta, self.nphi = nodes.value, nt.value, nph.value
        self.spectral_doubles = sd.value
        # realisations side by side: realisation r's memory nodes are
        # nodes r * radial_nodes ... of every node-indexed tensor
        real, radial = ctypes.c_int(), ctypes.c_int()
        check(lib.sl3_operator_realisations(self.handle, ctypes.byref(real),
                                            ctypes.byref(radial)))
        self.realisations, self.radial_nodes = real.value, radial.value
        self.radii = np.zeros(self.nodes)
        self.theta = np.zeros(self.ntheta)
        self.wtheta = np.zeros(self.ntheta)
        check(lib.sl3_operator_geometry(self.handle, self.radii.ctypes.data_as(c_double_p),
                                        self.theta.ctypes.data_as(c_double_p),
                                        self.wtheta.ctypes.data_as(c_double_p)))
        # the node blocks (transform degrees): every grid field lays the
        # blocks' grids end to end, grid_values values in all
        count, gv = ctypes.c_int(), ctypes.c_int64()
        check(lib.sl3_operator_blocks(self.handle, 0, ctypes.byref(count), None, None, None,
                                      None, None, ctypes.byref(gv)))
        n = count.value
        arrays = [(ctypes.c_int * n)() for _ in range(5)]
        check(lib.sl3_operator_blocks(self.handle, n, ctypes.byref(count), *arrays, None))
        self.grid_values = gv.value
        self.blocks = []
        offset = 0
        for i in range(n):
            first, nodes, lmax_b, nt_b, nph_b = (a[i] for a in arrays)
            th, wt = np.zeros(nt_b), np.zeros(nt_b)
            check(lib.sl3_operator_block_geometry(self.handle, i, th.ctypes.data_as(c_double_p),
                                                  wt.ctypes.data_as(c_double_p)))
            self.blocks.append({"first_node": first, "nodes": nodes, "lmax": lmax_b,
                                "ntheta": nt_b, "nphi": nph_b, "theta": th, "wtheta": wt,
                                "field_offset": offset})
            offset += nodes * nt_b * nph_b

    def __del__(self, _lib=lib):  # the default keeps the library through shutdown
        if getattr(self, "handle", None):
            _lib.sl3_operator_destroy(self.handle)
            self.handle = None

    def set_options(self, **overrides):
        for k, v in overrides.items():
            setattr(self.options, k, v)
        check(lib.sl3_operator_set_options(self.handle, ctypes.byref(self.options)))

    def set_delta(self, delta):
        import torch

        torch.cuda.synchronize()
        ptr = ctypes.c_void_p(delta.data_ptr()) if delta is not None else None
        check(lib.sl3_operator_set_delta(self.handle, ptr))
        self.delta = delta

    def delta_grid(self):
        """A host copy of the lateral field last set ([node][theta][phi]),
        or None for delta = 0."""
        d = getattr(self, "delta", None)
        return None if d is None else d.detach().cpu().numpy().copy()

    def zero_state(self):
        return State(self)

    def step(self, state, dt, scheme=EXPONENTIAL_TRAPEZOID, load=None, start=None, tide=None,
             guard=0.0):
        """Step (src/Viscoelastic3D.cpp:1982-2030) in place on the state;
        load and tide are LOAD_FN channels or None; start a device-layout
        warm start (a torch tensor) or None; guard > 0 drops a warm start
        whose first pass changes it by more than guard."""
        info = StepInfo()
        ptr = ctypes.c_void_p(start.data_ptr()) if start is not None else None
        check(lib.sl3_step_guarded(self.handle, state.handle, ctypes.c_double(dt), scheme,
                                   load, tide, None, ptr, ctypes.c_double(guard),
                                   ctypes.byref(info)))
        return info

    def realisation_infos(self):
        """Per realisation, the last step's own StepInfo."""
        count = ctypes.c_int()
        check(lib.sl3_step_realisation_infos(self.handle, 0, ctypes.byref(count), None))
        infos = (StepInfo * max(count.value, 1))()
        check(lib.sl3_step_realisation_infos(self.handle, count.value, ctypes.byref(count),
                                             infos))
        return list(infos[:count.value])

    def export(self, name):
        count = ctypes.c_int64()
        check(lib.sl3_operator_export(self.handle, name.encode(), None, ctypes.byref(count)))
        out = np.zeros(count.value)
        check(lib.sl3_operator_export(self.handle, name.encode(),
                                      out.ctypes.data_as(c_double_p), ctypes.byref(count)))
        return out

    def synthesise(self, spin, coefficients, field):
        """SurfaceGrid::Synthesise(Spin)Batch over the memory nodes, device
        tensors in the reference layouts."""
        import torch

        torch.cuda.synchronize()
        check(lib.sl3_gsh_synthesise(self.handle, spin, ctypes.c_void_p(coefficients.data_ptr()),
                                     ctypes.c_void_p(field.data_ptr())))

    def analyse(self, spin, field, coefficients):
        import torch

        torch.cuda.synchronize()
        check(lib.sl3_gsh_analyse(self.handle, spin, ctypes.c_void_p(field.data_ptr()),
                                  ctypes.c_void_p(coefficients.data_ptr())))

    # ---- OpenStep (include/sl3dpp/Viscoelastic3D.hpp:461-647) -------------

    def start_step(self, state, dt, scheme=EXPONENTIAL_TRAPEZOID, load=None, tide=None):
        """StartStep: the state's memory becomes m~ until finish_step."""
        import torch

        torch.cuda.synchronize()
        check(lib.sl3_step_start(self.handle, state.handle, ctypes.c_double(dt), scheme, load,
                                 tide, None))

    def step_pass(self, s, nxt):
        """One pass (Respond, AddChannels, StrainOf) of device tensor s into
        nxt."""
        check(lib.sl3_step_pass(self.handle, ctypes.c_void_p(s), ctypes.c_void_p(nxt)))

    def converge(self, pass_fn=None, start=None):
        """Converge with a pass (a PASS_FN or None for step_pass)."""
        info = StepInfo()
        ptr = ctypes.c_void_p(start.data_ptr()) if start is not None else None
        check(lib.sl3_step_converge(self.handle, pass_fn, None, ptr, ctypes.byref(info)))
        return info

    def close_step(self, start=None):
        info = StepInfo()
        ptr = ctypes.c_void_p(start.data_ptr()) if start is not None else None
        check(lib.sl3_step_close(self.handle, ptr, ctypes.byref(info)))
        return info

    def finish_step(self, strain=None):
        ptr = ctypes.c_void_p(strain.data_ptr()) if strain is not None else None
        check(lib.sl3_step_finish(self.handle, ptr))

    def cancel_step(self):
        check(lib.sl3_step_cancel(self.handle))

    def step_view(self, name):
        """(device pointer, doubles) of an open step's buffer."""
        ptr = ctypes.c_void_p()
        count = ctypes.c_int64()
        check(lib.sl3_step_view(self.handle, name.encode(), ctypes.byref(ptr),
                                ctypes.byref(count)))
        return ptr.value, count.value

    def add_force(self, analysed):
        check(lib.sl3_step_add_force(self.handle, ctypes.c_void_p(analysed.data_ptr())))

    def add_memory(self, blocks):
        arr = (ctypes.c_void_p * 3)(*[b.data_ptr() for b in blocks])
        check(lib.sl3_step_add_memory(self.handle, arr))

    def transposed_pass(self, z, out, src=None):
        """The discrete adjoint's transposed pass of the open step
        (src/AdjointDiscrete3D.cpp:998-1037 without the sea-level duals):
        out = Analyse(W Synthesise(StrainOf(SolveForces(Entry(z)))) + src),
        device-layout spectral tensors z and out, src three grid blocks or
        None."""
        arr = None
        if src is not None:
            arr = (ctypes.c_void_p * 3)(*[b.data_ptr() for b in src])
        check(lib.sl3_step_transposed_pass(self.handle, ctypes.c_void_p(z.data_ptr()), arr,
                                           ctypes.c_void_p(out.data_ptr())))

    def relative_change(self, nxt, now):
        out = ctypes.c_double()
        check(lib.sl3_step_relative_change(self.handle, ctypes.c_void_p(nxt.data_ptr()),
                                           ctypes.c_void_p(now.data_ptr()), ctypes.byref(out)))
        return out.value

    # ---- GridCoupling (include/sl3dpp/GridCoupling.hpp:166-384) -----------

    def solution_doubles(self):
        n = ctypes.c_int64()
        check(lib.sl3_coupling_solution_doubles(self.handle, ctypes.byref(n)))
        return n.value

    def coupling_force(self, weighted, forces):
        """SpheroidalForce + ToroidalForce of a device-layout tensor into the
        Solutions layout (every task end to end)."""
        check(lib.sl3_coupling_force(self.handle, ctypes.c_void_p(weighted.data_ptr()),
                                     ctypes.c_void_p(forces.data_ptr())))

    def coupling_strain(self, solutions, spectral):
        check(lib.sl3_coupling_strain(self.handle, ctypes.c_void_p(solutions.data_ptr()),
                                      ctypes.c_void_p(spectral.data_ptr())))

    def multiply(self, blocks, weight):
        check(lib.sl3_coupling_multiply(self.handle, *[ctypes.c_void_p(b.data_ptr())
                                                        for b in blocks],
                                        ctypes.c_void_p(weight.data_ptr())))

    def scale_by_node(self, factor, spectral):
        f = np.ascontiguousarray(factor, dtype=np.float64)
        check(lib.sl3_coupling_scale_by_node(self.handle, f.ctypes.data_as(c_double_p),
                                             ctypes.c_void_p(spectral.data_ptr())))

    def to_reference(self, device, reference):
        check(lib.sl3_spectral_to_reference(self.handle, ctypes.c_void_p(device.data_ptr()),
                                            ctypes.c_void_p(reference.data_ptr())))

    def from_reference(self, reference, device):
        check(lib.sl3_spectral_from_reference(self.handle, ctypes.c_void_p(reference.data_ptr()),
                                              ctypes.c_void_p(device.data_ptr())))

    def delta_field(self, coefficients):
        """delta on every node's grid from per-node spin-0 coefficients
        (delta_coefficients), synthesised on the device."""
        import torch

        coef = torch.from_numpy(np.ascontiguousarray(coefficients).view(np.float64).ravel()).cuda()
        field = torch.zeros(self.grid_values, dtype=torch.float64, device="cuda")
        self.synthesise(0, coef, field)
        return field


lib.sl3_spectral_combine.argtypes = [ctypes.c_void_p, ctypes.c_int, c_double_p,
                                     ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]
lib.sl3_state_spectral.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]


def lagrange_extrapolation_weights(times, target):
    """Weights c_j with p(target) = sum_j c_j p(times[j]) for every
    polynomial p of degree < len(times) (Lagrange interpolation through the
    points, evaluated at the target)."""
    out = []
    for j, tj in enumerate(times):
        w = 1.0
        for k, tk in enumerate(times):
            if k != j:
                w *= (target - tk) / (tj - tk)
        out.append(w)
    return out


def spectral_combine(op, weights, sources, dst):
    """dst = sum_j weights[j] sources[j] over op's spectral tensor
    (sl3_spectral_combine; sources are device pointers, dst a tensor)."""
    n = len(weights)
    c = (ctypes.c_double * n)(*weights)
    src = (ctypes.c_void_p * n)(*sources)
    check(lib.sl3_spectral_combine(op.handle, n, c, src, ctypes.c_void_p(dst.data_ptr())))


lib.sl3_extrapolator_create.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]
lib.sl3_extrapolator_destroy.argtypes = [ctypes.c_void_p]
lib.sl3_extrapolator_reset.argtypes = [ctypes.c_void_p]


class ExtrapolatedStart:
    """Warm starts along a trajectory (sl3_extrapolator). The next step's
    strain coefficients are extrapolated in time from the last converged
    steps' (the Lagrange polynomial through up to order + 1 of them, at
    their times) and handed to the step as the `start` of
    OpenStep::Converge (src/Viscoelastic3D.cpp:1916-1921). The fixed point
    then iterates to the same tolerance, so a step's result is the plain
    step's to that tolerance. With fewer than two recorded steps, for Sdirk2
    (which takes no start), and for a state other than the last one stepped
    (a new history), the step uses its own start. guard > 0 drops an
    extrapolated start whose first pass changes it by more than guard."""

    def __init__(self, op, order=3, guard=0.0):
        self.op, self.order, self.guard = op, order, guard
        self.handle = ctypes.c_void_p()
        check(lib.sl3_extrapolator_create(op.handle, order, ctypes.byref(self.handle)))

    def __del__(self, _lib=lib):  # the default keeps the library through shutdown
        if getattr(self, "handle", None):
            _lib.sl3_extrapolator_destroy(self.handle)
            self.handle = None

    def reset(self):
        check(lib.sl3_extrapolator_reset(self.handle))

    def step(self, state, dt, scheme=EXPONENTIAL_TRAPEZOID, load=None, tide=None):
        info = StepInfo()
        check(lib.sl3_extrapolated_step(self.handle, state.handle, ctypes.c_double(dt), scheme,
                                        load, tide, None, ctypes.c_double(self.guard),
                                        ctypes.byref(info)))
        return info
