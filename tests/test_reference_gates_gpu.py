"""The HIP path through the reference's own value gates, and the OpenStep and
GridCoupling boundary:

 * the zero-lateral pipeline (tests/TestGridCoupling.cpp:178-332) on the
   device: strain recovery (sl3_coupling_strain), synthesis, the product
   with a unit weight (sl3_coupling_multiply), analysis, the node scaling
   by mu (sl3_coupling_scale_by_node) and the force (sl3_coupling_force)
   equal the point-shear part of the assembled systems applied to the
   displacement (1e-11), and each stage equals the oracle's (1e-12);
 * the zero-lateral step (tests/TestViscoelastic3D.cpp:90-196): with
   delta = 0 the HIP 3D step's solutions over two steps equal the oracle's
   degree-basis stepper's (1e-10), exponential trapezoid and backward
   Euler, under a load with every order growing from zero;
 * the open step: sl3_step_start + sl3_step_converge + sl3_step_finish, with
   the library's pass or a caller's pass that calls sl3_step_pass, give
   sl3_step's state bit for bit; the views are the operator's buffers;
 * a non-finite pass ends the step unconverged with a NaN residual and the
   step runs no further pass (tests/TestViscoelastic3D.cpp:821-845)."""

import ctypes

import numpy as np
import pytest
from reference_gates import (DegreeBasisStepper, forces_gap, general_load, join,
                             oracle_pipeline_force, point_shear_force, random_displacement,
                             split, zero_lateral_problem)
from step_problem import amplitudes, small_earth_spec

import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def earth():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return st.Earth(small_earth_spec())


def cuda(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def device_pipeline(op, x_flat, mu):
    """strain -> synthesis -> x 1 -> analysis -> x mu -> force, all on the
    device through the C ABI; returns the forces and the strain
    coefficients (reference layout)."""
    import torch

    S = op.spectral_doubles
    sol = cuda(x_flat)
    spec = torch.zeros(S, dtype=torch.float64, device="cuda")
    op.coupling_strain(sol, spec)
    ref = torch.zeros_like(spec)
    op.to_reference(spec, ref)
    L, nodes, fs = op.lmax, op.nodes, op.ntheta * op.nphi
    sizes = [(L + 1) ** 2 - 4, (L + 1) ** 2 - 1, (L + 1) * (L + 2) // 2]
    grids, off = [], 0
    for b, spin in enumerate((-2, -1, 0)):
        n = 2 * sizes[b] * nodes
        g = torch.zeros((1 if spin == 0 else 2) * nodes * fs, dtype=torch.float64, device="cuda")
        op.synthesise(spin, ref[off: off + n], g)
        grids.append(g)
        off += n
    ones = torch.ones(nodes * fs, dtype=torch.float64, device="cuda")
    op.multiply(grids, ones)
    back = torch.zeros_like(ref)
    off = 0
    for b, spin in enumerate((-2, -1, 0)):
        n = 2 * sizes[b] * nodes
        op.analyse(spin, grids[b], back[off: off + n])
        off += n
    weighted = torch.zeros_like(spec)
    op.from_reference(back, weighted)
    op.scale_by_node(mu, weighted)
    forces = torch.zeros(op.solution_doubles(), dtype=torch.float64, device="cuda")
    op.coupling_force(weighted, forces)
    torch.cuda.synchronize()
    return forces.cpu().numpy(), ref.cpu().numpy()


@pytest.mark.parametrize("lmax", [8, 12])
def test_zero_lateral_pipeline_on_the_device(earth, oracle, lmax):
    print(f"[gates] pipeline lmax {lmax}", flush=True)
    L, T, M = earth.scales()
    op = st.Viscoelastic3D(earth, lmax, None)
    ref = zero_lateral_problem(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, 100 * st.YEAR / T,
                               theta=op.theta, wtheta=op.wtheta)
    assert op.solution_doubles() == sum(2 * s["ndof"] * (l + 1) for s, l in
                                        zip(ref.systems, list(range(lmax + 1))
                                            + list(range(1, lmax + 1))))
    x = random_displacement(ref.systems, lmax)
    forces, coef = device_pipeline(op, join(x), ref.mu_node)
    got = split(forces, ref.systems, lmax)
    want = point_shear_force(earth, lmax, x)
    pinned = [s["pinned"] for s in ref.systems]
    gap, scale = forces_gap(got, want, pinned)
    print(f"[gates] lmax {lmax}: device pipeline vs point shear {gap:.2e}", flush=True)
    assert scale > 0.0 and gap < 1e-11
    oracle_forces, oracle_coef = oracle_pipeline_force(oracle, ref, x)
    assert rel(coef, oracle_coef) < 1e-12  # strain recovery
    gap_o, _ = forces_gap(got, oracle_forces, pinned)
    assert gap_o < 1e-12


def test_force_of_a_random_tensor_matches_the_oracle(earth, oracle):
    """sl3_coupling_force alone on a random analysed tensor."""
    import torch

    lmax = 10
    L, T, M = earth.scales()
    op = st.Viscoelastic3D(earth, lmax, None)
    ref = zero_lateral_problem(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, 100 * st.YEAR / T,
                               theta=op.theta, wtheta=op.wtheta)
    rng = np.random.default_rng(3)
    coef = rng.normal(size=ref.spectral_doubles)
    # a real field's spin-0 m = 0 coefficients are real
    L1 = lmax + 1
    spin0 = coef[2 * (2 * L1 * L1 - 5) * op.nodes:].reshape(op.nodes, -1, 2)
    for l in range(lmax + 1):
        spin0[:, l * (l + 1) // 2, 1] = 0.0
    dev = torch.zeros(op.spectral_doubles, dtype=torch.float64, device="cuda")
    op.from_reference(cuda(coef), dev)
    forces = torch.zeros(op.solution_doubles(), dtype=torch.float64, device="cuda")
    op.coupling_force(dev, forces)
    want = np.zeros(op.solution_doubles())
    DP = ctypes.POINTER(ctypes.c_double)
    oracle.sl3o_memory_force(ctypes.byref(ref.problem), coef.ctypes.data_as(DP),
                             want.ctypes.data_as(DP))
    got = forces.cpu().numpy()
    pinned = [s["pinned"] for s in ref.systems]
    gap, _ = forces_gap(split(got, ref.systems, lmax), split(want, ref.systems, lmax), pinned)
    assert gap < 1e-12


@pytest.mark.parametrize("scheme", [st.EXPONENTIAL_TRAPEZOID, st.BACKWARD_EULER])
def test_zero_lateral_step_matches_the_degree_basis_stepper(earth, oracle, scheme):
    lmax = 8
    print(f"[gates] delta 0 step scheme {scheme}", flush=True)
    tau = 1.0 / np.max(earth.export("point_itau"))
    dt = 0.7 * tau
    op = st.Viscoelastic3D(earth, lmax, None,
                           st.StepOptions.default(max_iterations=10, rtol=1e-13))
    ref = zero_lateral_problem(oracle, earth, lmax, scheme, dt, theta=op.theta,
                               wtheta=op.wtheta)
    one = DegreeBasisStepper(oracle, ref)
    load = general_load(lmax, earth)
    state = op.zero_state()
    for k in range(2):
        info = op.step(state, dt, scheme, load)
        assert info.converged == 1 and info.iterations <= 3
        x1 = one.step(dt, scheme, amplitudes(load, lmax, (k + 1) * dt))
        scale = max(np.max(np.abs(x)) for x in x1)
        gap = max(np.max(np.abs(op.export(f"solution:{l}") - x1[l])) for l in range(lmax + 1))
        tor = max(np.max(np.abs(op.export(f"solution:{lmax + l}"))) for l in range(1, lmax + 1))
        print(f"[gates] step {k}: solutions vs degree basis {gap / scale:.2e}, "
              f"toroidal {tor / scale:.1e}", flush=True)
        assert gap < 1e-10 * scale
        assert tor < 1e-12 * scale


def make_lateral(earth, lmax, options):
    op = st.Viscoelastic3D(earth, lmax, None, options)
    op.set_delta(op.delta_field(st.delta_coefficients(6, op.nodes, seed=31, rms=1.2, lmax=lmax)))
    return op


def test_open_step_pieces_equal_the_step(earth):
    import torch

    lmax = 12
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    opts = st.StepOptions.default(max_iterations=60, rtol=1e-10, anderson_memory=6,
                                  anderson_rate=0.5)
    op = make_lateral(earth, lmax, opts)
    load, _ = st.zonal_load_rate(lmax, earth)
    a, b, c = op.zero_state(), op.zero_state(), op.zero_state()
    calls = []

    def pass_fn(s, nxt, user):
        calls.append(1)
        op.step_pass(s, nxt)
        return 0

    pass_c = st.PASS_FN(pass_fn)
    for k in range(2):
        print(f"[gates] open step {k}", flush=True)
        ia = op.step(a, dt, st.EXPONENTIAL_TRAPEZOID, load)
        op.start_step(b, dt, st.EXPONENTIAL_TRAPEZOID, load)
        weight, nw = op.step_view("weight")
        assert nw == op.nodes * op.ntheta * op.nphi
        ib = op.converge()
        op.finish_step()
        op.start_step(c, dt, st.EXPONENTIAL_TRAPEZOID, load)
        calls.clear()
        ic = op.converge(pass_c)
        assert len(calls) == ic.iterations
        op.finish_step()
        assert ia.iterations == ib.iterations == ic.iterations
        assert ia.residual == ib.residual == ic.residual
        for name in ["coefficients", "m0", "m1", "m2", "strain2"]:
            ea = a.export(name)
            assert np.array_equal(ea, b.export(name)), name
            assert np.array_equal(ea, c.export(name)), name
        assert a.t == b.t == c.t
    # the views: the step's weight and beta as the operator exports them
    op.start_step(b, dt, st.EXPONENTIAL_TRAPEZOID, load)
    for name in ("weight", "beta"):
        ptr, n = op.step_view(name)
        view = torch.empty(n, dtype=torch.float64, device="cuda")
        st.device_copy(view.data_ptr(), ptr, 8 * n)
        assert np.array_equal(view.cpu().numpy(), op.export(name)), name
    op.cancel_step()
    with pytest.raises(Exception, match="no step is open"):
        op.finish_step()


def test_add_force_moves_the_step(earth):
    """OpenStep::AddForce: a force added to the fixed force is what a step
    with that memory force sees; adding zero changes nothing."""
    import torch

    lmax = 8
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    opts = st.StepOptions.default(max_iterations=60, rtol=1e-11)
    op = make_lateral(earth, lmax, opts)
    load, _ = st.zonal_load_rate(lmax, earth)
    a, b = op.zero_state(), op.zero_state()
    op.step(a, dt, st.EXPONENTIAL_TRAPEZOID, load)
    op.start_step(b, dt, st.EXPONENTIAL_TRAPEZOID, load)
    op.add_force(torch.zeros(op.spectral_doubles, dtype=torch.float64, device="cuda"))
    op.converge()
    op.finish_step()
    assert rel(b.export("coefficients"), a.export("coefficients")) < 1e-9


def test_a_non_finite_pass_ends_the_step(earth):
    """tests/TestViscoelastic3D.cpp:821-845: a pass whose strain is not
    finite ends Converge at once, unconverged, with a NaN residual."""
    import torch

    lmax = 8
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    op = make_lateral(earth, lmax, st.StepOptions.default(max_iterations=30, rtol=1e-12))
    load, _ = st.zonal_load_rate(lmax, earth)
    state = op.zero_state()
    calls = []

    nan = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")

    def poisoned(s, nxt, user):
        calls.append(1)
        op.step_pass(s, nxt)
        if len(calls) == 2:
            st.device_copy(nxt, nan.data_ptr(), 8)  # one non-finite coefficient
        return 0

    fn = st.PASS_FN(poisoned)
    op.start_step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    info = op.converge(fn)
    assert info.converged == 0
    assert np.isnan(info.residual)
    assert info.iterations == 2 and len(calls) == 2
    op.cancel_step()
