"""The discrete 3D adjoint's transposed pass (sl3_step_transposed_pass;
src/AdjointDiscrete3D.cpp:998-1037 without the sea-level and measurement
duals): out = Analyse(W Synthesise(StrainOf(SolveForces(Entry(z)))) + src).

It is pinned by what it must be, the adjoint of the forward pass's linear
part L(x) = pass(x) - pass(0) (sl3_step_pass: Synthesise, x W, Analyse,
the forces, SolveForces, StrainOf) in the pairing of two real trace-free
tensors over the memory points,

  <a, b> = sum_nodes jac w r^2 sum_(l, m) g_N Re(conj(a) b),

g = 1, 2, 3 for the spin -2, -1, 0 blocks (the tensor contraction's
weights of the stored components, the +1, +2 ones folded in), doubled for
spin 0's m > 0 (the m < 0 half of a real field), jac w r^2 the radial
quadrature at a memory node. For random tensors and a random lateral
field, <L x, u> = <x, T u> to rounding: an operator that differs from the
transpose by a term, a sign, an index or the weight W fails it. (The
weights themselves were found by fitting a diagonal D with T = D^-1 L^T D
on the explicit matrices of a small case: every entry's ratio fits one
such D to rounding, and D is the pairing above.) The source term is the
analysis of src."""

import numpy as np
import pytest
from step_problem import small_earth_spec

import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu

LMAX = 6


def device_rows(lmax):
    """(block, l, m) of the device layout's rows: per block (spin -2, -1,
    0) the orders m-major (all m for the complex blocks, m >= 0 for spin
    0), l = max(|m|, |N|)..lmax."""
    rows = []
    for b, N in enumerate((-2, -1, 0)):
        for m in range(-lmax if N != 0 else 0, lmax + 1):
            for l in range(max(abs(m), abs(N)), lmax + 1):
                rows.append((b, l, m))
    return rows


def pairing_weights(earth, op):
    """The pairing's weight of every device-layout double ([row][node]
    complex), and the mask of the real field's imaginary m = 0 parts of
    spin 0 (no degree of freedom)."""
    elements, ngll = earth.mesh()
    np_ = ngll - 1
    first, count = (int(v) for v in earth.export("memory_elements"))
    jac = earth.export("jac")
    pw = earth.export("point_weight")
    radius = earth.export("point_radius").reshape(elements, np_)[first: first + count].ravel()
    node = np.array([jac[first + k // np_] * pw[k % np_] for k in range(op.nodes)]) * radius ** 2
    rows = device_rows(op.lmax)
    g = np.array([(1.0, 2.0, 3.0)[b] * (2.0 if b == 2 and m > 0 else 1.0) for b, l, m in rows])
    w = (g[:, None, None] * node[None, :, None]) * np.ones((1, 1, 2))
    dead = np.zeros((len(rows), op.nodes, 2), dtype=bool)
    for k, (b, l, m) in enumerate(rows):
        if b == 2 and m == 0:
            dead[k, :, 1] = True
    assert w.size == op.spectral_doubles
    return w.ravel(), dead.ravel()


@pytest.fixture(scope="module")
def open_step():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    earth = st.Earth(small_earth_spec())
    _, T, _ = earth.scales()
    op = st.Viscoelastic3D(earth, LMAX, None,
                           st.StepOptions.default(max_iterations=60, rtol=1e-12))
    coef = st.delta_coefficients(4, op.nodes, seed=17, rms=0.8, lmax=LMAX)
    op.set_delta(op.delta_field(coef))
    load, _ = st.zonal_load_rate(LMAX, earth)
    dt = 100 * st.YEAR / T
    state = op.zero_state()
    op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    op.start_step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    yield earth, op, state
    op.cancel_step()


def test_transposed_pass_is_the_adjoint_of_the_pass(open_step):
    import torch

    earth, op, _ = open_step
    w, dead = pairing_weights(earth, op)
    S = op.spectral_doubles
    zero = torch.zeros(S, dtype=torch.float64, device="cuda")
    p0 = torch.zeros_like(zero)
    op.step_pass(zero.data_ptr(), p0.data_ptr())
    rng = np.random.default_rng(3)
    worst = 0.0
    for trial in range(4):
        x, u = rng.normal(size=S), rng.normal(size=S)
        x[dead] = 0.0
        u[dead] = 0.0
        xd = torch.tensor(x, device="cuda")
        ud = torch.tensor(u, device="cuda")
        lx = torch.zeros_like(zero)
        tu = torch.zeros_like(zero)
        op.step_pass(xd.data_ptr(), lx.data_ptr())
        op.transposed_pass(ud, tu)
        torch.cuda.synchronize()
        lx = (lx - p0).cpu().numpy()
        tu = tu.cpu().numpy()
        lhs = np.sum(w * lx * u)
        rhs = np.sum(w * x * tu)
        scale = np.sqrt(np.sum(w * lx * lx) * np.sum(w * u * u))
        assert scale > 0.0
        gap = abs(lhs - rhs) / scale
        worst = max(worst, gap)
        assert gap < 1e-12, (trial, lhs, rhs)
    print(f"<L x, u> against <x, T u>: worst {worst:.2e} of |L x| |u|")


def test_the_source_enters_through_its_analysis(open_step):
    import torch

    earth, op, _ = open_step
    S = op.spectral_doubles
    G = op.nodes * op.ntheta * op.nphi
    rng = np.random.default_rng(8)
    src = [torch.tensor(rng.normal(size=n), device="cuda") for n in (2 * G, 2 * G, G)]
    z = torch.tensor(rng.normal(size=S), device="cuda")
    with_src = torch.zeros(S, dtype=torch.float64, device="cuda")
    without = torch.zeros_like(with_src)
    op.transposed_pass(z, with_src, src)
    op.transposed_pass(z, without)
    # the source's analysis through the transforms' own entry points
    sizes = [(LMAX + 1) ** 2 - 4, (LMAX + 1) ** 2 - 1, (LMAX + 1) * (LMAX + 2) // 2]
    parts = []
    for spin, size, field in zip((-2, -1, 0), sizes, src):
        c = torch.zeros(2 * size * op.nodes, dtype=torch.float64, device="cuda")
        op.analyse(spin, field, c)
        parts.append(c)
    reference = torch.cat(parts)
    want = torch.zeros_like(with_src)
    op.from_reference(reference, want)
    torch.cuda.synchronize()
    got = (with_src - without).cpu().numpy()
    want = want.cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
