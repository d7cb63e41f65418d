"""The Legendre kernels' tile-shape tuner: every shape the tuner picks from
(degrees per analysis workgroup 16, 32, 64, 128; colatitudes per
synthesis workgroup 32, 64, 112, 208) gives the same bits, for the transforms and for a
whole converged step, and an operator reports the shape it chose with the
timings that chose it."""

import numpy as np
import pytest
from step_problem import small_earth_spec

import sl3dpp_amd
import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu

SHAPES = [(16, 32), (32, 64), (64, 112), (128, 208)]


@pytest.fixture(scope="module")
def earth():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return st.Earth(small_earth_spec())


def run(earth, lmax, shape):
    import torch

    sl3dpp_amd.tuning_set(*shape)
    try:
        opts = st.StepOptions.default(max_iterations=60, rtol=1e-10, anderson_memory=6)
        op = st.Viscoelastic3D(earth, lmax, None, opts)
    finally:
        sl3dpp_amd.tuning_set(0, 0)
    tuning = op.export("tuning")
    assert tuple(int(v) for v in tuning[:2]) == shape
    coef = st.delta_coefficients(min(lmax, 6), op.nodes, seed=3, rms=1.0, lmax=lmax)
    op.set_delta(op.delta_field(coef))
    out = []
    # the transforms of every spin
    rng = np.random.default_rng(lmax)
    for spin in (-2, -1, 0):
        real = spin == 0
        size = (lmax + 1) * (lmax + 2) // 2 if real else (lmax + 1) ** 2 - spin * spin
        c = rng.normal(size=(op.nodes, size, 2))
        if real:
            for l in range(lmax + 1):
                c[:, l * (l + 1) // 2, 1] = 0.0
        cdev = torch.from_numpy(c.ravel().copy()).cuda()
        field = torch.zeros(op.nodes * op.ntheta * op.nphi * (1 if real else 2),
                            dtype=torch.float64, device="cuda")
        op.synthesise(spin, cdev, field)
        back = torch.zeros_like(cdev)
        op.analyse(spin, field, back)
        out += [field.cpu().numpy(), back.cpu().numpy()]
    # two converged steps (the pass's truncated transforms and the start's)
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    load, _ = st.zonal_load_rate(lmax, earth)
    state = op.zero_state()
    for k in range(2):
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        assert info.converged == 1
        out.append(state.export("coefficients"))
        out += [state.export(f"m{b}") for b in range(3)]
    return out, tuning


# lmax 20: one block of each shape; lmax 140: several degree and colatitude
# blocks (ntheta 212)
@pytest.mark.parametrize("lmax", [20, 140])
def test_every_tile_shape_gives_the_same_bits(earth, lmax):
    base = None
    for shape in SHAPES:
        print(f"[tuner] lmax {lmax} shape {shape}", flush=True)
        out, _ = run(earth, lmax, shape)
        if base is None:
            base = out
            continue
        for a, b in zip(base, out):
            assert np.array_equal(a, b)


def test_tuner_reports_its_choice(earth):
    sl3dpp_amd.tuning_set(0, 0)
    op = st.Viscoelastic3D(earth, 24, None, st.StepOptions.default())
    t = op.export("tuning")
    rows, theta = (16, 32, 64, 128), (32, 64, 112, 208)
    assert len(t) == 10
    assert int(t[0]) in rows and int(t[1]) in theta
    assert np.all(t[2:] > 0)
    # the chosen shapes are the fastest timed
    assert t[2 + rows.index(int(t[0]))] == np.min(t[2:6])
    assert t[6 + theta.index(int(t[1]))] == np.min(t[6:10])
