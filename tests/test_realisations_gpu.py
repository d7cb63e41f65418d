"""Realisations side by side (sl3_step_options.realisations): R independent
lateral fields on one Earth and degree stepped by one operator, every
kernel of a pass taking their memory nodes as extra columns. Each
realisation converges on its own residual and Anderson history, so it
takes the passes it takes alone: its StepInfo and its state equal those of
an operator holding that realisation alone, bit for bit, through a
trajectory with a change of dt, the tide switched on, a backward Euler and
an Sdirk2 step, realisations that need different numbers of passes (one of
them no lateral field), and a warm start. No reference counterpart: the
reference steps realisations one operator at a time."""

import numpy as np
import pytest
from step_problem import small_earth_spec, tidal_rate

import sl3dpp_amd as sl3
import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu

RMS = [0.7, 0.0, 1.3]  # a field, none, a strong one: different pass counts


@pytest.fixture(scope="module")
def earth():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return st.Earth(small_earth_spec())


def options(realisations=1, memory=6, damping=1.0):
    return st.StepOptions.default(max_iterations=80, rtol=1e-10, anderson_memory=memory,
                                  anderson_rate=0.5, damping=damping,
                                  realisations=realisations)


def alone(earth, lmax, rms, seed, **kw):
    op = st.Viscoelastic3D(earth, lmax, None, options(**kw))
    if rms != 0.0:
        coef = st.delta_coefficients(min(lmax, 6), op.nodes, seed=seed, rms=rms, lmax=lmax)
        op.set_delta(op.delta_field(coef))
    return op


def side_by_side(earth, lmax, singles, **kw):
    import torch

    R = len(singles)
    op = st.Viscoelastic3D(earth, lmax, None, options(realisations=R, **kw))
    assert op.realisations == R and op.radial_nodes == singles[0].nodes
    assert op.nodes == R * singles[0].nodes
    fields = []
    for s in singles:
        d = s.delta if s.delta is not None else torch.zeros(s.grid_values, dtype=torch.float64,
                                                            device="cuda")
        fields.append(d)
    op.set_delta(torch.cat(fields))
    assert op.grid_values == R * singles[0].grid_values
    return op


def snapshot(state):
    out = {"coefficients_device": state.export("coefficients_device")}
    for b in range(3):
        out[f"m{b}"] = state.export(f"m{b}")
        out[f"strain{b}"] = state.export(f"strain{b}")
    return out


def part(name, value, r, R, N):
    """Realisation r's share of a batched export."""
    if name == "coefficients_device":  # [row][node] complex, every block alike
        return value.reshape(-1, R * N, 2)[:, r * N:(r + 1) * N, :].ravel()
    return value.reshape(R * N, -1)[r * N:(r + 1) * N].ravel()  # [node][grid]


def plan(earth, lmax):
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    tide = tidal_rate(lmax, earth)
    return [  # (dt, scheme, tide)
        (dt, st.EXPONENTIAL_TRAPEZOID, None),
        (dt, st.EXPONENTIAL_TRAPEZOID, None),
        (0.5 * dt, st.EXPONENTIAL_TRAPEZOID, tide),
        (dt, st.BACKWARD_EULER, tide),
        (dt, st.SDIRK2, None),
    ]


def run(earth, op, lmax, warm=True):
    import torch

    load, _ = st.zonal_load_rate(lmax, earth)
    state = op.zero_state()
    infos, snaps = [], []
    for h, scheme, td in plan(earth, lmax):
        info = op.step(state, h, scheme, load, tide=td)
        infos.append((info, op.realisation_infos()))
        snaps.append(snapshot(state))
    if warm:
        L, T, M = earth.scales()
        start = torch.from_numpy(state.export("coefficients_device")).cuda()
        info = op.step(state, 100 * st.YEAR / T, st.EXPONENTIAL_TRAPEZOID, load, start=start)
        infos.append((info, op.realisation_infos()))
        snaps.append(snapshot(state))
    return infos, snaps


def same_info(a, b):
    assert a.converged == b.converged
    assert a.iterations == b.iterations
    assert a.mixed_passes == b.mixed_passes
    assert a.residual == b.residual


# lmax 8: the matrix-core ring kernels; lmax 12: the generic ring kernel
@pytest.mark.parametrize("lmax", [8, 12])
def test_realisations_side_by_side_equal_each_alone(earth, lmax):
    R = len(RMS)
    print(f"[realisations] lmax {lmax} R {R}", flush=True)
    singles = [alone(earth, lmax, rms, 40 + r) for r, rms in enumerate(RMS)]
    N = singles[0].nodes
    batch = side_by_side(earth, lmax, singles)
    b_infos, b_snaps = run(earth, batch, lmax)
    passes = []
    for r, single in enumerate(singles):
        s_infos, s_snaps = run(earth, single, lmax)
        counts = []
        for (bi, per), (si, own) in zip(b_infos, s_infos):
            assert len(per) == R and len(own) == 1
            same_info(per[r], si)
            same_info(own[0], si)
            counts.append(si.iterations)
        passes.append(counts)
        for bs, ss in zip(b_snaps, s_snaps):
            for name, value in ss.items():
                assert np.array_equal(part(name, bs[name], r, R, N), value), (r, name)
    # the batch's own info: the most passes, the largest residual, all converged
    for k, (bi, per) in enumerate(b_infos):
        assert bi.converged == 1
        assert bi.iterations == max(p.iterations for p in per)
        assert bi.residual == max(p.residual for p in per)
    print(f"passes per realisation: {passes}")
    assert len({tuple(p) for p in passes}) > 1  # the realisations take different passes


def test_damped_realisations_equal_each_alone(earth):
    """Damping: the sweeps a damped pass takes after its residual."""
    lmax = 8
    R = 2
    singles = [alone(earth, lmax, rms, 60 + r, damping=0.9) for r, rms in enumerate([0.7, 1.3])]
    N = singles[0].nodes
    batch = side_by_side(earth, lmax, singles, damping=0.9)
    b_infos, b_snaps = run(earth, batch, lmax, warm=False)
    for r, single in enumerate(singles):
        s_infos, s_snaps = run(earth, single, lmax, warm=False)
        for (bi, per), (si, own) in zip(b_infos, s_infos):
            same_info(per[r], si)
        for bs, ss in zip(b_snaps, s_snaps):
            for name, value in ss.items():
                assert np.array_equal(part(name, bs[name], r, R, N), value), (r, name)


def test_realisation_options_are_checked(earth):
    with pytest.raises(sl3.InvalidArgument, match="realisations"):
        st.Viscoelastic3D(earth, 8, None, options(realisations=0))
    with pytest.raises(sl3.InvalidArgument, match="realisations"):
        st.Viscoelastic3D(earth, 8, None, options(realisations=65))
    o = options(realisations=2)
    o.transform_degrees = [4]
    with pytest.raises(sl3.InvalidArgument, match="transform degrees"):
        st.Viscoelastic3D(earth, 8, None, o)
    o = options(realisations=2)
    o.solver = st.CONJUGATE_GRADIENTS
    with pytest.raises(sl3.InvalidArgument, match="fixed point"):
        st.Viscoelastic3D(earth, 8, None, o)
    op = st.Viscoelastic3D(earth, 8, None, options(realisations=2))
    with pytest.raises(sl3.InvalidArgument, match="fixed at construction"):
        op.set_options(realisations=3)
