"""Transforms cut with depth (Options::transform_degrees, include/sl3dpp/
GridCoupling.hpp:23-40, src/GridCoupling.cpp:163-190, docs/radial-
partitioning.md section 6): each memory element's nodes are transformed at
the least listed degree (or lmax) at or above the largest degree whose
subset reaches the element, on a grid of that degree; the nodes form
contiguous blocks whose grids every grid field lays end to end.

 * the blocks follow the reference's rule, restated here from the
   operator's own first node per degree;
 * one pass and whole converged steps with a lateral field, against the
   oracle given the same blocks (each node on its block's grid), 1e-11 and
   1e-8;
 * degrees that cut nothing leave the step as it is, bit for bit;
 * with a smooth lateral field the cut step stays close to the full one:
   the nodes' products are exact on their own grids up to the weight's
   components between the grids' degrees (docs/radial-partitioning.md:
   138-158)."""

import bisect

import numpy as np
import pytest
from step_problem import OracleStep, amplitudes, small_earth_spec

import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu

LMAX = 48


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def earth():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    # 12 mantle elements: the deep ones stop short of lmax 48 under the
    # truncation eps 1e-8
    return st.Earth(small_earth_spec(core=4, mantle=12, lithosphere=1))


def reaches(op):
    """Per memory node, the largest degree whose subset reaches it (from
    the operator's first memory node per degree)."""
    tasks = op.export("tasks").reshape(-1, 5)
    first = {int(t[0]): int(t[3]) for t in tasks if int(t[1]) == 0}
    return [max(l for l in range(op.lmax + 1) if first[l] <= node) for node in range(op.nodes)]


@pytest.fixture(scope="module")
def degrees(earth):
    """The transform degrees: the reaches of the three deepest memory
    elements, each then a block of its own."""
    op = make(earth, LMAX, [], 0.0)
    r = reaches(op)
    np_ = op.nodes // 12
    out = sorted({r[0], r[np_], r[2 * np_]})
    assert len(out) == 3 and out[-1] < LMAX
    return out


def make(earth, lmax, degrees, rms, seed=11, field_lmax=6, **kw):
    opts = st.StepOptions.default(anderson_memory=6, anderson_rate=0.5, **kw)
    opts.transform_degrees = degrees
    op = st.Viscoelastic3D(earth, lmax, None, opts)
    if rms != 0.0:
        coef = st.delta_coefficients(field_lmax, op.nodes, seed=seed, rms=rms, lmax=lmax)
        op.set_delta(op.delta_field(coef))
    return op


def test_blocks_follow_the_reference_rule(earth, degrees):
    op = make(earth, LMAX, degrees, 0.0)
    allowed = sorted(set(degrees + [LMAX]))
    want = [allowed[bisect.bisect_left(allowed, reach)] for reach in reaches(op)]
    got = []
    for k in op.blocks:
        got += [k["lmax"]] * k["nodes"]
    assert got == want
    assert [k["lmax"] for k in op.blocks] == allowed  # a block at each depth
    assert op.blocks[-1]["lmax"] == LMAX
    for k in op.blocks:
        nt = int(np.ceil(1.5 * (k["lmax"] + 1) - 1e-9))
        assert (k["ntheta"], k["nphi"]) == (nt, 2 * nt)
    assert op.grid_values == sum(k["nodes"] * k["ntheta"] * k["nphi"] for k in op.blocks)


def run_pair(earth, oracle, degrees, rms, steps, max_iterations, rtol):
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    op = make(earth, LMAX, degrees, rms, max_iterations=max_iterations, rtol=rtol)
    ref = OracleStep(oracle, earth, LMAX, st.EXPONENTIAL_TRAPEZOID, dt, op.delta_grid(),
                     theta=op.theta, wtheta=op.wtheta, blocks=op.blocks)
    assert ref.grid_values == op.grid_values
    load, _ = st.zonal_load_rate(LMAX, earth)
    state = op.zero_state()
    out = []
    for k in range(steps):
        print(f"[cut] degrees {degrees} rms {rms} step {k}", flush=True)
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        amp = amplitudes(load, LMAX, (k + 1) * dt)
        rinfo = ref.step(dt, st.EXPONENTIAL_TRAPEZOID, amp, max_iterations=max_iterations,
                         rtol=rtol, anderson_memory=6, anderson_rate=0.5)
        out.append((info, rinfo, state.export("coefficients"), ref.coef.copy(),
                    [state.export(f"m{b}") for b in range(3)], [m.copy() for m in ref.m],
                    [state.export(f"strain{b}") for b in range(3)],
                    [s.copy() for s in ref.strain]))
    return out


def test_one_pass_with_blocks_matches_oracle(earth, oracle, degrees):
    for info, rinfo, c, rc, m, rm, d, rd in run_pair(earth, oracle, degrees, 0.7, 2, 1, 1e-14):
        assert info.iterations == rinfo.iterations == 1
        assert abs(info.residual - rinfo.residual) <= 1e-10 * abs(rinfo.residual)
        assert rel(c, rc) < 1e-11
        for b in range(3):
            assert rel(m[b], rm[b]) < 1e-11
            assert rel(d[b], rd[b]) < 1e-11


def test_converged_steps_with_blocks_match_oracle(earth, oracle, degrees):
    for info, rinfo, c, rc, m, rm, d, rd in run_pair(earth, oracle, degrees, 1.0, 2, 150,
                                                     1e-10):
        assert info.converged == 1 and rinfo.converged == 1
        assert abs(info.iterations - rinfo.iterations) <= 1
        assert rel(c, rc) < 1e-8
        for b in range(3):
            assert rel(m[b], rm[b]) < 1e-8
            assert rel(d[b], rd[b]) < 1e-8


def trajectory(earth, degrees, rms, steps=2, field_lmax=6):
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    op = make(earth, LMAX, degrees, rms, field_lmax=field_lmax, max_iterations=150,
              rtol=1e-12)
    load, _ = st.zonal_load_rate(LMAX, earth)
    state = op.zero_state()
    for k in range(steps):
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        assert info.converged == 1
    return op, state.export("coefficients")


def test_degrees_that_cut_nothing_leave_the_step_unchanged(earth):
    op0, full = trajectory(earth, [], 0.7)
    op1, same = trajectory(earth, [2, LMAX], 0.7)  # every node reaches past degree 2
    assert len(op0.blocks) == len(op1.blocks) == 1
    assert np.array_equal(full, same)


def test_a_smooth_field_barely_notices_the_cut(earth, degrees):
    """A degree-2 lateral field: the weight's components above degree 2
    come only from exp(-delta)'s own products, tiny on these grids."""
    _, full = trajectory(earth, [], 0.3, field_lmax=2)
    op, cut = trajectory(earth, degrees, 0.3, field_lmax=2)
    assert len(op.blocks) >= 3
    gap = rel(cut, full)
    print(f"cut vs full, degree-2 field rms 0.3: {gap:.2e}")
    assert gap < 1e-6
