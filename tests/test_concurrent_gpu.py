"""Independent realisations stepped from different host threads: each thread
runs the library's work on its own stream, so the operators' steps overlap
on the device. Their results are those of the same steps taken one after
the other (bitwise: the same kernels on the same data), and the handles can
be made on one thread and stepped on another."""

import threading

import numpy as np
import pytest
from step_problem import small_earth_spec

import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu

LMAX = 16


@pytest.fixture(scope="module")
def earth():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return st.Earth(small_earth_spec())


def realisation(earth, seed):
    opts = st.StepOptions.default(max_iterations=60, rtol=1e-10, anderson_memory=6)
    op = st.Viscoelastic3D(earth, LMAX, None, opts)
    coef = st.delta_coefficients(6, op.nodes, seed=seed, rms=1.0, lmax=LMAX)
    op.set_delta(op.delta_field(coef))
    return op


def run(op, load, dt, steps):
    state = op.zero_state()
    infos = []
    for _ in range(steps):
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        infos.append((info.iterations, info.converged, info.residual))
    return infos, state.export("coefficients"), [state.export(f"m{b}") for b in range(3)]


def test_concurrent_realisations_match_sequential(earth):
    L, T, M = earth.scales()
    dt = 100 * st.YEAR / T
    load, _ = st.zonal_load_rate(LMAX, earth)
    seeds = [3, 5, 7, 9]
    ops = [realisation(earth, s) for s in seeds]
    print("[concurrent] sequential", flush=True)
    want = [run(op, load, dt, 3) for op in ops]
    got = [None] * len(ops)
    errors = []

    def work(i):
        try:
            got[i] = run(ops[i], load, dt, 3)
        except Exception as e:  # reported below, on the main thread
            errors.append(e)

    print("[concurrent] threads", flush=True)
    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(ops))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for i in range(len(ops)):
        assert got[i] is not None
        assert got[i][0] == want[i][0]  # the same passes and residuals
        assert np.array_equal(got[i][1], want[i][1])
        for b in range(3):
            assert np.array_equal(got[i][2][b], want[i][2][b])
    # the realisations differ: the test would not see a mix-up otherwise
    assert not np.array_equal(want[0][1], want[1][1])
