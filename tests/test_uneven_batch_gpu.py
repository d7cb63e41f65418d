"""An uneven batch of independent 3D steps (the uneven-batch configuration at
test size): realisations of different lmax (8 to 24) on Earths of different
radial meshes (12, 16 and 20 elements x 6 GLL nodes), each with its own
lateral field, stepped at once by three host threads (one library stream
each, the operators' work overlapping on the device), two converged steps
each (rtol 1e-10). Every realisation's state is compared with the oracle's
converged steps on the same inputs: strain coefficients, memory and strain
grids to 1e-8 relative (the converged-step bar of test_step_gpu.py), and
the pass counts within one of the oracle's. The per-realisation errors are
printed."""

import queue
import threading

import numpy as np
import pytest
from step_problem import OracleStep, amplitudes, small_earth_spec

import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu

MESHES = {12: (6, 5, 1), 16: (8, 7, 1), 20: (10, 9, 1)}
# (lmax, elements, delta seed): the lmax sizes hit the matrix-core ring
# kernels with R = 4 (8, 16, 24) and R = 2 (10, 12: nphi 34 = 2 x 17 and
# 40 = 4 x 10, the generic kernel)
CASES = [(8, 12, 21), (16, 20, 22), (24, 16, 23), (12, 12, 24), (10, 20, 25), (16, 16, 26)]
STEPS = 2
RTOL = 1e-10


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def earths():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return {e: st.Earth(small_earth_spec(*MESHES[e], ngll=6)) for e in MESHES}


def make(earth, lmax, seed):
    opts = st.StepOptions.default(max_iterations=150, rtol=RTOL, anderson_memory=6,
                                  anderson_rate=0.5)
    op = st.Viscoelastic3D(earth, lmax, None, opts)
    coef = st.delta_coefficients(min(lmax, 6), op.nodes, seed=seed, rms=1.0, lmax=lmax)
    op.set_delta(op.delta_field(coef))
    return op


def test_uneven_batch_matches_the_oracle_per_realisation(earths, oracle):
    cases = []
    for lmax, elements, seed in CASES:
        earth = earths[elements]
        _, T, _ = earth.scales()
        op = make(earth, lmax, seed)
        load, _ = st.zonal_load_rate(lmax, earth)
        cases.append({"lmax": lmax, "elements": elements, "earth": earth, "op": op,
                      "state": op.zero_state(), "dt": 100 * st.YEAR / T, "load": load,
                      "infos": []})
    work = queue.Queue()
    for i in range(len(cases)):
        work.put(i)
    errors = []

    def worker():
        while True:
            try:
                i = work.get_nowait()
            except queue.Empty:
                return
            c = cases[i]
            try:
                for _ in range(STEPS):
                    c["infos"].append(c["op"].step(c["state"], c["dt"],
                                                   st.EXPONENTIAL_TRAPEZOID, c["load"]))
            except Exception as e:  # reported on the main thread
                errors.append(e)

    print(f"[uneven] {len(cases)} realisations on 3 threads", flush=True)
    threads = [threading.Thread(target=worker) for _ in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not errors, errors
    for c in cases:
        lmax, op = c["lmax"], c["op"]
        ref = OracleStep(oracle, c["earth"], lmax, st.EXPONENTIAL_TRAPEZOID, c["dt"],
                         op.delta_grid(), theta=op.theta, wtheta=op.wtheta)
        passes = []
        for k in range(STEPS):
            amp = amplitudes(c["load"], lmax, (k + 1) * c["dt"])
            passes.append(ref.step(c["dt"], st.EXPONENTIAL_TRAPEZOID, amp, max_iterations=150,
                                   rtol=RTOL, anderson_memory=6, anderson_rate=0.5).iterations)
        gaps = [rel(c["state"].export("coefficients"), ref.coef)]
        gaps += [rel(c["state"].export(f"m{b}"), ref.m[b]) for b in range(3)]
        gaps += [rel(c["state"].export(f"strain{b}"), ref.strain[b]) for b in range(3)]
        mine = [info.iterations for info in c["infos"]]
        print(f"[uneven] lmax {lmax}, {c['elements']} elements: passes {mine} (oracle "
              f"{passes}), worst relative gap {max(gaps):.2e}", flush=True)
        assert all(info.converged for info in c["infos"])
        assert all(abs(a - b) <= 1 for a, b in zip(mine, passes))
        assert max(gaps) < 1e-8
