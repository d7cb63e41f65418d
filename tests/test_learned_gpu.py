"""The learned warm start on the GPU:
 * the network's BF16 matrix-core kernels against the FP64 oracle
   (oracle/mlp.py) on the same BF16-rounded weights and inputs: the forward
   pass, every gradient of the mean-squared loss, and Adam's update; ragged
   widths and row counts exercise the tile edges. Tolerances: the device
   rounds hidden activations and output gradients to BF16 (relative 2^-9
   each), so the forward output is held to 2e-2 and the gradients to 4e-2
   of their norms, Adam to 1e-5;
 * training lowers the loss of a teacher network's targets;
 * the spectral gather and scatter move exactly the rows of degree <= cut;
 * the guarded step: a warm start above the guard is dropped and the step
   converges from its own start to the plain step's state; the exact
   solution as the start closes in one pass;
 * the trained warm start: the same converged states as the plain fixed
   point, and what the network gains over the start without it."""

import os
import sys

import numpy as np
import pytest

import sl3dpp_amd.step as st

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp as omlp  # noqa: E402


def norm_rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) /
                 max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def learned():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    import sl3dpp_amd.learned as ln

    return ln


def _network(ln, widths, rows, seed):
    import torch

    net = ln.Mlp(widths, max_rows=rows, seed=seed)
    W = [omlp.round_bf16(net.export(f"weight:{k}").reshape(widths[k], widths[k - 1]))
         for k in range(1, len(widths))]
    b = [net.export(f"bias:{k}") for k in range(1, len(widths))]
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(rows, widths[0])).astype(np.float32)
    t = rng.normal(size=(rows, widths[-1])).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(t).cuda()
    return net, W, b, omlp.round_bf16(x), t, xd, td


@pytest.mark.parametrize("widths,rows", [([40, 64, 48, 24], 100), ([130, 200, 72, 33], 257),
                                         ([512, 512, 512, 512, 256], 1024)])
def test_forward_and_gradients_against_oracle(learned, widths, rows):
    import torch

    ln = learned
    net, W, b, x, t, xd, td = _network(ln, widths, rows, seed=5)
    y = torch.empty(rows, widths[-1], dtype=torch.float32, device="cuda")
    net.forward(xd, y)
    want, _, _ = omlp.forward(W, b, x)
    assert norm_rel(y.cpu().numpy(), want) < 2e-2
    ln._sync()
    loss = net.backward(rows, xd, td)
    ref_loss, dW, db = omlp.backward(W, b, x, t)
    assert abs(loss - ref_loss) <= 2e-2 * ref_loss
    for k in range(1, len(widths)):
        gw = net.export(f"grad_weight:{k}").reshape(widths[k], widths[k - 1])
        gb = net.export(f"grad_bias:{k}")
        assert norm_rel(gw, dW[k - 1]) < 4e-2, k
        assert norm_rel(gb, db[k - 1]) < 4e-2, k


def test_slice_depths_give_the_same_bits(learned):
    """The products staged 32 and 128 deep, and the tuner's per-shape
    choice, sum in the same order: outputs and gradients agree bit for
    bit."""
    import torch

    ln = learned
    widths, rows = [130, 200, 72, 33], 257
    results = []
    for chunk in (32, 128, 0):
        net, W, b, x, t, xd, td = _network(ln, widths, rows, seed=5)
        net.set_chunk(chunk)
        y = torch.empty(rows, widths[-1], dtype=torch.float32, device="cuda")
        net.forward(xd, y)
        ln._sync()
        net.backward(rows, xd, td, want_loss=False)
        grads = [net.export(f"grad_weight:{k}") for k in range(1, len(widths))]
        grads += [net.export(f"grad_bias:{k}") for k in range(1, len(widths))]
        results.append((y.cpu().numpy(), grads))
        if chunk == 0:
            chosen = net.chunks()
            assert chosen and all(c[4] in (32, 128) for c in chosen)
    for y, grads in results[1:]:
        np.testing.assert_array_equal(y, results[0][0])
        for g, g0 in zip(grads, results[0][1]):
            np.testing.assert_array_equal(g, g0)


def test_adam_update_against_oracle(learned):
    ln = learned
    widths = [24, 40, 16]
    net, W, b, x, t, xd, td = _network(ln, widths, 50, seed=9)
    w0 = [net.export(f"weight:{k}").astype(np.float64) for k in (1, 2)]
    m = [np.zeros_like(w) for w in w0]
    v = [np.zeros_like(w) for w in w0]
    for step in (1, 2, 3):
        ln._sync()
        net.backward(50, xd, td, want_loss=False)
        g = [net.export(f"grad_weight:{k}").astype(np.float64) for k in (1, 2)]
        net.adam(lr=1e-2)
        for i in range(2):
            w0[i], m[i], v[i] = omlp.adam(w0[i], g[i], m[i], v[i], step, lr=1e-2)
        for i, k in enumerate((1, 2)):
            got = net.export(f"weight:{k}")
            np.testing.assert_allclose(got, w0[i], rtol=1e-5, atol=1e-6)
        # the next gradients start from the device's own FP32 weights
        w0 = [net.export(f"weight:{k}").astype(np.float64) for k in (1, 2)]


def test_training_lowers_the_loss(learned):
    import torch

    ln = learned
    rng = np.random.default_rng(1)
    teacher = ln.Mlp([32, 64, 8], max_rows=4096, seed=3)
    x = torch.from_numpy(rng.normal(size=(4096, 32)).astype(np.float32)).cuda()
    t = torch.empty(4096, 8, dtype=torch.float32, device="cuda")
    teacher.forward(x, t)
    net = ln.Mlp([32, 128, 128, 8], max_rows=256, seed=4)
    ln._sync()
    first = np.mean([net.backward(256, x, t, True, r0) for r0 in range(0, 4096, 256)])
    for epoch in range(30):
        for r0 in range(0, 4096, 256):
            net.backward(256, x, t, False, r0)
            net.adam(lr=3e-3)
    last = np.mean([net.backward(256, x, t, True, r0) for r0 in range(0, 4096, 256)])
    assert last < 0.05 * first, (first, last)


@pytest.fixture(scope="module")
def small_problem(learned):
    from step_problem import small_earth_spec

    earth = st.Earth(small_earth_spec())
    _, tscale, _ = earth.scales()
    dt = 100.0 * st.YEAR / tscale
    opts = st.StepOptions.default(max_iterations=200, rtol=1e-9, anderson_memory=6,
                                  anderson_rate=0.5)
    lmax = 16
    op = st.Viscoelastic3D(earth, lmax, None, opts)
    load, _ = st.zonal_load_rate(lmax, earth)
    return earth, op, dt, load


def test_gather_scatter_rows(learned, small_problem):
    import torch

    ln = learned
    _, op, dt, load = small_problem
    coef = st.delta_coefficients(8, op.nodes, seed=3, rms=0.7, lmax=op.lmax)
    op.set_delta(op.delta_field(coef))
    state = op.zero_state()
    op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    x = torch.from_numpy(state.export("coefficients_device")).cuda()
    base = torch.from_numpy(np.random.default_rng(2).normal(size=x.numel())).cuda()
    # every row at the full cut: the gather is a permutation of the tensor
    full = ln.spectral_rows(op, op.lmax)
    assert 2 * full * op.nodes == op.spectral_doubles
    cut = 6
    rows = ln.spectral_rows(op, cut)
    feats = torch.empty(op.nodes, 2 * rows, dtype=torch.float32, device="cuda")
    ln.gather(op, ln._ptr(x), base, cut, feats, 2 * rows)
    out = torch.empty_like(x)
    ln.scatter(op, feats, 2 * rows, cut, base, out)
    ln._sync()
    o, xb, bb = out.cpu().numpy(), x.cpu().numpy(), base.cpu().numpy()
    moved = np.abs(o - bb) > 0
    # the moved entries are x (to float32 rounding of x - base), the others base
    np.testing.assert_allclose(o[moved], xb[moved], rtol=0, atol=1e-6 * np.max(np.abs(xb - bb)))
    assert moved.sum() <= 2 * rows * op.nodes
    # the rows of degree <= cut are 2 (cut + 1)^2 - 5 + (cut + 1)(cut + 2) / 2 per node
    assert rows == 2 * (cut + 1) ** 2 - 5 + (cut + 1) * (cut + 2) // 2


def test_guarded_step_falls_back_to_the_plain_step(learned, small_problem):
    import torch

    _, op, dt, load = small_problem
    coef = st.delta_coefficients(8, op.nodes, seed=4, rms=0.7, lmax=op.lmax)
    op.set_delta(op.delta_field(coef))
    state = op.zero_state()
    cold = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    x_cold = state.export("coefficients")
    exact = torch.from_numpy(state.export("coefficients_device")).cuda()
    # a start that is wrong everywhere, under a guard it exceeds: dropped
    wrong = exact * 3.0
    state = op.zero_state()
    info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load, start=wrong, guard=1e-3)
    assert info.warm_fallback == 1
    # the dropped pass counts; the cold sequence after it may end a pass
    # either side of the plain step's (its first pass recomputes the start
    # force the plain step takes from StartStep, equal to rounding)
    assert info.converged and abs(info.iterations - (cold.iterations + 1)) <= 1
    x = state.export("coefficients")
    assert np.max(np.abs(x - x_cold)) <= 1e-7 * np.max(np.abs(x_cold))
    # the converged state as the start: one pass, kept
    state = op.zero_state()
    info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load, start=exact, guard=1e-3)
    assert info.warm_fallback == 0 and info.iterations <= 2 and info.converged


def test_learned_warm_start_end_to_end(learned, small_problem):
    ln = learned
    _, op, dt, load = small_problem
    ws = ln.WarmStart(op, dt, st.EXPONENTIAL_TRAPEZOID, load, lin=8, lrad=4, lcut=8,
                      width=256, layers=4, max_rows=512, seed=3)
    sample = ln.delta_sampler(op, lmax_field=8)
    X, T, infos = ws.make_triples(sample, list(range(100, 124)))
    history = ws.fit(X, T, infos, epochs=40, lr=1e-3, batch=64)
    assert history[-1] < 0.5 * history[0]
    records = ln.evaluate(ws, sample, [900, 901, 902])
    s = ln.summary(records)
    print("learned warm start (lmax 16):", s)
    for r in records:
        # the same converged state as the plain fixed point (rtol 1e-9)
        assert r["cold_converged"] and r["learned_converged"]
        assert r["state_difference"] < 1e-6
    # the network predicts the held-out correction better than none at all
    assert s["network_error"] < 1.0
    # and its start beats the radial reference's (the start without it)
    assert np.mean([r["learned_first_residual"] for r in records]) < \
        np.mean([r["radial_first_residual"] for r in records])
