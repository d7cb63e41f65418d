"""The learned warm start of the step's fixed point.

There is no counterpart in the reference: the prediction is handed to the
step as the `start` of OpenStep::Converge (src/Viscoelastic3D.cpp:1916-1921),
and the fixed point then iterates to the same tolerance, so the result is the
plain step's to that tolerance.

A multilayer perceptron (hidden layers of `width`, GELU) acts on every memory
node alike (shared weights, one row per node):
  * input row: the log-viscosity field's spin-0 coefficients of degree
    <= lin at the node (re, im) and (context) their mean over the column of
    memory nodes, since the degree systems couple every node; the
    radial-reference response's rows of
    degree <= lrad at the node (the step's converged strain with delta = 0,
    spins -2, -1, 0), and the node's radius;
  * output row: the lateral correction at the node, the converged strain
    minus the radial reference over the rows of degree <= lcut.
The warm start is the radial reference plus the predicted correction, and
the step drops it (the guard) when its first pass changes it by more than the
cold start's second pass typically does (the median over the training steps).

Training triples (delta, radial response, converged correction) are made by
the solver itself, one first step from rest per delta sample. Every kernel of
the network (BF16 products on the matrix cores, FP32 accumulation, the loss,
backpropagation and Adam) and of the feature moves is the library's own.
"""

import ctypes
import time

import numpy as np

from . import check, lib

c_int_p = ctypes.POINTER(ctypes.c_int)
c_float_p = ctypes.POINTER(ctypes.c_float)

lib.sl3_mlp_create.argtypes = [ctypes.c_int, c_int_p, ctypes.c_int, ctypes.c_uint64,
                               ctypes.POINTER(ctypes.c_void_p)]
lib.sl3_mlp_destroy.argtypes = [ctypes.c_void_p]
lib.sl3_mlp_forward.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                ctypes.c_void_p]
lib.sl3_mlp_backward.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
lib.sl3_mlp_adam.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double, ctypes.c_double]
lib.sl3_mlp_export.argtypes = [ctypes.c_void_p, ctypes.c_char_p, c_float_p,
                               ctypes.POINTER(ctypes.c_int64)]
lib.sl3_mlp_import.argtypes = [ctypes.c_void_p, ctypes.c_char_p, c_float_p, ctypes.c_int64]
lib.sl3_mlp_set_chunk.argtypes = [ctypes.c_void_p, ctypes.c_int]
lib.sl3_column_moments.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
lib.sl3_column_standardise.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
lib.sl3_spectral_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
lib.sl3_spectral_gather.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
lib.sl3_spectral_scatter.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]
lib.sl3_state_spectral.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _sync():
    import torch

    torch.cuda.synchronize()
    check(lib.sl3_synchronize())


class Mlp:
    """sl3_mlp: widths[0] -> ... -> widths[-1], GELU between layers, BF16
    products with FP32 accumulation, Adam. Tensors are torch float32 on the
    device, row-major."""

    def __init__(self, widths, max_rows, seed=0):
        self.widths = [int(w) for w in widths]
        self.max_rows = int(max_rows)
        arr = (ctypes.c_int * len(self.widths))(*self.widths)
        self.handle = ctypes.c_void_p()
        check(lib.sl3_mlp_create(len(self.widths) - 1, arr, self.max_rows, seed,
                                 ctypes.byref(self.handle)))

    def __del__(self, _lib=lib):  # the default keeps the library through shutdown
        if getattr(self, "handle", None):
            _lib.sl3_mlp_destroy(self.handle)
            self.handle = None

    @property
    def layers(self):
        return len(self.widths) - 1

    def forward(self, x, y):
        _sync()
        check(lib.sl3_mlp_forward(self.handle, x.shape[0], _ptr(x), _ptr(y)))
        _sync()

    def backward(self, rows, x, target, want_loss=True, row0=0):
        """Gradients of mean((y - target)^2) over rows row0..row0+rows of
        x and target (already synchronised with torch)."""
        loss = ctypes.c_double()
        xo = ctypes.c_void_p(x.data_ptr() + 4 * row0 * x.shape[1])
        to = ctypes.c_void_p(target.data_ptr() + 4 * row0 * target.shape[1])
        check(lib.sl3_mlp_backward(self.handle, rows, xo, to,
                                   ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib.sl3_mlp_adam(self.handle, lr, beta1, beta2, eps))

    def export(self, name):
        count = ctypes.c_int64()
        check(lib.sl3_mlp_export(self.handle, name.encode(), None, ctypes.byref(count)))
        out = np.zeros(count.value, dtype=np.float32)
        check(lib.sl3_mlp_export(self.handle, name.encode(), out.ctypes.data_as(c_float_p),
                                 ctypes.byref(count)))
        return out

    def set_chunk(self, chunk):
        """0: per-shape tuning (default); 32 or 128: a fixed slice depth."""
        check(lib.sl3_mlp_set_chunk(self.handle, chunk))

    def chunks(self):
        """The tuner's choices: [(kind, M, N, K, depth), ...]."""
        v = self.export("chunks").astype(int)
        return [tuple(v[i:i + 5]) for i in range(0, v.size, 5)]

    def import_(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float32).ravel()
        check(lib.sl3_mlp_import(self.handle, name.encode(), v.ctypes.data_as(c_float_p),
                                 v.size))


def column_moments(x, floor=1e-6):
    """(mean, scale) per column of a rows x cols float32 device tensor."""
    import torch

    mean = torch.empty(x.shape[1], dtype=torch.float32, device=x.device)
    scale = torch.empty_like(mean)
    _sync()
    check(lib.sl3_column_moments(x.shape[0], x.shape[1], _ptr(x), floor, _ptr(mean),
                                 _ptr(scale)))
    _sync()
    return mean, scale


def standardise(x, mean, scale, inverse=False):
    _sync()
    check(lib.sl3_column_standardise(x.shape[0], x.shape[1], _ptr(x), _ptr(mean), _ptr(scale),
                                     1 if inverse else 0))
    _sync()


def spectral_rows(op, cut):
    n = ctypes.c_int64()
    check(lib.sl3_spectral_rows(op.handle, cut, ctypes.byref(n)))
    return n.value


def state_spectral(state):
    """The device pointer of a state's strain coefficients (device layout)."""
    p = ctypes.c_void_p()
    check(lib.sl3_state_spectral(state.handle, ctypes.byref(p)))
    if not p.value:
        raise ValueError("the state has no strain coefficients yet")
    return p


def gather(op, spectral_ptr, base, cut, out, ld, col0=0, mean=None, scale=None):
    _sync()
    check(lib.sl3_spectral_gather(op.handle, spectral_ptr, _ptr(base), cut, _ptr(mean),
                                  _ptr(scale), _ptr(out), ld, col0))


def scatter(op, values, ld, cut, base, out, mean=None, scale=None):
    _sync()
    check(lib.sl3_spectral_scatter(op.handle, _ptr(values), ld, cut, _ptr(mean), _ptr(scale),
                                   _ptr(base), _ptr(out)))


class WarmStart:
    """The learned warm start of one operator's first steps from rest under a
    fixed load, scheme and dt (the radial reference is theirs). Computing
    the radial reference steps the operator with delta = 0 and leaves it
    there: set the lateral field again (set_delta) before stepping. One
    thread at a time per instance, like the operator it drives."""

    def __init__(self, op, dt, scheme, load, lin=32, lrad=8, lcut=32, width=512, layers=4,
                 max_rows=1024, seed=7, context=True):
        import torch

        self.op, self.dt, self.scheme, self.load = op, dt, scheme, load
        self.lin, self.lrad, self.lcut = min(lin, op.lmax), min(lrad, op.lmax), min(lcut, op.lmax)
        self.nodes = op.nodes
        self.n_delta = (self.lin + 1) * (self.lin + 2) // 2
        self.rows_rad = spectral_rows(op, self.lrad)
        self.rows_out = spectral_rows(op, self.lcut)
        self.context = bool(context)
        self.f_in = (4 if self.context else 2) * self.n_delta + 2 * self.rows_rad + 1
        self.f_out = 2 * self.rows_out
        self.mlp = Mlp([self.f_in] + [width] * (layers - 1) + [self.f_out], max_rows, seed)
        self.guard = 0.0
        # the radial reference: the converged first step with delta = 0
        op.set_delta(None)
        ref = op.zero_state()
        info = op.step(ref, dt, scheme, load)
        self.radial_passes = info.iterations
        self.xrad = torch.from_numpy(ref.export("coefficients_device")).cuda()
        _sync()
        # the static part of a node's row: radial response rows, radius
        rad = torch.zeros(self.nodes, 2 * self.rows_rad, dtype=torch.float32, device="cuda")
        gather(op, _ptr(self.xrad), None, self.lrad, rad, 2 * self.rows_rad)
        _sync()
        radius = np.asarray(op.radii, dtype=np.float32) / float(np.max(op.radii))
        self.static = np.concatenate([rad.cpu().numpy(), radius[:, None]], axis=1)
        self.in_mean = self.in_scale = self.out_mean = self.out_scale = None
        self.history = []

    # ---- features
    def features(self, coefficients):
        """Input rows (nodes x f_in, float32 on the device) for per-node
        spin-0 delta coefficients (nodes x (L+1)(L+2)/2 complex)."""
        import torch

        c = np.asarray(coefficients)[:, :self.n_delta]
        d = np.empty((self.nodes, 2 * self.n_delta), dtype=np.float32)
        d[:, 0::2] = c.real
        d[:, 1::2] = c.imag
        parts = [d]
        if self.context:
            # the column's mean field: the degree systems couple every node
            parts.append(np.broadcast_to(d.mean(axis=0), d.shape))
        rows = np.concatenate(parts + [self.static], axis=1)
        return torch.from_numpy(np.ascontiguousarray(rows)).cuda()

    def set_delta(self, coefficients):
        self.op.set_delta(self.op.delta_field(coefficients))

    # ---- triples made by the solver
    def make_triples(self, coefficient_fn, seeds, log=None):
        """(inputs, targets, cold infos): per seed, one first step from rest
        with delta from coefficient_fn(seed), converged from the step's own
        start; the targets are the converged correction rows (raw)."""
        import torch

        n = len(seeds)
        X = torch.empty(n * self.nodes, self.f_in, dtype=torch.float32, device="cuda")
        T = torch.empty(n * self.nodes, self.f_out, dtype=torch.float32, device="cuda")
        infos = []
        t0 = time.perf_counter()
        for i, seed in enumerate(seeds):
            coef = coefficient_fn(seed)
            self.set_delta(coef)
            state = self.op.zero_state()
            info = self.op.step(state, self.dt, self.scheme, self.load)
            res = self.op.export("residuals")
            infos.append((info.iterations, float(res[1]) if res.size > 1 else float(res[0])))
            rows = self.features(coef)
            _sync()
            X[i * self.nodes:(i + 1) * self.nodes].copy_(rows)
            tview = T[i * self.nodes:(i + 1) * self.nodes]
            gather(self.op, state_spectral(state), self.xrad, self.lcut, tview, self.f_out)
            _sync()
            if log and (i + 1) % max(1, n // 8) == 0:
                log(f"triples: {i + 1}/{n}, {time.perf_counter() - t0:.1f} s")
        return X, T, infos

    # ---- training
    def fit(self, X, T, infos, epochs=50, lr=1e-3, batch=None, log=None, seed=0):
        """Standardises X and T in place (moments kept for prediction), then
        trains with Adam over minibatches of whole rows; the guard is the
        median of the training steps' second cold residual."""
        batch = batch or self.mlp.max_rows
        self.in_mean, self.in_scale = column_moments(X)
        self.out_mean, self.out_scale = column_moments(T)
        standardise(X, self.in_mean, self.in_scale)
        standardise(T, self.out_mean, self.out_scale)
        self.guard = float(np.median([r for _, r in infos]))
        rows = X.shape[0]
        starts = list(range(0, rows, batch))
        rng = np.random.default_rng(seed)
        _sync()
        t0 = time.perf_counter()
        for epoch in range(epochs):
            total, count = 0.0, 0
            for b in rng.permutation(len(starts)):
                r0 = starts[b]
                nb = min(batch, rows - r0)
                loss = self.mlp.backward(nb, X, T, True, r0)
                self.mlp.adam(lr)
                total += loss * nb
                count += nb
            self.history.append(total / count)
            if log and (epoch % 10 == 0 or epoch == epochs - 1):
                log(f"epoch {epoch}: loss {self.history[-1]:.4e} "
                    f"({time.perf_counter() - t0:.1f} s)")
        _sync()
        return self.history

    # ---- prediction
    def predict_rows(self, coefficients):
        """The predicted correction rows (nodes x f_out, raw units)."""
        import torch

        x = self.features(coefficients)
        standardise(x, self.in_mean, self.in_scale)
        y = torch.empty(self.nodes, self.f_out, dtype=torch.float32, device="cuda")
        self.mlp.forward(x, y)
        standardise(y, self.out_mean, self.out_scale, inverse=True)
        return y

    def start(self, coefficients):
        """The warm start (device layout, float64): the radial reference plus
        the predicted correction."""
        import torch

        x = self.features(coefficients)
        standardise(x, self.in_mean, self.in_scale)
        y = torch.empty(self.nodes, self.f_out, dtype=torch.float32, device="cuda")
        self.mlp.forward(x, y)
        out = torch.empty_like(self.xrad)
        scatter(self.op, y, self.f_out, self.lcut, self.xrad, out, self.out_mean,
                self.out_scale)
        _sync()
        return out

    def step(self, state, coefficients, guard=None):
        """A first step with the learned warm start (guarded)."""
        start = self.start(coefficients)
        return self.op.step(state, self.dt, self.scheme, self.load, start=start,
                            guard=self.guard if guard is None else guard)

    def correction_error(self, state, coefficients):
        """The network's relative error on the correction rows against a
        converged state's: |pred - true| / |true|, and |true| / |xrad rows|."""
        import torch

        true = torch.empty(self.nodes, self.f_out, dtype=torch.float32, device="cuda")
        gather(self.op, state_spectral(state), self.xrad, self.lcut, true, self.f_out)
        pred = self.predict_rows(coefficients)
        _sync()
        t = true.double().cpu().numpy()
        p = pred.double().cpu().numpy()
        return float(np.linalg.norm(p - t) / max(np.linalg.norm(t), 1e-300))


def evaluate(ws, coefficient_fn, seeds, log=None, radial=True):
    """Per held-out seed: the cold step (the plain fixed point from its own
    start), the radial-reference start (no network) and the learned start,
    each a first step from rest to the same tolerance; passes, wall time
    (the learned one with its inference), the network's error, fallbacks."""
    out = []
    op = ws.op
    for seed in seeds:
        coef = coefficient_fn(seed)
        ws.set_delta(coef)
        rec = {"seed": int(seed)}
        state = op.zero_state()
        _sync()
        t0 = time.perf_counter()
        info = op.step(state, ws.dt, ws.scheme, ws.load)
        rec["cold_s"] = time.perf_counter() - t0
        rec["cold_passes"] = info.iterations
        rec["cold_converged"] = bool(info.converged)
        cold = state.export("coefficients")
        rec["error"] = ws.correction_error(state, coef)
        if radial:
            state = op.zero_state()
            info = op.step(state, ws.dt, ws.scheme, ws.load, start=ws.xrad)
            rec["radial_passes"] = info.iterations
            rec["radial_first_residual"] = float(op.export("residuals")[0])
        state = op.zero_state()
        _sync()
        t0 = time.perf_counter()
        info = ws.step(state, coef)
        rec["learned_s"] = time.perf_counter() - t0
        rec["learned_passes"] = info.iterations
        rec["learned_converged"] = bool(info.converged)
        rec["fallback"] = bool(info.warm_fallback)
        rec["learned_first_residual"] = float(op.export("residuals")[0])
        learned = state.export("coefficients")
        rec["state_difference"] = float(np.max(np.abs(learned - cold)) /
                                        max(np.max(np.abs(cold)), 1e-300))
        out.append(rec)
        if log:
            log(f"seed {seed}: cold {rec['cold_passes']} passes {1e3 * rec['cold_s']:.1f} ms, "
                + (f"radial {rec['radial_passes']}, " if radial else "")
                + f"learned {rec['learned_passes']} {1e3 * rec['learned_s']:.1f} ms (first residual "
                f"{rec['learned_first_residual']:.2e}, fallback {rec['fallback']}), network error "
                f"{rec['error']:.3f}")
    return out


def summary(records):
    """Aggregates of evaluate(): mean passes, steps/s, fallback rate."""
    n = len(records)
    cold_s = sum(r["cold_s"] for r in records)
    learned_s = sum(r["learned_s"] for r in records)
    s = {
        "samples": n,
        "cold_passes": float(np.mean([r["cold_passes"] for r in records])),
        "learned_passes": float(np.mean([r["learned_passes"] for r in records])),
        "cold_steps_per_s": n / cold_s,
        "learned_steps_per_s": n / learned_s,
        "speedup": cold_s / learned_s,
        "fallback_rate": float(np.mean([r["fallback"] for r in records])),
        "network_error": float(np.mean([r["error"] for r in records])),
        "max_state_difference": float(max(r["state_difference"] for r in records)),
    }
    if "radial_passes" in records[0]:
        s["radial_passes"] = float(np.mean([r["radial_passes"] for r in records]))
    return s


def delta_sampler(op, lmax_field=64, rms=0.7, ar=0.9):
    """coefficient_fn(seed) for the 3D step configurations' delta field."""
    from . import step as st

    return lambda seed: st.delta_coefficients(min(lmax_field, op.lmax), op.nodes, seed=seed,
                                              rms=rms, ar=ar, lmax=op.lmax)


__all__ = ["Mlp", "WarmStart", "evaluate", "summary", "delta_sampler", "column_moments",
           "standardise", "spectral_rows", "gather", "scatter", "state_spectral"]
