"""mlp.py -- CPU restatement (FP64 numpy) of the learned warm start's
network: the multilayer perceptron's forward pass, its mean-squared loss,
backpropagation and one Adam step.

TEST INFRASTRUCTURE ONLY. The HIP network (sl3dpp_amd/csrc/learned.hip) is
compared against these functions; nothing in the product imports them.

The network has no counterpart in the reference (it only supplies the warm
start of OpenStep::Converge, src/Viscoelastic3D.cpp:1916-1921), so nothing
of the reference pins it. The pins are mathematical, in
tests/test_oracle_mlp.py: central finite differences of the loss pin every
gradient; GELU and its derivative are checked against mpmath's erf and
numerical derivative; the first Adam step against its closed form
(w - lr g / (|g| + eps) with the bias corrections); bf16 rounding against
exhaustive tie cases.
"""

import math

import numpy as np

_erf = np.vectorize(math.erf, otypes=[np.float64])


def round_bf16(x):
    """float32 values rounded to bfloat16 (nearest, ties to even), returned
    as float32 (NaN kept as a quiet NaN)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32).reshape(x.shape)


def gelu(z):
    """z Phi(z), Phi the standard normal distribution function."""
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    """Phi(z) + z phi(z)."""
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(weights, biases, x):
    """Layer k: z_k = h_{k-1} W_k^T + b_k; h_k = gelu(z_k) for hidden layers,
    the last z is the output. Returns (y, zs, hs) with hs[0] = x."""
    h = np.asarray(x, dtype=np.float64)
    hs, zs = [h], []
    L = len(weights)
    for k in range(L):
        z = h @ np.asarray(weights[k], dtype=np.float64).T + np.asarray(biases[k], np.float64)
        zs.append(z)
        if k < L - 1:
            h = gelu(z)
            hs.append(h)
        else:
            h = z
    return h, zs, hs


def loss(weights, biases, x, target):
    y, _, _ = forward(weights, biases, x)
    d = y - np.asarray(target, dtype=np.float64)
    return float(np.mean(d * d))


def backward(weights, biases, x, target):
    """(loss, dW list, db list) of mean((y - target)^2)."""
    y, zs, hs = forward(weights, biases, x)
    t = np.asarray(target, dtype=np.float64)
    n = y.size
    d = 2.0 * (y - t) / n
    L = len(weights)
    dW, db = [None] * L, [None] * L
    for k in range(L - 1, -1, -1):
        dW[k] = d.T @ hs[k]
        db[k] = d.sum(axis=0)
        if k > 0:
            d = (d @ np.asarray(weights[k], dtype=np.float64)) * gelu_grad(zs[k - 1])
    return float(np.mean((y - t) ** 2)), dW, db


def adam(w, g, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma and Ba 2015, Algorithm 1) at step number
    `step` (1-based); returns (w, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
