"""FP64 oracle of the 3-64-64-3 SiLU colour surrogate.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  No reference
implementation exists (the reference has no learned colour path); this is the
textbook multilayer perceptron with SiLU activations (Elfwing et al. 2018,
x * sigmoid(x)), mean-squared-error loss, reverse-mode gradients and Adam
(Kingma & Ba 2015).

Numerics of the backend it checks: weights, biases and the two hidden
activations are rounded to bfloat16 (round to nearest even) in the forward
pass -- what the MFMA inference kernel consumes -- and gradients pass through
the rounding unchanged (straight-through).  ``quant=False`` gives the plain
FP64 network, whose gradients the finite-difference check pins.

Parameter vector (4,611 values, the layout the kernels use):
  W1[3][64] @0, b1[64] @192, W2[64][64] @256, b2[64] @4352, W3[64][3] @4416,
  b3[3] @4608.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 3, 64, 3
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 192, 256, 4352, 4416, 4608
N_PARAMS = 4611


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(N_IN, N_HID), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(N_HID, N_HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(N_HID, N_OUT), p[OFF_B3:N_PARAMS])


def pack(w1, b1, w2, b2, w3, b3):
    return np.concatenate([np.ravel(w1), b1, np.ravel(w2), b2, np.ravel(w3), b3]).astype(np.float64)


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def sigmoid(a):
    return 0.5 * (1.0 + np.tanh(0.5 * a))


def silu(a):
    return a * sigmoid(a)


def silu_grad(a):
    s = sigmoid(a)
    return s * (1.0 + a * (1.0 - s))


def init_params(seed: int = 0):
    """He-style uniform initialisation for tests that start the GPU trainer
    and this oracle from the same parameters."""
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (N_IN, N_HID)) * np.sqrt(6.0 / N_IN)
    w2 = rng.uniform(-1, 1, (N_HID, N_HID)) * np.sqrt(6.0 / N_HID)
    w3 = rng.uniform(-1, 1, (N_HID, N_OUT)) * np.sqrt(6.0 / N_HID) * 0.1
    return pack(w1, np.zeros(N_HID), w2, np.zeros(N_HID), w3, np.full(N_OUT, 0.5))


def forward(p, x, quant: bool = True):
    """Returns y and the intermediates (a1, h1, a2, h2) for ``backward``."""
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float64))
    w1, b1, w2, b2, w3, b3 = (q(t) for t in unpack(p))
    x = np.asarray(x, dtype=np.float64)
    a1 = x @ w1 + b1
    h1 = q(silu(a1))
    a2 = h1 @ w2 + b2
    h2 = q(silu(a2))
    y = h2 @ w3 + b3
    return y, (x, a1, h1, a2, h2)


def loss_and_grad(p, x, t, quant: bool = True):
    """Mean squared error over all samples and channels and its gradient
    with respect to the parameter vector."""
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float64))
    w1, b1, w2, b2, w3, b3 = (q(t_) for t_ in unpack(p))
    y, (x, a1, h1, a2, h2) = forward(p, x, quant)
    n = x.shape[0]
    r = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(r * r))
    dy = 2.0 * r / (n * N_OUT)
    dw3 = h2.T @ dy
    db3 = dy.sum(0)
    da2 = (dy @ w3.T) * silu_grad(a2)
    dw2 = h1.T @ da2
    db2 = da2.sum(0)
    da1 = (da2 @ w2.T) * silu_grad(a1)
    dw1 = x.T @ da1
    db1 = da1.sum(0)
    return loss, pack(dw1, db1, dw2, db2, dw3, db3)


def adam_step(p, g, m, v, step: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam update (bias-corrected), step counted from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
