"""CPU restatement of the surrogate network -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's cpu_baseline leg import this module; the product
(ph_brush_amd/) never does.  No reference counterpart: the reference has no
learned part.  This module restates, in numpy FP64, what the device kernels
compute:

    x  = bf16(d * in_scale)                      d: 48 shell-count differences
    z1 = x @ bf16(W1) + b1          h1 = bf16(gelu(z1))
    z2 = h1 @ bf16(W2) + b2         h2 = bf16(gelu(z2))
    y  = h2 @ bf16(w3) + b3
    dE_star = out_mean + out_std * y

gelu(z) = z/2 (1 + erf(z/sqrt 2)).  The loss is the mean of (y - t_hat)^2 with
t_hat = (t - out_mean) / out_std.  The backward pass rounds the two matrix
operands the kernels round (dz2 for dW2 and dh1, dz1 for dW1) when `bf16` is
on; with `bf16=False` every step is plain FP64 and the gradient is the exact
gradient of the FP64 loss (pinned by central finite differences in the tests).
Adam follows Kingma & Ba (2015), Algorithm 1, with bias correction.

Flat parameter layout (float32 on the device):
    W1[48][128] (input-major), b1[128], W2[128][128] (input-major), b2[128],
    w3[128], b3  -> 22,913 trainable values, then in_scale[48], out_mean, out_std.
"""
from __future__ import annotations

import math

import numpy as np

NIN, NH = 48, 128
NTRAIN = NIN * NH + NH + NH * NH + NH + NH + 1          # 22,913
NPARAM = NTRAIN + NIN + 2                              # + in_scale, out_mean, out_std

_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest, ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    o = 0
    W1 = t[o:o + NIN * NH].reshape(NIN, NH); o += NIN * NH
    b1 = t[o:o + NH]; o += NH
    W2 = t[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = t[o:o + NH]; o += NH
    w3 = t[o:o + NH]; o += NH
    b3 = t[o]; o += 1
    in_scale = t[o:o + NIN]; o += NIN
    out_mean, out_std = t[o], t[o + 1]
    return W1, b1, W2, b2, w3, b3, in_scale, out_mean, out_std


def init_params(seed: int) -> np.ndarray:
    """He-uniform style initial weights (float32), zero biases, unit scales."""
    rng = np.random.default_rng(seed)
    th = np.zeros(NPARAM, dtype=np.float32)
    W1 = rng.uniform(-1, 1, (NIN, NH)) * math.sqrt(6.0 / NIN)
    W2 = rng.uniform(-1, 1, (NH, NH)) * math.sqrt(6.0 / NH)
    w3 = rng.uniform(-1, 1, NH) * math.sqrt(6.0 / NH) * 0.1
    o = 0
    th[o:o + NIN * NH] = W1.ravel(); o += NIN * NH + NH
    th[o:o + NH * NH] = W2.ravel(); o += NH * NH + NH
    th[o:o + NH] = w3; o += NH + 1
    th[o:o + NIN] = 1.0; o += NIN
    th[o], th[o + 1] = 0.0, 1.0
    return th


def forward(theta, d, bf=True):
    """Returns (dE_star, cache).  d: (B, 48) descriptor rows."""
    W1, b1, W2, b2, w3, b3, sc, om, osd = unpack(theta)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    x = r(np.asarray(d, dtype=np.float64) * sc)
    z1 = x @ r(W1) + b1
    h1 = r(gelu(z1))
    z2 = h1 @ r(W2) + b2
    h2 = r(gelu(z2))
    y = h2 @ r(w3) + b3
    return om + osd * y, (x, z1, h1, z2, h2, y)


def loss_and_grad(theta, d, t, bf=True):
    """Mean-squared loss on normalised targets and its gradient w.r.t. the
    22,913 trainable values (the kernels' definition when bf=True)."""
    W1, b1, W2, b2, w3, b3, sc, om, osd = unpack(theta)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    _, (x, z1, h1, z2, h2, y) = forward(theta, d, bf)
    B = len(y)
    th = (np.asarray(t, dtype=np.float64) - om) / osd
    e = y - th
    loss = float(np.mean(e * e))
    gy = 2.0 * e / B
    dw3 = h2.T @ gy
    db3 = gy.sum()
    dz2 = np.outer(gy, r(w3)) * gelu_grad(z2)
    dz2r = r(dz2)
    dW2 = h1.T @ dz2r
    db2 = dz2.sum(0)
    dh1 = dz2r @ r(W2).T
    dz1 = dh1 * gelu_grad(z1)
    dz1r = r(dz1)
    dW1 = x.T @ dz1r
    db1 = dz1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])
    return loss, g


def adam_step(theta, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update of the trainable part (step counts from 1)."""
    th = np.array(theta, dtype=np.float64)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    th[:NTRAIN] -= lr * mh / (np.sqrt(vh) + eps)
    return th, m, v
