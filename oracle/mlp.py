"""CPU oracle of the learned-probe MLP (test infrastructure only).

Float64 numpy restatement of the network the HIP kernels train and run:
    h1 = GELU(x W1^T + b1)        1024 -> 512
    h2 = GELU(h1 W2^T + b2)       512  -> 256
    o  = h2 W3^T + b3             256  -> 6   (5 nprobe logits + 1 cutoff)
    loss = mean_b CE(softmax(o[:, :5]), y) + mean_b (o[:, 5] - r)^2
with GELU(z) = z * Phi(z) (exact, erf form) and Adam (beta1 0.9, beta2 0.999,
eps 1e-8, bias-corrected).  The reference has no learned component; this is
the oracle the GPU gradients and updates are checked against.
"""
from __future__ import annotations

import math

import numpy as np

NPROBE_CLASSES = (2, 4, 8, 16, 32)
DIMS = (1024, 512, 256, 6)


def _erf(x):
    return np.vectorize(math.erf)(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    cdf = 0.5 * (1.0 + _erf(z / math.sqrt(2.0)))
    pdf = np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)
    return cdf + z * pdf


def init_params(seed: int, dims=DIMS):
    """W ~ U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from a seeded numpy generator, b = 0."""
    rng = np.random.default_rng(seed)
    params = []
    for i in range(len(dims) - 1):
        bound = 1.0 / math.sqrt(dims[i])
        params.append(rng.uniform(-bound, bound, (dims[i + 1], dims[i])).astype(np.float32))
        params.append(np.zeros(dims[i + 1], np.float32))
    return params


def forward(params, x):
    W1, b1, W2, b2, W3, b3 = [np.asarray(p, np.float64) for p in params]
    x = np.asarray(x, np.float64)
    z1 = x @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    o = h2 @ W3.T + b3
    return o, (x, z1, h1, z2, h2)


def loss_and_grads(params, x, y, r):
    """Loss and gradients (float64) for labels y (class index) and targets r."""
    W1, b1, W2, b2, W3, b3 = [np.asarray(p, np.float64) for p in params]
    o, (x, z1, h1, z2, h2) = forward(params, x)
    B = x.shape[0]
    logits = o[:, :5]
    m = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - m)
    p = e / e.sum(axis=1, keepdims=True)
    ce = -np.log(p[np.arange(B), y]).mean()
    diff = o[:, 5] - r
    mse = (diff ** 2).mean()
    do = np.zeros_like(o)
    do[:, :5] = p
    do[np.arange(B), y] -= 1.0
    do[:, :5] /= B
    do[:, 5] = 2.0 * diff / B
    gW3 = do.T @ h2
    gb3 = do.sum(axis=0)
    dz2 = (do @ W3) * gelu_grad(z2)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * gelu_grad(z1)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(axis=0)
    return ce + mse, [gW1, gb1, gW2, gb2, gW3, gb3]


def adam_step(params, grads, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (t counts from 1); returns new params, m, v."""
    out_p, out_m, out_v = [], [], []
    for p, g, mi, vi in zip(params, grads, m, v):
        mi = b1 * mi + (1 - b1) * g
        vi = b2 * vi + (1 - b2) * g * g
        mh = mi / (1 - b1 ** t)
        vh = vi / (1 - b2 ** t)
        out_p.append(p - lr * mh / (np.sqrt(vh) + eps))
        out_m.append(mi)
        out_v.append(vi)
    return out_p, out_m, out_v


def decide(o, conf=0.6, fallback=16):
    """nprobe per query: argmax class if its softmax >= conf, else fallback."""
    logits = np.asarray(o, np.float64)[:, :5]
    e = np.exp(logits - logits.max(axis=1, keepdims=True))
    p = e / e.sum(axis=1, keepdims=True)
    cls = p.argmax(axis=1)
    return np.where(p.max(axis=1) >= conf, np.array(NPROBE_CLASSES)[cls], fallback)
