"""numpy reference of the prefilter MLP (test infrastructure): the forward with
the same bf16 roundings as the device kernels (features, W1, W2 and the layer-1
activations are bf16; accumulation, biases, w3 and the layer-2 activations are
fp32) and the analytic gradient of the class-weighted logistic loss, in
float64. With bf16=False the forward is plain float64 (for the finite
difference check that pins the gradient formulas)."""
import numpy as np

from saber_amd.prefilter import bf16_bits, bf16_float, NPARAMS, W1, B1, W2, B2, W3, B3


def unpack(params):
    p = np.asarray(params, np.float64)
    return (p[W1:B1].reshape(64, 256), p[B1:W2], p[W2:B2].reshape(256, 256), p[B2:W3], p[W3:B3], p[B3])


def rb(a):
    return bf16_float(bf16_bits(np.asarray(a, np.float32))).astype(np.float64)


def forward(params, x, bf16=True):
    w1, b1, w2, b2, w3, b3 = unpack(params)
    if bf16:
        # the device keeps biases and w3 in fp32
        b1, b2, w3, b3 = [np.asarray(np.float32(v), np.float64) for v in (b1, b2, w3, b3)]
    x = np.asarray(x, np.float64)
    if bf16:
        x, w1, w2 = rb(x), rb(w1), rb(w2)
    z1 = x @ w1 + b1
    h1 = np.maximum(z1, 0.0)
    h1b = rb(h1) if bf16 else h1
    z2 = h1b @ w2 + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    return y, (x, w1, w2, w3, h1, h1b, h2)


def loss_and_grad(params, x, labels, pos_weight=1.0, bf16=True):
    y, (xx, w1, w2, w3, h1, h1b, h2) = forward(params, x, bf16)
    t = np.asarray(labels, np.float64)
    n = len(t)
    wt = np.where(t > 0.5, pos_weight, 1.0)
    sp = np.where(y > 0, y + np.log1p(np.exp(-np.abs(y))), np.log1p(np.exp(y)))
    loss = float(np.sum(wt * (sp - t * y)) / n)
    p = 1.0 / (1.0 + np.exp(-y))
    dy = wt * (p - t) / n
    g = np.zeros(NPARAMS, np.float64)
    g[W3:B3] = dy @ h2
    g[B3] = dy.sum()
    dz2 = np.outer(dy, w3) * (h2 > 0)
    g[W2:B2] = (h1b.T @ dz2).ravel()
    g[B2:W3] = dz2.sum(0)
    dz1 = (dz2 @ w2.T) * (h1 > 0)
    g[W1:B1] = (xx.T @ dz1).ravel()
    g[B1:W2] = dz1.sum(0)
    return loss, g
