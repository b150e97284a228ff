"""FP64 restatement of the warm-start network -- TEST INFRASTRUCTURE ONLY.

The checker of the MI355X network kernels (ps_mlp.hip): a dense MLP with
GELU hidden layers (tanh form) and a linear output, parameters in the
order W0 [n_in][n_hidden], b0, W1, b1, ..., W_out [n_hidden][n_out], b_out,
mean-squared-error loss over rows x outputs, and Adam (beta 0.9 / 0.999,
eps 1e-8, bias-corrected).  Loaded only by tests/.  Its gradients are pinned
by central finite differences (tests/test_mlp_ref.py).
"""
from __future__ import annotations

import numpy as np

S = np.sqrt(2.0 / np.pi)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(S * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    u = S * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * S * (1.0 + 3.0 * 0.044715 * x * x)


class Mlp:
    def __init__(self, n_in, n_hidden, n_layers, n_out):
        self.dims = [n_in] + [n_hidden] * n_layers + [n_out]
        self.L = n_layers

    def unpack(self, p):
        p = np.asarray(p, np.float64)
        out, o = [], 0
        for k, n in zip(self.dims[:-1], self.dims[1:]):
            W = p[o:o + k * n].reshape(k, n)
            o += k * n
            b = p[o:o + n]
            o += n
            out.append((W, b))
        assert o == len(p)
        return out

    def n_params(self):
        return sum(k * n + n for k, n in zip(self.dims[:-1], self.dims[1:]))

    def forward(self, p, x):
        layers = self.unpack(p)
        h = np.asarray(x, np.float64)
        zs, hs = [], [h]
        for i, (W, b) in enumerate(layers):
            z = h @ W + b
            zs.append(z)
            if i < self.L:
                h = gelu(z)
                hs.append(h)
        return zs[-1], zs, hs

    def loss(self, p, x, t):
        y, _, _ = self.forward(p, x)
        return float(np.mean((y - np.asarray(t, np.float64)) ** 2))

    def loss_grad(self, p, x, t):
        y, zs, hs = self.forward(p, x)
        t = np.asarray(t, np.float64)
        rows, n_out = y.shape
        dz = 2.0 * (y - t) / (rows * n_out)
        layers = self.unpack(p)
        grads = [None] * len(layers)
        for i in range(len(layers) - 1, -1, -1):
            W, _ = layers[i]
            grads[i] = (hs[i].T @ dz, dz.sum(axis=0))
            if i > 0:
                dz = (dz @ W.T) * gelu_grad(zs[i - 1])
        g = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
        return float(np.mean((y - t) ** 2)), g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
