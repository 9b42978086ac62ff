"""FP64 restatement of the selector MLP (forward, loss, analytic gradient).

TEST INFRASTRUCTURE ONLY: the checker for the device kernels in gary_amd/csrc/selector.hip
(mlp_forward_kernel, mlp_backward_kernel, mlp_reduce_kernel). Pinned against central finite
differences of its own loss (tests/test_selector.py). The network has no counterpart in the
reference.

Layout (GARY_MLP_PARAMS floats): W1[32][6] b1[32] W2[32][32] b2[32] W3[3][32] b3[3]; ReLU after
the two hidden layers; loss = mean over samples of 0.5 * Σ_o mask[o] * (y_o - t_o)^2.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 6, 32, 3
SIZES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
         ("W3", (OUT, HID)), ("b3", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], np.float64).reshape(shape)
        o += k
    return out


def forward(p, X):
    w = unpack(p)
    X = np.asarray(X, np.float64)
    a1 = X @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ w["W3"].T + w["b3"]
    return y, (X, a1, h1, a2, h2)


def loss_grad(p, X, T, M):
    w = unpack(p)
    y, (X, a1, h1, a2, h2) = forward(p, X)
    n = X.shape[0]
    r = (y - T) * M
    loss = 0.5 * float((r * r).sum()) / n
    dy = r / n
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    d2 = (dy @ w["W3"]) * (a2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ w["W2"]) * (a1 > 0)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SIZES])
