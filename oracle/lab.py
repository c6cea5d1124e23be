"""FP64 restatement of the learned-boundary networks and their gradient.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
training kernels (k_lab_fwdbwd / k_lab_adam) and inference kernel; never
by the product.

One net maps a 61-feature column record x to 5 ghost values:
y = Ws x + W3 tanh(W2 tanh(W1 x + b1) + b2) + b3, with the flat parameter
layout W1[32][61] b1[32] W2[32][32] b2[32] W3[5][32] b3[5] Ws[5][61].
The loss of a batch is the mean square error over the samples and the
first nout outputs.  The gradient is pinned against central finite
differences in tests/test_lab_oracle.py.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NO = 61, 32, 5
REC = 72


def offsets():
    o = {"W1": 0}
    o["b1"] = o["W1"] + NH * NIN
    o["W2"] = o["b1"] + NH
    o["b2"] = o["W2"] + NH * NH
    o["W3"] = o["b2"] + NH
    o["b3"] = o["W3"] + NO * NH
    o["Ws"] = o["b3"] + NO
    o["n"] = o["Ws"] + NO * NIN
    return o


NP = offsets()["n"]


def unpack(p):
    o = offsets()
    return (p[o["W1"]:o["b1"]].reshape(NH, NIN), p[o["b1"]:o["W2"]], p[o["W2"]:o["b2"]].reshape(NH, NH),
            p[o["b2"]:o["W3"]], p[o["W3"]:o["b3"]].reshape(NO, NH), p[o["b3"]:o["Ws"]],
            p[o["Ws"]:o["n"]].reshape(NO, NIN))


def bf16(a):
    """Nearest-even BF16 of FP32 values, as FP64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(a))


def forward(p, X, hidden_bf16=False):
    """y[n, 5] in FP64 for features X[n, 61].  hidden_bf16 rounds the
    inputs of both hidden layers to BF16 (the time step's matrix-core
    form; the training loss and gradient below are the plain form)."""
    W1, b1, W2, b2, W3, b3, Ws = (np.asarray(a, np.float64) for a in unpack(np.asarray(p, np.float64)))
    X = np.asarray(X, np.float64)
    Xh = bf16(X) if hidden_bf16 else X
    h1 = np.tanh(Xh @ W1.T + b1)
    if hidden_bf16:
        h1 = bf16(h1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3.T + b3 + X @ Ws.T


def loss_grad(p, X, T, nout):
    """(loss, grad[NP]) of the mean square error over n samples and the
    first nout outputs, by reverse-mode differentiation in FP64."""
    p = np.asarray(p, np.float64)
    W1, b1, W2, b2, W3, b3, Ws = unpack(p)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    n = len(X)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ W3.T + b3 + X @ Ws.T
    e = y - T
    e[:, nout:] = 0.0
    loss = float((e ** 2).sum() / (n * nout))
    d3 = 2.0 * e / (n * nout)
    d2 = (d3 @ W3) * (1.0 - h2 ** 2)
    d1 = (d2 @ W2) * (1.0 - h1 ** 2)
    g = np.zeros(NP)
    gW1, gb1, gW2, gb2, gW3, gb3, gWs = unpack(g)
    gW1[...] = d1.T @ X
    gb1[...] = d1.sum(0)
    gW2[...] = d2.T @ h1
    gb2[...] = d2.sum(0)
    gW3[...] = d3.T @ h2
    gb3[...] = d3.sum(0)
    gWs[...] = d3.T @ X
    return loss, g


def loss_only(p, X, T, nout):
    y = forward(p, X)
    e = (y - np.asarray(T, np.float64))[:, :nout]
    return float((e ** 2).sum() / (len(X) * nout))
