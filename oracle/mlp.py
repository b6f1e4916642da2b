"""CPU restatement (numpy, float64) of the (h, m) proposer's network and of the
adaptive Krylov exponential's controller, for checking the library's HIP
kernels and host logic.

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by the product
package. Neither piece has a counterpart in the reference (its exponential
route takes a fixed step and dimension, crates/rapidfem-td/src/
session.rs:700-726); the network follows the parameter layout and loss
declared in include/rfhip.h (rfhip_mlp_*), the estimate follows
propagator.rs:278-294.

Parity pins: forward/backward against central finite differences of the
loss (tests/test_mlp_oracle.py), the estimate against a direct dense
exponential (scipy.linalg.expm) of the Hessenberg matrix.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 12, 64, 2
N_PARAMS = N_H * N_IN + N_H + N_H * N_H + N_H + N_OUT * N_H + N_OUT


def unpack(p: np.ndarray):
    """W1 (64,12), b1, W2 (64,64), b2, W3 (2,64), b3 from the flat block."""
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((N_H, N_IN), (N_H,), (N_H, N_H), (N_H,), (N_OUT, N_H), (N_OUT,)):
        k = int(np.prod(shape))
        out.append(p[o:o + k].reshape(shape))
        o += k
    assert o == N_PARAMS
    return out


def split_norm(norm: np.ndarray):
    norm = np.asarray(norm, np.float64)
    return norm[:N_IN], norm[N_IN:2 * N_IN], norm[2 * N_IN:2 * N_IN + N_OUT], norm[2 * N_IN + N_OUT:]


def forward(p, norm, x):
    """Raw outputs y = yn * out_scale + out_mean of rows x (n, 12)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    im, iinv, om, osc = split_norm(norm)
    xn = (np.asarray(x, np.float64) - im) * iinv
    h1 = np.tanh(xn @ w1.T + b1)
    h2 = np.tanh(h1 @ w2.T + b2)
    yn = h2 @ w3.T + b3
    return yn * osc + om


def loss_grad(p, norm, x, t, w=None):
    """sum_s w_s |yn_s - tn_s|^2 / sum_s w_s and its gradient (flat, float64)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    im, iinv, om, osc = split_norm(norm)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    n = x.shape[0]
    w = np.ones(n) if w is None else np.asarray(w, np.float64)
    ws = w / w.sum()
    xn = (x - im) * iinv
    h1 = np.tanh(xn @ w1.T + b1)
    h2 = np.tanh(h1 @ w2.T + b2)
    yn = h2 @ w3.T + b3
    tn = (t - om) / osc
    e = yn - tn
    loss = float(np.sum(ws[:, None] * e * e))
    d3 = 2.0 * ws[:, None] * e                      # (n, 2)
    g_w3 = d3.T @ h2
    g_b3 = d3.sum(0)
    d2 = (d3 @ w3) * (1.0 - h2 * h2)                # (n, 64)
    g_w2 = d2.T @ h1
    g_b2 = d2.sum(0)
    d1 = (d2 @ w2) * (1.0 - h1 * h1)
    g_w1 = d1.T @ xn
    g_b1 = d1.sum(0)
    g = np.concatenate([g_w1.ravel(), g_b1, g_w2.ravel(), g_b2, g_w3.ravel(), g_b3])
    return loss, g


def krylov_estimate(hess: np.ndarray, j: int, beta: float, h: float) -> float:
    """beta * H[j][j-1] * |[exp(h H_j)]_{j-1,0}| (propagator.rs:278-294,
    scaled by the propagated norm), by scipy's dense exponential."""
    from scipy.linalg import expm

    hn = float(hess[j, j - 1])
    if hn == 0.0:
        return 0.0
    e = expm(h * np.asarray(hess[:j, :j], np.float64))
    return beta * hn * abs(float(e[j - 1, 0]))
