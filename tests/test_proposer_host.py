"""CPU checks of the adaptive Krylov exponential's host pieces and of the
proposer network's CPU restatement (no GPU).

  * the numpy network (oracle/mlp.py): its gradient against central finite
    differences of its own loss, in float64 (pins the backward formulas);
  * the library's Padé exponential (rfhip_expm_pade, host code) against
    scipy.linalg.expm and the closed forms of diagonal, nilpotent and
    rotation generators;
  * the hindsight labels (rfhip_krylov_labels, host code): every reported
    step is admissible under the scipy estimate, a slightly longer one is
    not, and the chosen dimension maximises h_j / cost(j).
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import mlp as M


def _net(seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=0.3, size=M.N_PARAMS)
    norm = np.concatenate([rng.normal(size=12), rng.uniform(0.5, 2.0, 12), rng.normal(size=2), rng.uniform(0.5, 3.0, 2)])
    return p, norm


def test_oracle_gradient_matches_finite_differences():
    p, norm = _net(1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(7, 12))
    t = rng.normal(size=(7, 2))
    w = rng.uniform(0.2, 2.0, 7)
    loss, g = M.loss_grad(p, norm, x, t, w)
    eps = 1e-6
    idx = np.concatenate([rng.choice(M.N_PARAMS, 60, replace=False), np.arange(M.N_PARAMS - 2, M.N_PARAMS)])
    for q in idx:
        dp = np.zeros_like(p)
        dp[q] = eps
        fd = (M.loss_grad(p + dp, norm, x, t, w)[0] - M.loss_grad(p - dp, norm, x, t, w)[0]) / (2 * eps)
        assert abs(fd - g[q]) <= 1e-7 * max(1.0, abs(fd)), (q, fd, g[q])
    # the loss is the weighted mean squared error of the normalised outputs
    yn = (M.forward(p, norm, x) - norm[24:26]) / norm[26:28]
    tn = (t - norm[24:26]) / norm[26:28]
    assert math.isclose(loss, float(np.sum(w[:, None] * (yn - tn) ** 2) / w.sum()), rel_tol=1e-12)


def test_oracle_forward_of_a_linear_identity_network():
    """With tanh in its linear range, W3 W2 W1 = I-like products pass x through."""
    p = np.zeros(M.N_PARAMS)
    w1, b1, w2, b2, w3, b3 = M.unpack(p)
    s = 1e-4
    w1[0, 3] = s
    w2[0, 0] = 1.0
    w3[1, 0] = 1.0 / s
    b3[0] = 0.5
    norm = np.concatenate([np.zeros(12), np.ones(12), np.zeros(2), np.ones(2)])
    x = np.zeros((3, 12))
    x[:, 3] = [0.1, -0.2, 0.3]
    y = M.forward(p, norm, x)
    assert np.allclose(y[:, 0], 0.5)
    assert np.allclose(y[:, 1], x[:, 3], rtol=1e-6)


def test_pade_exponential_matches_scipy_and_closed_forms():
    from rapidfem_amd.proposer import expm_pade

    rng = np.random.default_rng(3)
    for n, scale in ((1, 0.3), (5, 0.1), (12, 3.0), (30, 20.0), (60, 200.0)):
        a = rng.normal(size=(n, n)) * scale / math.sqrt(n)
        want = expm(a)
        got = expm_pade(a)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want) * max(1.0, scale), (n, scale)
    # upper Hessenberg (the Arnoldi shape), dissipative and oscillatory
    h = np.triu(rng.normal(size=(40, 40)), -1) * 5.0 - 3.0 * np.eye(40)
    assert np.linalg.norm(expm_pade(h) - expm(h)) <= 1e-11 * np.linalg.norm(expm(h))
    # diagonal, nilpotent, rotation
    d = np.diag([-2.0, 0.0, 1.5])
    assert np.allclose(expm_pade(d), np.diag(np.exp([-2.0, 0.0, 1.5])), rtol=1e-14, atol=0)
    nil = np.diag([1.0, 1.0], 1)
    assert np.allclose(expm_pade(nil), np.array([[1, 1, 0.5], [0, 1, 1], [0, 0, 1]]), rtol=0, atol=1e-15)
    th = 7.3
    rot = np.array([[0.0, -th], [th, 0.0]])
    assert np.allclose(expm_pade(rot), [[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]], atol=1e-14)


def _arnoldi(a, v, m):
    """Plain Arnoldi (two Gram-Schmidt passes) of a dense matrix: (m+1, m) H."""
    n = a.shape[0]
    V = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    V[:, 0] = v / np.linalg.norm(v)
    for j in range(m):
        w = a @ V[:, j]
        for _ in range(2):
            c = V[:, :j + 1].T @ w
            H[:j + 1, j] += c
            w -= V[:, :j + 1] @ c
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
    return H


def test_labels_are_admissible_and_maximise_efficiency():
    import ctypes as C

    from rapidfem_amd import _lib as L
    from rapidfem_amd.proposer import labels

    rng = np.random.default_rng(4)
    n, m_log = 300, 24
    # a dissipative, oscillatory test operator: skew part plus a negative diagonal
    s = rng.normal(size=(n, n))
    a = (s - s.T) * 2.0 - np.diag(rng.uniform(0.0, 30.0, n))
    hs, steps = [], []
    for q in range(3):
        v = rng.normal(size=n)
        H = _arnoldi(a, v, m_log)
        hs.append(H)
        steps.append((np.linalg.norm(v), 1e-3 * (q + 1)))
    hess = np.stack(hs)
    rec = np.zeros(3, L.EXP_STEP_DTYPE)
    rec["m"] = m_log
    rec["beta"] = [b for b, _ in steps]
    rec["h"] = [h for _, h in steps]
    rec["t_left"] = 10.0
    rate = 1e-6
    cost = (0.5, 1.0, 0.05)
    grid = (4, 8, 12, 16, 20, 24)
    hb, mb, eff = labels(rec, hess, rate, cost, m_grid=grid, nthreads=2)
    for q in range(3):
        H, beta = hess[q], rec["beta"][q]
        best = None
        for j in grid:
            # recompute h_j independently: admissible at the label, not at 1.01x
            hj = None
            lo, hi = 1e-12, 10.0
            for _ in range(200):
                mid = math.sqrt(lo * hi)
                if M.krylov_estimate(H, j, beta, mid) <= rate * mid:
                    lo = mid
                else:
                    hi = mid
            hj = lo
            e = hj / (cost[0] + cost[1] * j + cost[2] * 0.5 * j * (j + 1))
            if best is None or e > best[0]:
                best = (e, j, hj)
        assert M.krylov_estimate(H, int(mb[q]), beta, hb[q]) <= rate * hb[q] * (1 + 1e-9)
        assert mb[q] == best[1]
        assert abs(hb[q] - best[2]) <= 0.01 * best[2]
        assert abs(eff[q] - best[0]) <= 0.01 * best[0]
