"""Learned (h, m) proposer of the adaptive Krylov exponential.

No reference counterpart: the reference's exponential route propagates each
frame with a fixed step and Krylov dimension (crates/rapidfem-td/src/
session.rs:700-726, propagator.rs:186-326). Here every step's (h, m) is
either chosen by the plain error-controlled controller or proposed by a
12 -> 64 -> 64 -> 2 perceptron, and judged by the a-posteriori estimate of
propagator.rs:278-294 either way (GpuOperator.expmv_adaptive).

The network runs and trains on the GPU through the library's own HIP
kernels (rfhip_mlp_*). Training data are steps the plain controller logged
with a wide basis (m_log columns); their labels are the hindsight-best
(h, m) under a cost model of one step: for each grid dimension j the largest
step the estimate admits, h_j, and the j with the most time advanced per
unit of cost, h_j / cost(j) (rfhip_krylov_labels).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib as L

N_FEATURES = 12
N_PARAMS = 5122
N_NORM = 28
M_GRID = (10, 12, 14, 16, 18, 20, 24, 28, 32, 36, 40, 44, 48, 52, 56, 60)


@dataclass
class ExpController:
    """rfhip_expctl. tol: relative error target of the whole run; m: the plain
    controller's dimension (the session's krylov_dim default, session.rs:208);
    [m_min, m_max]: the proposer's range."""

    tol: float = 1e-6
    m: int = 40
    m_min: int = 10
    m_max: int = 60
    h0: float = 0.0
    safety: float = 0.9
    growth_limit: float = 5.0
    shrink_limit: float = 0.2
    max_steps: int = 1_000_000
    h_scale: float = 0.0  # the proposer's unit of time (the CFL step); 0: 1/|H[1][0]| of the first basis

    def c(self) -> L.ExpCtlC:
        return L.ExpCtlC(float(self.tol), int(self.m), int(self.m_min), int(self.m_max), float(self.h0),
                         float(self.safety), float(self.growth_limit), float(self.shrink_limit), int(self.max_steps),
                         float(self.h_scale))


def expm_pade(a) -> np.ndarray:
    """exp(a), degree-13 Padé with scaling and squaring (the adaptive path's
    Hessenberg exponential), computed by the library on the host."""
    a = np.ascontiguousarray(a, np.float64)
    out = np.empty_like(a)
    L.check(L.lib.rfhip_expm_pade(a, int(a.shape[0]), out))
    return out


def features_and_targets(steps: np.ndarray, h_best: np.ndarray, m_best: np.ndarray, h_scale: float):
    """Rows (features, (log(h* / h_scale), m*)) of the logged steps that have
    a proposal (no breakdown within two columns) and a positive label; h_scale
    is the run's ExpController.h_scale."""
    keep = (steps["m"] >= 2) & (h_best > 0) & ~((steps["breakdown"] == 1) & (steps["m"] <= 2))
    x = steps["feat"][keep].astype(np.float32)
    t = np.stack([np.log(h_best[keep] / h_scale), m_best[keep].astype(np.float64)], axis=1).astype(np.float32)
    return x, t


def labels(steps: np.ndarray, hess: np.ndarray, rate: float, cost_model, m_grid=M_GRID, nthreads: int = 16):
    """Hindsight-best (h, m) of each logged step (rfhip_krylov_labels).
    steps/hess: what expmv_adaptive returned for a run with hess_dim = m_log;
    rate: tol * |v| / t_end of that run; cost_model: (c0, c1, c2) of
    cost(j) = c0 + c1 j + c2 j (j+1) / 2."""
    n = len(steps)
    m_log = hess.shape[2]
    grid = np.array([j for j in m_grid if j <= m_log], np.int32)
    hb = np.zeros(n)
    mb = np.zeros(n, np.int32)
    eff = np.zeros(n)
    hessc = np.ascontiguousarray(hess, np.float64)
    L.check(L.lib.rfhip_krylov_labels(hessc, int(m_log), int(n), np.ascontiguousarray(steps["m"], np.int32),
                                      np.ascontiguousarray(steps["beta"]), np.full(n, float(rate)),
                                      np.ascontiguousarray(steps["h"]), np.ascontiguousarray(steps["t_left"]),
                                      grid, int(len(grid)), np.asarray(cost_model, np.float64), int(nthreads),
                                      hb, mb, eff.ctypes.data_as(C.c_void_p)))
    return hb, mb, eff


class Proposer:
    """The perceptron (rfhip_mlp): FP32 parameters on the GPU."""

    def __init__(self, ctx, seed: int = 0):
        self._ctx = ctx  # keeps the context alive
        h = C.c_void_p()
        L.check(L.lib.rfhip_mlp_create(ctx._h, C.c_uint64(seed), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib.rfhip_mlp_destroy(h)
            self._h = None

    def params(self) -> tuple[np.ndarray, np.ndarray]:
        p = np.empty(N_PARAMS, np.float32)
        nm = np.empty(N_NORM, np.float32)
        L.check(L.lib.rfhip_mlp_get_params(self._h, p.ctypes.data_as(C.c_void_p), nm.ctypes.data_as(C.c_void_p)))
        return p, nm

    def set_params(self, params=None, norm=None) -> None:
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        nm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        if p is not None and p.shape != (N_PARAMS,):
            raise ValueError(f"expected {N_PARAMS} parameters, got {p.shape}")
        if nm is not None and nm.shape != (N_NORM,):
            raise ValueError(f"expected {N_NORM} normalisation entries, got {nm.shape}")
        L.check(L.lib.rfhip_mlp_set_params(self._h, None if p is None else p.ctypes.data_as(C.c_void_p),
                                           None if nm is None else nm.ctypes.data_as(C.c_void_p)))

    def fit_normalisation(self, x: np.ndarray, t: np.ndarray) -> None:
        """Feature and target standardisation from training data."""
        xm, xs = x.mean(0), x.std(0)
        tm, ts = t.mean(0), t.std(0)
        norm = np.concatenate([xm, 1.0 / np.where(xs > 1e-6, xs, 1.0), tm, np.where(ts > 1e-6, ts, 1.0)])
        self.set_params(norm=norm.astype(np.float32))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_FEATURES)
        y = np.empty((x.shape[0], 2), np.float32)
        L.check(L.lib.rfhip_mlp_forward(self._h, x, int(x.shape[0]), y))
        return y

    def loss_grad(self, x, t, w=None) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_FEATURES)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        w = None if w is None else np.ascontiguousarray(w, np.float32)
        loss = C.c_double()
        g = np.empty(N_PARAMS, np.float32)
        L.check(L.lib.rfhip_mlp_loss_grad(self._h, x, t, None if w is None else w.ctypes.data_as(C.c_void_p),
                                          int(x.shape[0]), C.byref(loss), g.ctypes.data_as(C.c_void_p)))
        return loss.value, g

    def train(self, x, t, w=None, epochs: int = 50, batch: int = 256, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_FEATURES)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        w = None if w is None else np.ascontiguousarray(w, np.float32)
        hist = np.zeros(epochs)
        L.check(L.lib.rfhip_mlp_train(self._h, x, t, None if w is None else w.ctypes.data_as(C.c_void_p),
                                      int(x.shape[0]), int(epochs), int(batch), C.c_float(lr), C.c_uint64(seed),
                                      hist.ctypes.data_as(C.c_void_p)))
        return hist

    def save(self, path: str) -> None:
        p, nm = self.params()
        np.savez(path, params=p, norm=nm)

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"], z["norm"])


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "proposer_weights.npz")
