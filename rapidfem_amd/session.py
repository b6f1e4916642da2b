"""TdSession GPU routes (crates/rapidfem-td/src/session.rs) over GpuOperator.

Only the device paths of the reference session are mirrored: explicit LSERK4
(session.rs:885-927), adaptive KCL (:940-970) and exponential Krylov
(:867-874), plus cfl_dt (:514-569), run on the GPU operator in FP64.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .gpu import Controller, Drive, GpuContext, GpuOperator, KclStats
from .maxwell import MaxwellOperator

# constants.rs:58 KRYLOV_TOL and the CFL bracket constants (constants.rs:231-262).
KRYLOV_TOL = 1e-10
CFL_POWER_ITERS = 40
CFL_PROBE_STEPS = 64
CFL_BISECT_ITERS = 7
CFL_Z_STABLE = 3.0
CFL_Z_UNSTABLE = 15.0
CFL_GROWTH_FACTOR = 10.0
CFL_GROWTH_RATE_TOL = 1e-3
CFL_SAFETY = 0.8


@dataclass(frozen=True)
class GaussianPulse:
    """session.rs:242-261: g(t) = exp(-((t-t0)/tau)^2) cos(2 pi f0 (t-t0))."""

    t0: float
    tau: float
    f0: float | None = None

    def eval(self, t: float) -> float:
        x = (t - self.t0) / self.tau
        env = math.exp(-(x * x))
        return env if self.f0 is None else env * math.cos(2.0 * math.pi * self.f0 * (t - self.t0))

    __call__ = eval


def normal_vector(n: int, seed: int) -> np.ndarray:
    """session.rs:415-431: splitmix64 uniforms, Box-Muller normals (vectorised)."""
    mask = (1 << 64) - 1
    k = np.arange(1, 2 * n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)) & np.uint64(mask)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = ((z >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


class TdSession:
    """Session over one GPU operator (session.rs:281-452)."""

    def __init__(self, op: MaxwellOperator, c: float = 1.0, form: str = "fused_f64", ctx: GpuContext | None = None):
        self.op = op
        self.c = c
        self.ctx = ctx or GpuContext()
        self.gop = GpuOperator(self.ctx, op, form)
        self._cfl_h: float | None = None

    def n_dof(self) -> int:
        return self.gop.n_dof()

    # session.rs:514-569
    def cfl_dt(self, recompute: bool = False) -> float:
        if self._cfl_h is not None and not recompute:
            return self._cfl_h / self.c
        n = self.n_dof()
        v = normal_vector(n, 0)
        v /= np.sqrt(np.sum(v * v))
        rho = 1.0
        for _ in range(CFL_POWER_ITERS):
            av = self.gop.apply_f64(v)
            rho = float(np.sqrt(np.sum(av * av)))
            v = av / rho
        probe = v
        n0 = float(np.sqrt(np.sum(probe * probe)))

        def stable(z: float) -> bool:
            self.gop.state_load(probe)
            self.gop.state_lserk(z / rho, CFL_PROBE_STEPS)
            y = self.gop.state_store()
            if not np.all(np.isfinite(y)):
                return False
            with np.errstate(over="ignore", invalid="ignore"):
                nf = float(np.sqrt(np.sum(y * y)))
            if not math.isfinite(nf):
                return False
            if nf > CFL_GROWTH_FACTOR * n0:
                return False
            rate = (nf / n0) ** (1.0 / CFL_PROBE_STEPS) if nf > 0 else 0.0
            return rate <= 1.0 + CFL_GROWTH_RATE_TOL

        lo, hi = CFL_Z_STABLE, CFL_Z_UNSTABLE
        while not stable(lo) and lo > 0.1:
            lo *= 0.5
        while stable(hi) and hi < 1e3:
            hi *= 1.5
        for _ in range(CFL_BISECT_ITERS):
            mid = 0.5 * (lo + hi)
            if stable(mid):
                lo = mid
            else:
                hi = mid
        self._cfl_h = CFL_SAFETY * lo / rho
        self.rho = rho
        return self._cfl_h / self.c

    # ---- single steps (session.rs:479-510) ----
    def step(self, y, h, krylov_dim=40, tol=0.0):
        return self.gop.expmv(self.ctx, y, self.c * h, krylov_dim, tol)

    def step_explicit(self, y, h):
        return self.gop.transient(self.ctx, y, self.c * h, 1, 1, record=False, precision="f64")

    def step_adaptive(self, y, h):
        """One KCL step of h with no controller: the fourth-order state."""
        ctl = Controller(atol=1e300, rtol=0.0)
        traj, _ = self.gop.transient_kcl(self.ctx, y, self.c * h, 1, ctl=ctl, precision="f64")
        return traj[-1]

    # ---- runs (session.rs:620-777) ----
    def transient(self, y0, dt, steps, drive: Drive | None = None, waveform=None, method="explicit",
                  krylov_dim=40, controller: Controller | None = None):
        """Runs `steps` frames of dt; returns (frames (steps+1, n), KclStats or None)."""
        drive = drive or Drive.free()
        n = self.n_dof()
        y0 = np.zeros(n) if y0 is None else np.ascontiguousarray(y0, np.float64)
        if not (math.isfinite(dt) and dt > 0):
            raise ValueError(f"dt must be positive, got {dt}")
        if drive.kind == 1 and not 0 <= drive.dof < n:
            raise ValueError(f"source DOF {drive.dof} out of range (n_dof = {n})")
        if drive.kind == 2 and drive.vec is not None and drive.vec.shape != (n,):
            raise ValueError(f"the source vector has length {drive.vec.shape[0]}, expected {n}")
        if drive.kind != 0 and waveform is None:
            raise ValueError("a driven run needs a waveform")
        if method == "explicit":
            nsub = max(1, math.ceil(dt / self.cfl_dt()))
            h_sub = dt / nsub
            g = None
            if drive.kind != 0:
                g = np.array([waveform(k * h_sub) for k in range(steps * nsub)], np.float64)
            traj = self.gop.transient(self.ctx, y0, self.c * dt, steps, nsub, drive, g, True, precision="f64")
            return traj, None
        if method == "adaptive":
            g = None if drive.kind == 0 else np.array([waveform(k * dt) for k in range(steps)], np.float64)
            traj, st = self.gop.transient_kcl(self.ctx, y0, self.c * dt, steps, drive, g, controller,
                                              precision="f64")
            return traj, KclStats(st.accepted, st.rejected, st.h_min / self.c, st.h_max / self.c)
        if method == "exponential":
            if drive.kind == 2 and drive.vec is None:
                raise ValueError("the exponential route needs the source pattern on the host: pass "
                                 "Drive.vector(gop.port_source(port)) instead of Drive.port_source()")
            frames = [y0]
            y = y0
            for k in range(steps):
                if drive.kind == 0:
                    y = self.gop.expmv(self.ctx, y, self.c * dt, krylov_dim, KRYLOV_TOL)
                else:
                    gv = waveform(k * dt)
                    if drive.kind == 1:
                        b = np.zeros(n)
                        b[drive.dof] = gv
                    else:
                        b = drive.vec * gv
                    y = self.gop.etd_step(self.ctx, y, b, self.c * dt, krylov_dim, KRYLOV_TOL)
                frames.append(y)
            return np.stack(frames), None
        raise ValueError(f"unknown method {method!r}")
