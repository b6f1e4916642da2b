"""TdSession::transient routing (crates/rapidfem-td/src/session.rs:620-777)
over the GPU operator against the oracle.

  explicit     nsub = ceil(dt / cfl_dt) LSERK4 substeps per frame, the
               waveform sampled at k*h_sub (session.rs:673-678, 885-927),
               operator time scaled by c;
  adaptive     KCL with the PI controller, the waveform held per frame and
               h_min/h_max reported in physical time (session.rs:940-970);
  exponential  one ETD step per frame with the source b*g(k dt) (a point or
               vector drive), or expmv for a free run, at Arnoldi dimension
               krylov_dim and KRYLOV_TOL (session.rs:700-726, 867-874).

Meshes are the oracle's 3x3x3 jittered Kuhn box (162 tets); a wave speed
c = 2 makes every route scale its operator time. FP64 forms: 1e-11
relative L2 per frame.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

C_WAVE = 2.0


@pytest.fixture(scope="module")
def ctx():
    from rapidfem_amd import GpuContext

    return GpuContext()


def _session(ctx, oracle, order=2, form="fused_f64"):
    """A TdSession on the package's MaxwellOperator of the jittered box, and
    the oracle operator of the same mesh."""
    import rapidfem_amd as R
    from rapidfem_amd import maxwell

    xyz, tets = oracle.box_mesh(3, 3, 3, 1.0, 1.0, 1.0, 0.1, 7, True)
    op = maxwell.MaxwellOperator(maxwell.Mesh(xyz, tets), order, 1.0)
    oop = oracle.Operator(xyz, tets, order, 1.0)
    x = op.node_coords()
    y0 = np.zeros((op.n_elem, x.shape[1], 6))
    y0[..., 1] = np.sin(np.pi * x[..., 0]) * np.sin(np.pi * x[..., 2])
    y0[..., 5] = 0.2 * np.cos(np.pi * x[..., 1])
    return R.TdSession(op, C_WAVE, form, ctx), oop, y0.reshape(-1)


@pytest.mark.parametrize("form", ["plain_f64", "fused_f64"])
def test_explicit_route_substeps_and_samples_the_pulse_per_substep(ctx, oracle, form):
    from rapidfem_amd import Drive, GaussianPulse

    sess, oop, y0 = _session(ctx, oracle, form=form)
    cfl = sess.cfl_dt()
    assert abs(cfl * C_WAVE - oop.cfl_dt()) <= 1e-9 * oop.cfl_dt()
    dt = 2.5 * cfl
    nsub = math.ceil(dt / cfl)
    assert nsub == 3
    steps, dof = 4, sess.n_dof() // 3 + 5
    pulse = GaussianPulse(t0=2 * dt, tau=dt, f0=1.5)
    frames, stats = sess.transient(y0, dt, steps, Drive.point(dof), pulse, method="explicit")
    assert stats is None and frames.shape == (steps + 1, sess.n_dof())
    h_sub = dt / nsub
    y = y0
    assert rel_l2(frames[0], y0) == 0.0
    for k in range(steps):
        g = np.array([pulse(float(k * nsub + j) * h_sub) for j in range(nsub)])
        y = oop.lserk(y, C_WAVE * h_sub, nsub, ("point", dof), g)
        err = rel_l2(frames[k + 1], y)
        print(f"explicit route {form} frame {k + 1}: rel L2 vs oracle {err:.3e}", flush=True)
        assert err < 1e-11


def test_explicit_route_rejects_bad_drives(ctx, oracle):
    from rapidfem_amd import Drive, GaussianPulse

    sess, _, y0 = _session(ctx, oracle)
    pulse = GaussianPulse(0.1, 0.05)
    with pytest.raises(ValueError):
        sess.transient(y0, 0.01, 2, Drive.point(sess.n_dof()), pulse)
    with pytest.raises(ValueError):
        sess.transient(y0, 0.01, 2, Drive.vector(np.ones(7)), pulse)
    with pytest.raises(ValueError):
        sess.transient(y0, 0.01, 2, Drive.point(3), None)
    with pytest.raises(ValueError):
        sess.transient(y0, -1.0, 2)
    with pytest.raises(ValueError):
        sess.transient(y0, 0.01, 2, Drive.port_source(), pulse, method="exponential")


def test_adaptive_route_holds_the_waveform_per_frame(ctx, oracle):
    from rapidfem_amd import Controller, Drive, GaussianPulse

    sess, oop, y0 = _session(ctx, oracle)
    n = sess.n_dof()
    b = np.sin(np.arange(n) * 0.013)
    dt, steps = 0.04, 3
    pulse = GaussianPulse(t0=0.05, tau=0.03, f0=4.0)
    ctl = Controller(atol=1e-8, rtol=1e-6)
    frames, st = sess.transient(y0, dt, steps, Drive.vector(b), pulse, method="adaptive", controller=ctl)
    octl = oracle.Controller.default()
    octl.atol, octl.rtol = 1e-8, 1e-6
    g = np.array([pulse(k * dt) for k in range(steps)])
    y, acc, rej, hmin, hmax = oop.adaptive(y0, C_WAVE * dt, steps, octl, ("vector", b), g)
    err = rel_l2(frames[-1], y)
    print(f"adaptive route: {st.accepted}/{st.rejected} vs oracle {acc}/{rej}, rel L2 {err:.3e}", flush=True)
    assert (st.accepted, st.rejected) == (acc, rej)
    assert err < 1e-11
    # The next step is h * safety * e^-0.175 ... with e the embedded error
    # norm, a sum of (bhat - b) f terms that cancel to O(h^4) of their size:
    # round-off of 1e-16 in the applies moves e by ~1e-7 relative and the
    # proposed steps by ~2e-8 (measured 2.3e-8 on h_min), so h_min and h_max
    # are compared at 1e-6.
    assert abs(st.h_min - hmin / C_WAVE) <= 1e-6 * st.h_min
    assert abs(st.h_max - hmax / C_WAVE) <= 1e-6 * st.h_max


@pytest.mark.parametrize("drive", ["free", "point", "vector"])
def test_exponential_route_steps_etd_per_frame(ctx, oracle, drive):
    from rapidfem_amd import Drive, GaussianPulse
    from rapidfem_amd.session import KRYLOV_TOL

    sess, oop, y0 = _session(ctx, oracle)
    n = sess.n_dof()
    dt, steps, m = 0.03, 3, 30
    pulse = GaussianPulse(t0=0.04, tau=0.03, f0=3.0)
    if drive == "free":
        d, wave = Drive.free(), None
    elif drive == "point":
        d, wave = Drive.point(n // 2 + 1), pulse
    else:
        bvec = np.cos(np.arange(n) * 0.021)
        d, wave = Drive.vector(bvec), pulse
    frames, _ = sess.transient(y0, dt, steps, d, wave, method="exponential", krylov_dim=m)
    y = y0
    for k in range(steps):
        if drive == "free":
            y, _ = oop.expmv(y, C_WAVE * dt, m, KRYLOV_TOL)
        else:
            gv = pulse(k * dt)
            if drive == "point":
                b = np.zeros(n)
                b[n // 2 + 1] = gv
            else:
                b = bvec * gv
            y = oop.etd_step(y, b, C_WAVE * dt, m, KRYLOV_TOL)
        err = rel_l2(frames[k + 1], y)
        print(f"exponential route {drive} frame {k + 1}: rel L2 vs oracle {err:.3e}", flush=True)
        assert err < 1e-11
