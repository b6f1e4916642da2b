"""The reference's own GPU gates (crates/rapidfem-td/src/gpu/operator.rs:1514-1959),
restated case by case: same mesh (structured_box(3, 3, 3), N=2, upwind),
same inputs, same step selection (dt = 1/rho from 30 power iterations of the
CPU apply) and step counts, with the CPU side taken from the oracle and the
GPU operator built by the package's own host builder, as GpuOperator::new
builds it from a MaxwellOperator.

Tolerances: the reference's GPU_REL_TOL = 1e-3 (constants.rs:50) for the
plain FP32 form, which keeps the reference's kernel structure and
precision; 1e-10 for the FP64 forms; 1e-6 for the compressed-register form
on LSERK4 runs (its registers hold 2^-32 of each value; its apply and
Krylov paths are FP64).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GPU_REL_TOL = 1e-3
FORMS = ["plain_f32", "plain_f64", "fused_f64", "fused_f64c"]


def _tol(form, lserk=False):
    if form == "plain_f32":
        return GPU_REL_TOL
    if form == "fused_f64c" and lserk:
        return 1e-6
    return 1e-10


@pytest.fixture(scope="module")
def ctx():
    from rapidfem_amd import GpuContext

    return GpuContext()


@pytest.fixture(scope="module")
def cavity(oracle):
    """structured_box(3, 3, 3, 1, 1, 1), N=2, upwind, vacuum: the package
    operator and the oracle's."""
    import rapidfem_amd as R

    mesh = R.structured_box(3, 3, 3, 1.0, 1.0, 1.0)
    op = R.MaxwellOperator(mesh, 2, 1.0)
    oop = oracle.Operator(mesh.nodes, mesh.tets, 2, 1.0)
    return op, oop


def _rho(oop, v0):
    """30 power iterations of the CPU apply (operator.rs:1592-1601)."""
    v = v0.copy()
    rho = 1.0
    for _ in range(30):
        av = oop.apply(v)
        rho = float(np.sqrt(np.sum(av * av)))
        v = av * (1.0 / rho)
    return rho


def _sin(n, a, b):
    return np.sin(a + np.arange(n) * b)


@pytest.mark.parametrize("form", FORMS)
def test_gpu_apply_matches_cpu(ctx, oracle, form):
    """operator.rs:1514-1561: vacuum cavity and a dielectric fill (eps_r = 4)."""
    import rapidfem_amd as R

    mesh = R.structured_box(3, 3, 3, 1.0, 1.0, 1.0)
    K = mesh.n_tets
    for label, eps in (("vacuum", None), ("dielectric er=4", 4.0)):
        mats = None if eps is None else {"eps": np.full((K, 3), eps)}
        op = R.MaxwellOperator(mesh, 2, 1.0, materials=mats)
        oop = oracle.Operator(mesh.nodes, mesh.tets, 2, 1.0, eps=eps)
        n = op.n_dof
        y = _sin(n, 0.3, 0.017)
        want = oop.apply(y)
        g = R.GpuOperator(ctx, op, form)
        assert g.n_dof() == n
        got = g.apply(ctx, y) if form == "plain_f32" else g.apply_f64(y)
        rel = rel_l2(got, want)
        print(f"apply [{label}] form={form}: rel L2 {rel:.3e}", flush=True)
        assert rel < (GPU_REL_TOL if form == "plain_f32" else 1e-12)


@pytest.mark.parametrize("form", FORMS)
def test_gpu_kcl_adaptive_matches_cpu(ctx, oracle, cavity, form):
    """operator.rs:1563-1643: 60 frames of dt = 1/rho, atol 1e-6, rtol 1e-3,
    final snapshot within 10 GPU_REL_TOL (the reference's bound)."""
    import rapidfem_amd as R

    op, oop = cavity
    n = op.n_dof
    y0 = _sin(n, 0.2, 0.011)
    dt = 1.0 / _rho(oop, y0)
    steps = 60
    octl = oracle.Controller.default()
    octl.atol, octl.rtol = 1e-6, 1e-3
    want, acc, rej, _, _ = oop.adaptive(y0, dt, steps, octl)
    g = R.GpuOperator(ctx, op, form)
    ctl = R.Controller(atol=1e-6, rtol=1e-3)
    prec = "f32" if form == "plain_f32" else "f64"
    traj, st = g.transient_kcl(ctx, y0, dt, steps, ctl=ctl, precision=prec)
    rel = rel_l2(traj[-1], want)
    print(f"kcl form={form}: gpu {st.accepted}/{st.rejected}, cpu {acc}/{rej}, rel L2 {rel:.3e}", flush=True)
    assert rel < (10 * GPU_REL_TOL if form == "plain_f32" else 1e-9)
    if form != "plain_f32":
        assert (st.accepted, st.rejected) == (acc, rej)


@pytest.mark.parametrize("form", FORMS)
def test_gpu_transient_matches_cpu(ctx, oracle, cavity, form):
    """operator.rs:1645-1701: 200 free LSERK4 steps of dt = 1/rho."""
    import rapidfem_amd as R

    op, oop = cavity
    n = op.n_dof
    y0 = _sin(n, 0.2, 0.011)
    dt = 1.0 / _rho(oop, y0)
    steps = 200
    want = oop.lserk(y0, dt, steps)
    g = R.GpuOperator(ctx, op, form)
    prec = "f32" if form == "plain_f32" else "f64"
    got = g.transient(ctx, y0, dt, steps, 1, record=False, precision=prec)
    rel = rel_l2(got, want)
    print(f"transient form={form}: rel L2 {rel:.3e}", flush=True)
    assert rel < _tol(form, lserk=True)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("drive", ["point", "vector"])
def test_gpu_driven_transient_matches_cpu(ctx, oracle, cavity, form, drive):
    """operator.rs:1703-1836: 150 steps from rest with a point source at
    n/3 (g_k = sin(0.3 k)) or the spread pattern b_i = 0.5 cos(0.09 i)."""
    import rapidfem_amd as R

    op, oop = cavity
    n = op.n_dof
    dt = 1.0 / _rho(oop, _sin(n, 0.1, 0.007))
    steps = 150
    wave = np.sin(0.3 * np.arange(steps))
    if drive == "point":
        sdof = n // 3
        d, od = R.Drive.point(sdof), ("point", sdof)
    else:
        b = 0.5 * np.cos(0.09 * np.arange(n))
        d, od = R.Drive.vector(b), ("vector", b)
    want = oop.lserk(np.zeros(n), dt, steps, drive=od, g=wave)
    g = R.GpuOperator(ctx, op, form)
    prec = "f32" if form == "plain_f32" else "f64"
    got = g.transient(ctx, np.zeros(n), dt, steps, 1, drive=d, g=wave, record=False, precision=prec)
    rel = rel_l2(got, want)
    print(f"driven transient [{drive}] form={form}: rel L2 {rel:.3e}", flush=True)
    assert rel < _tol(form, lserk=True)


@pytest.mark.parametrize("form", FORMS)
def test_gpu_expmv_matches_cpu(ctx, oracle, cavity, form):
    """operator.rs:1838-1887: v_i = sin(0.3 + 0.013 i), t = 0.02, m = 40."""
    import rapidfem_amd as R

    op, oop = cavity
    n = op.n_dof
    v = _sin(n, 0.3, 0.013)
    want, _ = oop.expmv(v, 0.02, 40)
    got = R.GpuOperator(ctx, op, form).expmv(ctx, v, 0.02, 40)
    rel = rel_l2(got, want)
    print(f"expmv form={form}: rel L2 {rel:.3e}", flush=True)
    assert rel < _tol(form)


@pytest.mark.parametrize("form", FORMS)
def test_gpu_etd_step_matches_cpu(ctx, oracle, cavity, form):
    """operator.rs:1890-1958: y_i = sin(0.25 + 0.013 i), h = 0.02, m = 40,
    a point source (1.7 at n/3) and a vector source 0.4 cos(0.07 i)."""
    import rapidfem_amd as R

    op, oop = cavity
    n = op.n_dof
    y = _sin(n, 0.25, 0.013)
    b_point = np.zeros(n)
    b_point[n // 3] = 1.7
    b_vec = 0.4 * np.cos(0.07 * np.arange(n))
    g = R.GpuOperator(ctx, op, form)
    for label, b in (("point", b_point), ("vector", b_vec)):
        want = oop.etd_step(y, b, 0.02, 40)
        got = g.etd_step(ctx, y, b, 0.02, 40)
        rel = rel_l2(got, want)
        print(f"etd_step [{label}] form={form}: rel L2 {rel:.3e}", flush=True)
        assert rel < _tol(form)
