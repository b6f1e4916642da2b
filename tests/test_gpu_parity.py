"""GPU parity: every kernel form against the CPU oracle, through the C ABI.

Operator data comes from the oracle (the reference's f64 values), so these
compare kernels on identical inputs. Tolerances: FP64 forms 1e-12 relative
L2 per operator apply; the f32 form the reference's own GPU gate
GPU_REL_TOL = 1e-3 (constants.rs:50).
"""
import numpy as np
import pytest

from conftest import oracle_desc, rel_l2

pytestmark = pytest.mark.gpu

FORMS64 = ["plain_f64", "fused_f64", "fused_f64c"]
ALL_FORMS = ["plain_f32"] + FORMS64


@pytest.fixture(scope="module")
def ctx():
    from rapidfem_amd import GpuContext

    return GpuContext()


def _mesh(oracle, n=3, jitter=True, lx=1.0):
    return oracle.box_mesh(n, n, n, lx, 1.0, 1.0, 0.1 if jitter else 0.0, 7, jitter)


def _gop(ctx, oop, form):
    from rapidfem_amd import GpuOperator

    d, arrs, re_, ex = oracle_desc(oop)
    g = GpuOperator.from_desc(ctx, d, form, mass=re_["mass"], det=ex["det"])
    return g


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("form", ALL_FORMS)
def test_apply_matches_oracle(ctx, oracle, order, form):
    xyz, tets = _mesh(oracle, 3)
    K = tets.shape[0]
    eps = np.ones((K, 3))
    eps[::3] = 4.0
    sig = np.zeros(K)
    sig[::5] = 0.3
    oop = oracle.Operator(xyz, tets, order, 1.0, eps=eps, sigma=sig)
    n = oop.n_dof
    y = np.sin(0.3 + 0.017 * np.arange(n))
    want = oop.apply(y)
    g = _gop(ctx, oop, form)
    if form == "plain_f32":
        got = g.apply(ctx, y.astype(np.float32))
        tol = 1e-3
    else:
        got = g.apply_f64(y)
        tol = 1e-12
    err = rel_l2(got, want)
    print(f"apply order={order} form={form} rel_l2={err:.3e}")
    assert err < tol


def _te101(oracle, n=3, order=2, jitter=True, alpha=1.0):
    xyz, tets = _mesh(oracle, n, jitter)
    oop = oracle.Operator(xyz, tets, order, alpha)
    from rapidfem_amd import maxwell
    mesh = maxwell.Mesh(xyz, tets)
    x = maxwell.MaxwellOperator(mesh, order, alpha).node_coords()
    y = np.zeros((oop.K, x.shape[1], 6))
    y[..., 1] = np.sin(np.pi * x[..., 0]) * np.sin(np.pi * x[..., 2])
    y[..., 3] = 0.3 * np.sin(2 * np.pi * x[..., 1])  # not an eigenmode: exercise all terms
    return oop, y.reshape(-1)


@pytest.mark.parametrize("form", ALL_FORMS)
@pytest.mark.parametrize("drive", ["free", "point", "vector"])
def test_lserk_transient_matches_oracle(ctx, oracle, form, drive):
    """GpuOperator::transient (gpu/operator.rs:906-953) vs the oracle's LSERK4 run
    (explicit.rs:100-135): 3 frames x 4 substeps, held source per substep."""
    from rapidfem_amd import Drive

    oop, y0 = _te101(oracle, 3, 2)
    g = _gop(ctx, oop, form)
    n = oop.n_dof
    steps, sub, dt = 3, 4, 0.08
    h = dt / sub
    wave = np.cos(0.7 * np.arange(steps * sub))
    if drive == "free":
        d, od, gg = Drive.free(), None, None
    elif drive == "point":
        d, od, gg = Drive.point(n // 3 + 1), ("point", n // 3 + 1), wave
    else:
        b = np.sin(np.arange(n) * 0.01)
        d, od, gg = Drive.vector(b), ("vector", b), wave
    prec = "f32" if form == "plain_f32" else "f64"
    traj = g.transient(ctx, y0, dt, steps, sub, d, gg, True, precision=prec)
    want = [y0]
    y = y0
    for k in range(steps):
        y = oop.lserk(y, h, sub, od, None if gg is None else gg[k * sub:(k + 1) * sub])
        want.append(y)
    tol = {"plain_f32": 1e-3, "plain_f64": 1e-12, "fused_f64": 1e-12, "fused_f64c": 1e-7}[form]
    for k in range(steps + 1):
        err = rel_l2(traj[k], want[k])
        print(f"transient form={form} drive={drive} frame={k} rel_l2={err:.3e}")
        assert err < tol


@pytest.mark.parametrize("form", FORMS64)
def test_kcl_adaptive_matches_oracle(ctx, oracle, form):
    """GpuOperator::transient_kcl vs the oracle's KCL + PI controller: same
    accept/reject counts and states (session.rs:781-827)."""
    from rapidfem_amd import Controller

    oop, y0 = _te101(oracle, 3, 2)
    g = _gop(ctx, oop, form)
    dt, steps = 0.05, 3
    ctl = Controller(atol=1e-8, rtol=1e-6)
    traj, st = g.transient_kcl(ctx, y0, dt, steps, ctl=ctl, precision="f64")
    octl = oracle.Controller.default()
    octl.atol, octl.rtol = 1e-8, 1e-6
    y, acc, rej, hmin, hmax = oop.adaptive(y0, dt, steps, octl)
    err = rel_l2(traj[-1], y)
    print(f"kcl form={form}: gpu acc/rej {st.accepted}/{st.rejected}, oracle {acc}/{rej}, rel_l2 {err:.3e}")
    assert (st.accepted, st.rejected) == (acc, rej)
    assert err < (1e-7 if form == "fused_f64c" else 1e-11)
    assert abs(st.h_min - hmin) <= 1e-9 * hmin and abs(st.h_max - hmax) <= 1e-9 * hmax


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("drive", ["free", "vector"])
def test_kcl_compressed_registers_match_oracle(ctx, oracle, order, drive):
    """KCL with compressed registers (y and its backup f32 + bf16, err f32,
    stage input FP64) on the compressed-register form: the oracle's
    accept/reject counts, the state at the form's precision, and a
    rejection restores y from its compressed backup."""
    from rapidfem_amd import Controller, Drive, RfhipError

    oop, y0 = _te101(oracle, 3, order)
    n = oop.n_dof
    g = _gop(ctx, oop, "fused_f64c")
    g.set_kcl_registers("compressed")
    kw, okw = {}, {}
    if drive == "vector":
        b = np.cos(0.37 * np.arange(n))
        wave = np.array([0.4, -0.3, 0.2])
        kw = dict(drive=Drive.vector(b), g=wave)
        okw = dict(drive=("vector", b), g=wave)
    # a tight controller and a long first step force rejections
    ctl = Controller(atol=1e-9, rtol=1e-7)
    dt, steps = 0.05, 3
    traj, st = g.transient_kcl(ctx, y0, dt, steps, ctl=ctl, precision="f64", **kw)
    octl = oracle.Controller.default()
    octl.atol, octl.rtol = 1e-9, 1e-7
    y, acc, rej, _, _ = oop.adaptive(y0, dt, steps, octl, **okw)
    errs = []
    for k in range(1, steps + 1):
        yk, _, _, _, _ = oop.adaptive(y0, dt, k, octl, **okw)
        errs.append(rel_l2(traj[k], yk))
    print(f"kcl compressed order={order} drive={drive}: gpu {st.accepted}/{st.rejected}, oracle {acc}/{rej}, "
          f"rel_l2 per frame {[f'{e:.2e}' for e in errs]}", flush=True)
    assert (st.accepted, st.rejected) == (acc, rej)
    assert rej > 0
    assert max(errs) < 1e-7
    g.set_kcl_registers("f64")
    traj64, st64 = g.transient_kcl(ctx, y0, dt, steps, ctl=ctl, precision="f64", **kw)
    assert (st64.accepted, st64.rejected) == (acc, rej)
    assert rel_l2(traj64[-1], y) < 1e-11
    g64 = _gop(ctx, oop, "fused_f64")
    with pytest.raises(RfhipError):
        g64.set_kcl_registers("compressed")


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("variant", [5, 7])
@pytest.mark.parametrize("drive", ["free", "vector"])
def test_kcl_on_stage_kernel_variants_matches_oracle(ctx, oracle, order, variant, drive):
    """KCL with its applies on the block-wide stage kernel (variant 5) and
    the whole-elements-per-wave one (variant 7), each stage followed by the
    merged register update and next-stage build, free and vector-driven:
    accept/reject counts and states equal the oracle's."""
    from rapidfem_amd import Controller, Drive

    oop, y0 = _te101(oracle, 3, order)
    n = oop.n_dof
    g = _gop(ctx, oop, "fused_f64")
    g.set_variant(variant)
    g.poison_lds()
    dt, steps = 0.05, 3
    kw, okw = {}, {}
    if drive == "vector":
        b = np.cos(0.37 * np.arange(n))
        wave = np.array([0.4, -0.3, 0.2])
        kw = dict(drive=Drive.vector(b), g=wave)
        okw = dict(drive=("vector", b), g=wave)
    ctl = Controller(atol=1e-8, rtol=1e-6)
    traj, st = g.transient_kcl(ctx, y0, dt, steps, ctl=ctl, precision="f64", **kw)
    octl = oracle.Controller.default()
    octl.atol, octl.rtol = 1e-8, 1e-6
    y, acc, rej, _, _ = oop.adaptive(y0, dt, steps, octl, **okw)
    err = rel_l2(traj[-1], y)
    print(f"kcl order={order} variant={variant} drive={drive}: gpu {st.accepted}/{st.rejected}, "
          f"oracle {acc}/{rej}, rel_l2 {err:.3e}", flush=True)
    assert (st.accepted, st.rejected) == (acc, rej)
    assert err < 1e-11


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("variant", [5, 6, 7])
@pytest.mark.parametrize("drive", ["free", "vector", "point"])
def test_kcl_fused_update_matches_separate_pass(ctx, oracle, order, variant, drive):
    """Compressed KCL registers with the update in the stage kernel's
    epilogue (no k stored; the stage input alternates between two buffers;
    the error norm summed per wave) against the same registers with the
    update as a pass of its own after each apply (mode
    "compressed_separate"), on the operator's variant (5: the fused update
    runs on variant 7's kernel), free, vector- and point-driven (the point
    drive lands on one entry of one node's epilogue): the same
    accept/reject sequence and step
    range, the same states to the registers' rounding, and the oracle's
    counts and state within that rounding. The tight
    controller forces rejections, whose restore rebuilds the stage input in
    whichever buffer the next substep reads."""
    from rapidfem_amd import Controller, Drive

    oop, y0 = _te101(oracle, 3, order)
    n = oop.n_dof
    g = _gop(ctx, oop, "fused_f64c")
    g.set_variant(variant)
    kw, okw = {}, {}
    if drive == "vector":
        b = np.cos(0.37 * np.arange(n))
        wave = np.array([0.4, -0.3, 0.2])
        kw = dict(drive=Drive.vector(b), g=wave)
        okw = dict(drive=("vector", b), g=wave)
    elif drive == "point":
        wave = np.array([0.4, -0.3, 0.2])
        kw = dict(drive=Drive.point(n // 3 + 1), g=wave)
        okw = dict(drive=("point", n // 3 + 1), g=wave)
    ctl = Controller(atol=1e-9, rtol=1e-7)
    dt, steps = 0.05, 3
    runs = {}
    for mode in ["compressed_separate", "compressed"]:
        g.set_kcl_registers(mode)
        g.poison_lds()
        runs[mode] = g.transient_kcl(ctx, y0, dt, steps, ctl=ctl, precision="f64", **kw)
    (ts, ss), (tf, sf) = runs["compressed_separate"], runs["compressed"]
    octl = oracle.Controller.default()
    octl.atol, octl.rtol = 1e-9, 1e-7
    y, acc, rej, _, _ = oop.adaptive(y0, dt, steps, octl, **okw)
    d = max(rel_l2(tf[k], ts[k]) for k in range(1, steps + 1))
    err = rel_l2(tf[-1], y)
    print(f"kcl fused update order={order} variant={variant} drive={drive}: fused {sf.accepted}/{sf.rejected}, "
          f"separate {ss.accepted}/{ss.rejected}, oracle {acc}/{rej}; fused vs separate {d:.2e}, vs oracle {err:.2e}",
          flush=True)
    assert (sf.accepted, sf.rejected) == (ss.accepted, ss.rejected)
    # the registers' rounding can move a decision near the controller's
    # threshold against the FP64 oracle (N=4, vector drive: 193/5 against
    # 196/4), for both paths alike
    assert abs(sf.accepted - acc) <= 4 and abs(sf.rejected - rej) <= 2
    assert sf.rejected > 0
    # the two paths sum the error norm in different orders; a difference in
    # the last bit of h moves some compressed registers across a rounding
    # boundary of their bf16 part (2^-32 relative), and at rtol 1e-7 the
    # embedded error estimate sees that rounding at about 1e-3 of its size,
    # so the step sequences drift apart (by 2.4e-3 in h_min over 127
    # substeps at N=4), the states by the integration error that makes
    assert abs(sf.h_min - ss.h_min) <= 1e-2 * ss.h_min and abs(sf.h_max - ss.h_max) <= 1e-2 * ss.h_max
    assert d < 1e-8
    assert err < 1e-7


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("variant", [5, 7])
@pytest.mark.parametrize("drive", ["free", "vector"])
def test_kcl_fused_fp64_update_matches_separate_pass(ctx, oracle, order, variant, drive):
    """FP64 KCL registers (the default) with the update in the stage kernel's
    epilogue against the same registers with the update as a pass of its own
    (mode "f64_separate"), on the FP64 fused form: equal accept/reject
    counts, the step range and the states to FP64 rounding (the paths sum
    the error norm in different orders), and the oracle's counts and state."""
    from rapidfem_amd import Controller, Drive

    oop, y0 = _te101(oracle, 3, order)
    n = oop.n_dof
    g = _gop(ctx, oop, "fused_f64")
    g.set_variant(variant)
    kw, okw = {}, {}
    if drive == "vector":
        b = np.cos(0.37 * np.arange(n))
        wave = np.array([0.4, -0.3, 0.2])
        kw = dict(drive=Drive.vector(b), g=wave)
        okw = dict(drive=("vector", b), g=wave)
    ctl = Controller(atol=1e-9, rtol=1e-7)
    dt, steps = 0.05, 3
    runs = {}
    for mode in ["f64_separate", "f64"]:
        g.set_kcl_registers(mode)
        g.poison_lds()
        runs[mode] = g.transient_kcl(ctx, y0, dt, steps, ctl=ctl, precision="f64", **kw)
    (ts, ss), (tf, sf) = runs["f64_separate"], runs["f64"]
    octl = oracle.Controller.default()
    octl.atol, octl.rtol = 1e-9, 1e-7
    y, acc, rej, _, _ = oop.adaptive(y0, dt, steps, octl, **okw)
    d = max(rel_l2(tf[k], ts[k]) for k in range(1, steps + 1))
    err = rel_l2(tf[-1], y)
    print(f"kcl fused FP64 update order={order} variant={variant} drive={drive}: fused {sf.accepted}/{sf.rejected}, "
          f"separate {ss.accepted}/{ss.rejected}, oracle {acc}/{rej}; fused vs separate {d:.2e}, vs oracle {err:.2e}",
          flush=True)
    assert (sf.accepted, sf.rejected) == (ss.accepted, ss.rejected) == (acc, rej)
    assert sf.rejected > 0
    # the embedded error estimate, a difference of stage terms about 1e4
    # times its size, carries the summation-order difference into h at
    # about 1e-9 relative over 127 substeps (N=4)
    assert abs(sf.h_min - ss.h_min) <= 1e-8 * ss.h_min and abs(sf.h_max - ss.h_max) <= 1e-8 * ss.h_max
    assert d < 1e-12
    assert err < 1e-11


def test_kcl_collapse_is_an_error(ctx, oracle):
    """A step collapse is reported (session.rs:817-824), not a hang."""
    from rapidfem_amd import Controller, RfhipError

    oop, y0 = _te101(oracle, 2, 2)
    g = _gop(ctx, oop, "fused_f64")
    ctl = Controller(atol=1e-30, rtol=1e-30, min_step_factor=0.5)
    with pytest.raises(RfhipError) as ei:
        g.transient_kcl(ctx, y0, 0.05, 2, ctl=ctl, precision="f64")
    assert ei.value.code == 5


@pytest.mark.parametrize("form", ["plain_f32", "plain_f64", "fused_f64"])
@pytest.mark.parametrize("tol", [0.0, 1e-10])
def test_expmv_matches_oracle(ctx, oracle, form, tol):
    """GpuOperator::expmv vs propagator.rs expmv_into (CGS2 Arnoldi, Taylor-18 expm)."""
    oop, y0 = _te101(oracle, 3, 2)
    g = _gop(ctx, oop, form)
    want, wdim = oop.expmv(y0, 0.3, 30, tol)
    got = g.expmv(ctx, y0, 0.3, 30, tol)
    err = rel_l2(got, want)
    print(f"expmv form={form} tol={tol}: dim gpu {g.last_dim} oracle {wdim} rel_l2 {err:.3e}")
    assert g.last_dim == wdim
    assert err < (1e-3 if form == "plain_f32" else 1e-10)


@pytest.mark.parametrize("m", [3, 9, 48, 56])
def test_expmv_every_gram_schmidt_width_matches_oracle(ctx, oracle, m):
    """Arnoldi dimensions that cross every column bucket of the fused
    Gram-Schmidt sweep (4-column steps up to 48) and the two-sweep form
    above 48; no early stop, so the basis reaches m columns."""
    oop, y0 = _te101(oracle, 3, 2)
    g = _gop(ctx, oop, "fused_f64")
    want, wdim = oop.expmv(y0, 0.3, m, 0.0)
    got = g.expmv(ctx, y0, 0.3, m, 0.0)
    err = rel_l2(got, want)
    print(f"expmv m={m}: dim gpu {g.last_dim} oracle {wdim} rel_l2 {err:.3e}", flush=True)
    assert g.last_dim == wdim == m
    assert err < 1e-10


@pytest.mark.parametrize("form", ["plain_f64", "fused_f64", "fused_f64c"])
def test_expmv_on_graded_slab_matches_oracle(ctx, oracle, form):
    """The stiff Krylov workload at oracle size: the 5,184-tet N=3 box graded
    20:1 toward an eps_r 10, sigma 50 slab, built by the package's host
    builder; t = 3 / rho, m = 40 (the benchmark's step), FP64 Arnoldi on
    every form."""
    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    op, y0 = W.slab_box(12, 9, 8, 3)
    oop = oracle.Operator(op.mesh.nodes, op.mesh.tets, 3, 1.0, eps=1.0 / op.inv_eps,
                          sigma=op.sigma_eps[:, 0] / op.inv_eps[:, 0])
    v = y0 / np.linalg.norm(y0)
    for _ in range(20):
        av = oop.apply(v)
        rho = float(np.linalg.norm(av))
        v = av / rho
    t = 3.0 / rho
    want, wdim = oop.expmv(y0, t, 40)
    g = R.GpuOperator(ctx, op, form)
    got = g.expmv(ctx, y0, t, 40)
    err = rel_l2(got, want)
    print(f"slab expmv form={form}: t {t:.3e}, dim {g.last_dim}/{wdim}, rel_l2 {err:.3e}", flush=True)
    assert g.last_dim == wdim
    assert err < 1e-10
    assert rel_l2(got, y0) > 1e-3


@pytest.mark.parametrize("form", ["plain_f64", "fused_f64"])
def test_etd_matches_oracle(ctx, oracle, form):
    """GpuOperator::etd_step vs propagator.rs etd_step_into (augmented Arnoldi)."""
    oop, y0 = _te101(oracle, 3, 2)
    g = _gop(ctx, oop, form)
    b = np.zeros(oop.n_dof)
    b[oop.n_dof // 2] = 1.0
    b[100:160] = 0.1
    want = oop.etd_step(y0, b, 0.2, 25, 0.0)
    got = g.etd_step(ctx, y0, b, 0.2, 25, 0.0)
    err = rel_l2(got, want)
    print(f"etd form={form}: rel_l2 {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("form", FORMS64)
def test_energy_matches_oracle_and_central_flux_conserves(ctx, oracle, form):
    """field_energy (rhs.rs:1509) of the resident state matches the oracle
    before and after 200 LSERK4 steps; with the central flux (M~A skew,
    rhs.rs:1830) the discrete energy only moves by LSERK4's own O(dt^6)
    dissipation, far below the upwind flux's loss."""
    oop, y0 = _te101(oracle, 3, 2, alpha=0.0)
    g = _gop(ctx, oop, form)
    g.state_load(y0)
    e0 = g.state_energy()
    assert abs(e0 - oop.field_energy(y0)) <= 1e-12 * e0
    dt = 0.3 * oop.cfl_dt()
    g.state_lserk(dt, 200)
    e1 = g.state_energy()
    e1_or = oop.field_energy(oop.lserk(y0, dt, 200))
    drift = abs(e1 - e0) / e0
    g.state_load(y0)
    g.state_lserk(dt / 2, 400)
    drift_half = abs(g.state_energy() - e0) / e0
    print(f"energy form={form}: E0={e0:.12e} E200={e1:.12e} oracle {e1_or:.12e} drift={drift:.3e} "
          f"drift(dt/2)={drift_half:.3e}")
    assert abs(e1 - e1_or) <= (1e-8 if form == "fused_f64c" else 1e-11) * e0
    # A non-skew surface term would leave an O(T) energy change independent of dt.
    assert drift_half < drift / 16
    oup, _ = _te101(oracle, 3, 2, alpha=1.0)
    gu = _gop(ctx, oup, form)
    gu.state_load(y0)
    gu.state_lserk(dt, 200)
    assert (e0 - gu.state_energy()) / e0 > 10 * drift


def test_errors_are_status_codes(ctx, oracle):
    """Length mismatches and bad drives are EINVAL, never a crash
    (gpu/operator.rs:518,552 asserts become status codes)."""
    from rapidfem_amd import Drive, RfhipError

    oop, y0 = _te101(oracle, 2, 2)
    g = _gop(ctx, oop, "fused_f64")
    with pytest.raises(ValueError):
        g.apply_f64(y0[:-1])
    with pytest.raises(RfhipError) as ei:
        g.transient(ctx, y0, 0.01, 2, 1, Drive.point(oop.n_dof + 5), np.ones(2), precision="f64")
    assert ei.value.code == 1
    with pytest.raises(RfhipError) as ei:
        g.transient(ctx, y0, 0.01, 2, 1, Drive.point(3), np.ones(5), precision="f64")
    assert ei.value.code == 1


def test_single_element_and_ports(ctx, oracle):
    """K=1 (every face PEC) and characteristic port faces (rhs.rs:1380-1386)."""
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    tets = np.array([[0, 1, 2, 3]], np.int64)
    for port in (None, np.array([0, -1, 1, -1], np.int32)):
        oop = oracle.Operator(xyz, tets, 3, 1.0, face_port=port)
        y = np.cos(np.arange(oop.n_dof) * 0.37)
        want = oop.apply(y)
        for form in FORMS64:
            got = _gop(ctx, oop, form).apply_f64(y)
            assert rel_l2(got, want) < 1e-12, (form, port)
    xyz, tets = _mesh(oracle, 3)
    K = tets.shape[0]
    fp = -np.ones((K, 4), np.int32)
    oop0 = oracle.Operator(xyz, tets, 2, 1.0)
    ex = oop0.export()
    fp[ex["nbr"] < 0] = 0  # every boundary face is port 0 (absorbing box)
    oop = oracle.Operator(xyz, tets, 2, 1.0, face_port=fp.reshape(-1))
    y = np.cos(np.arange(oop.n_dof) * 0.11)
    want = oop.apply(y)
    for form in FORMS64:
        got = _gop(ctx, oop, form).apply_f64(y)
        assert rel_l2(got, want) < 1e-12, form


def test_product_builder_operator_matches_oracle(ctx, oracle):
    """The package's own host builder + fused kernel vs the oracle on the
    config-1 cavity (6x6x6 jittered, N=2, TE101), 50 LSERK4 steps."""
    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    op, y0 = W.cavity_te101(6, 2)
    oop = oracle.Operator(op.mesh.nodes, op.mesh.tets, 2, 1.0)
    dt = 0.016
    want = oop.lserk(y0, dt, 50)
    for form in FORMS64:
        g = R.GpuOperator(ctx, op, form)
        got = g.transient(ctx, y0, dt, 50, 1, record=False, precision="f64")
        err = rel_l2(got, want)
        print(f"builder+{form}: rel_l2 {err:.3e}")
        assert err < (1e-7 if form == "fused_f64c" else 1e-11)


def test_product_builder_debye_operator_matches_oracle(ctx, oracle):
    """The package's MaxwellOperator with Debye elements (DebyeMaterial,
    eps = eps_inf on them) through TdSession's GPU operator: the reference
    refuses this on the GPU (session.rs:592-593); here the fused and plain
    FP64 forms step it and match the oracle over 40 LSERK4 steps."""
    import rapidfem_amd as R

    mesh = R.structured_box_jittered(4, 3, 3, 1.0, 1.0, 1.0, 0.1, 5)
    K = mesh.n_tets
    deb = R.DebyeMaterial(2.5, 7.0, 0.3)
    elems = list(range(1, K, 3))
    mats = [R.ElemMaterial.isotropic(2.5 if e in set(elems) else 1.0) for e in range(K)]
    op = R.MaxwellOperator(mesh, 2, 1.0, mats, dispersive=[(e, deb) for e in elems])
    eps = np.array([[m.eps] * 3 if np.isscalar(m.eps) else m.eps for m in mats], np.float64).reshape(K, 3)
    oop = oracle.Operator(mesh.nodes, mesh.tets, 2, 1.0, eps=eps,
                          dispersive=(np.array(elems), 2.5, 7.0, 0.3))
    assert op.n_dof == oop.n_dof
    y0 = np.cos(0.2 + 0.013 * np.arange(op.n_dof))
    want = oop.lserk(y0, 0.01, 40)
    for form in FORMS64:
        s = R.TdSession(op, form=form)
        assert s.gop.n_dispersive == len(elems)
        got = s.gop.transient(s.ctx, y0, 0.01, 40, 1, record=False, precision="f64")
        err = rel_l2(got, want)
        print(f"builder+debye {form}: rel_l2 {err:.3e}")
        assert err < (1e-7 if form == "fused_f64c" else 1e-11)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("form", ["fused_f64", "fused_f64c"])
def test_every_kernel_variant_matches_oracle(ctx, oracle, order, form):
    """Every fused kernel variant (the autotuner may pick any) against the
    oracle: one operator apply and 6 LSERK4 steps with ports, PEC walls,
    materials and conductivity."""
    xyz, tets = _mesh(oracle, 3)
    K = tets.shape[0]
    eps = np.ones((K, 3))
    eps[::3] = 4.0
    sig = np.zeros(K)
    sig[::5] = 0.3
    port = np.full(4 * K, -1, np.int32)
    port[0:40:4] = 0  # local face 0 of the first tets, kept only where it is a boundary face
    oop = oracle.Operator(xyz, tets, order, 1.0, eps=eps, sigma=sig, face_port=port)
    n = oop.n_dof
    y = np.sin(0.3 + 0.017 * np.arange(n))
    want_apply = oop.apply(y)
    dt = 0.01
    want = oop.lserk(y, dt, 6)
    from rapidfem_amd import RfhipError

    g = _gop(ctx, oop, form)
    ran = 0
    for v in range(g.n_variants()):
        if not g.variant_available(v):
            # the one variant limited to Np <= 20 is refused, not run as another
            assert order == 4 and v == 3
            with pytest.raises(RfhipError):
                g.set_variant(v)
            continue
        ran += 1
        g.set_variant(v)
        assert g.variant == v
        # NaN in every CU's LDS first: a variant that reads shared memory it
        # did not write cannot pass on what the previous variant left there.
        g.poison_lds()
        e_apply = rel_l2(g.apply_f64(y), want_apply)
        g.poison_lds()
        g.state_load(y)
        g.state_lserk(dt, 6)
        e_step = rel_l2(g.state_store(), want)
        print(f"variant order={order} form={form} v={v} apply {e_apply:.3e} lserk {e_step:.3e}", flush=True)
        assert e_apply < 1e-12
        assert e_step < (1e-7 if form == "fused_f64c" else 1e-12)
    assert ran == g.n_variants() - (1 if order == 4 else 0)


@pytest.mark.parametrize("form", ["fused_f64", "fused_f64c"])
def test_variants_on_a_mesh_larger_than_the_resident_grid(ctx, oracle, form):
    """18x18x18 cells x 6 = 34,992 tets at N=3: more element groups than
    blocks resident on the device, so grid-stride variants loop."""
    xyz, tets = oracle.box_mesh(18, 18, 18, 1.0, 1.0, 1.0, 0.1, 3, True)
    oop = oracle.Operator(xyz, tets, 3, 1.0)
    n = oop.n_dof
    y = np.cos(0.1 + 0.013 * np.arange(n))
    want_apply = oop.apply(y)
    want = oop.lserk(y, 0.002, 3)
    g = _gop(ctx, oop, form)
    for v in range(g.n_variants()):
        assert g.variant_available(v)
        g.set_variant(v)
        g.poison_lds()
        e_apply = rel_l2(g.apply_f64(y), want_apply)
        g.poison_lds()
        g.state_load(y)
        g.state_lserk(0.002, 3)
        e_step = rel_l2(g.state_store(), want)
        print(f"large-mesh variant form={form} v={v} apply {e_apply:.3e} lserk {e_step:.3e}", flush=True)
        assert e_apply < 1e-12
        assert e_step < (1e-7 if form == "fused_f64c" else 1e-12)


@pytest.mark.parametrize("form", ALL_FORMS)
def test_probe_recording_matches_full_record(ctx, oracle, form):
    """rfhip_op_transient_probes_f64 (probe rows gathered on the device)
    equals the full-state trajectory at the probe DOFs, frame by frame, and
    the final state equals the oracle's."""
    from rapidfem_amd import Drive

    oop, y0 = _te101(oracle, 3, 2)
    g = _gop(ctx, oop, form)
    n = oop.n_dof
    steps, sub, dt = 4, 3, 0.06
    wave = np.cos(0.5 * np.arange(steps * sub))
    probes = np.array([0, 5, n // 2, n - 1, 17, 17], np.int64)
    d = Drive.point(n // 3)
    traj = g.transient(ctx, y0, dt, steps, sub, d, wave, True, precision="f64")
    rows, yf = g.transient_probes(ctx, y0, dt, steps, probes, sub, d, wave)
    assert rows.shape == (steps + 1, probes.size)
    assert np.array_equal(rows, traj[:, probes])
    assert np.array_equal(yf, traj[-1])
    y = y0
    for k in range(steps):
        y = oop.lserk(y, dt / sub, sub, ("point", n // 3), wave[k * sub:(k + 1) * sub])
    tol = {"plain_f32": 1e-3, "plain_f64": 1e-12, "fused_f64": 1e-12, "fused_f64c": 1e-7}[form]
    assert rel_l2(yf, y) < tol


def test_probe_recording_rejects_bad_dofs(ctx, oracle):
    from rapidfem_amd import RfhipError

    oop, y0 = _te101(oracle, 2, 1)
    g = _gop(ctx, oop, "fused_f64")
    with pytest.raises(RfhipError) as ei:
        g.transient_probes(ctx, y0, 0.01, 2, [0, oop.n_dof])
    assert ei.value.code == 1


@pytest.mark.parametrize("form", ALL_FORMS)
def test_port_signals_on_device_match_oracle(ctx, oracle, form):
    """rfhip_op_transient_ports_f64: (P_e, P_h) of two ports recorded per
    frame on the device against the oracle's projections (rhs.rs:1211-1242)
    of the oracle trajectory; ports on the x = 0 and x = 1 walls, point
    drive."""
    from rapidfem_amd import Drive, maxwell

    oop, y0 = _te101(oracle, 3, 2)
    xyz, tets = _mesh(oracle, 3)
    mop = maxwell.MaxwellOperator(maxwell.Mesh(xyz, tets), 2, 1.0)
    g = _gop(ctx, oop, form)

    def prof(x):
        et = np.stack([0 * x[:, 0], np.sin(np.pi * x[:, 2]), 0 * x[:, 0]], axis=1)
        ht = np.stack([0 * x[:, 0], 0 * x[:, 0], np.sin(np.pi * x[:, 2])], axis=1)
        return et, ht

    entries = []
    for port, val in enumerate((0.0, 1.0)):
        faces = maxwell.boundary_faces_on_plane(mop, 0, val)
        elem, lf, pr, w = maxwell.port_face_data(mop, faces, prof)
        g.set_port(port, elem, lf, pr, w)
        dof = ((elem[:, None].astype(np.int64) * mop.re.n_nodes + mop.re.face_nodes[lf]) * 6).reshape(-1)
        entries.append((dof, pr, w))
    assert g.n_ports == 2
    n = oop.n_dof
    steps, sub, dt = 4, 3, 0.06
    wave = np.cos(0.5 * np.arange(steps * sub))
    sig, yf = g.transient_ports(ctx, y0, dt, steps, sub, Drive.point(n // 3), wave)
    assert sig.shape == (steps + 1, 2, 2)
    y = y0
    tol = {"plain_f32": 1e-3, "plain_f64": 1e-11, "fused_f64": 1e-11, "fused_f64c": 1e-6}[form]
    for k in range(steps + 1):
        if k > 0:
            y = oop.lserk(y, dt / sub, sub, ("point", n // 3), wave[(k - 1) * sub:k * sub])
        for p, (dof, pr, w) in enumerate(entries):
            want = np.array(oracle.port_projections(y, dof, pr, w))
            scale = max(np.abs(want).max(), 1e-3)
            assert np.abs(sig[k, p] - want).max() < tol * scale, (k, p, sig[k, p], want)
    assert rel_l2(yf, y) < (1e-3 if form == "plain_f32" else 1e-6)


# ---- Debye dispersive block (rhs.rs:469-500, 963-985, 1295-1324, 1456-1472)

DEBYE = dict(eps_inf=2.0, eps_s=6.0, tau=0.3)


def _debye_pair(ctx, oracle, form, order=2, n=3, every=2, start=0):
    """Oracle operator with every `every`-th element Debye (eps = eps_inf
    there, 1 elsewhere) and the GPU operator on the same data."""
    xyz, tets = _mesh(oracle, n)
    K = tets.shape[0]
    elems = np.arange(start, K, every)
    eps = np.ones((K, 3))
    eps[elems] = DEBYE["eps_inf"]
    oop = oracle.Operator(xyz, tets, order, 1.0, eps=eps,
                          dispersive=(elems, DEBYE["eps_inf"], DEBYE["eps_s"], DEBYE["tau"]))
    from rapidfem_amd import DebyeMaterial, GpuOperator

    a, gg = DebyeMaterial(DEBYE["eps_inf"], DEBYE["eps_s"], DEBYE["tau"]).relaxation_coeffs()
    d, arrs, re_, ex = oracle_desc(oop)
    g = GpuOperator.from_desc(ctx, d, form, mass=re_["mass"], det=ex["det"],
                              dispersive=(elems, a, gg, 1.0 / DEBYE["eps_inf"]))
    return oop, g, elems


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("form", ALL_FORMS)
def test_debye_apply_matches_oracle(ctx, oracle, order, form):
    """The augmented apply (E rows with -Pdot/eps_inf, appended P rows) on a
    mesh where every second element is Debye."""
    oop, g, elems = _debye_pair(ctx, oracle, form, order)
    n = oop.n_dof
    assert g.n_dof() == n == 6 * oop.K * g_np(order) + 3 * len(elems) * g_np(order)
    assert g.n_dispersive == len(elems)
    y = np.sin(0.3 + 0.017 * np.arange(n))
    want = oop.apply(y)
    if form == "plain_f32":
        got, tol = g.apply(ctx, y.astype(np.float32)), 1e-3
    else:
        got, tol = g.apply_f64(y), 1e-12
    err = rel_l2(got, want)
    eh = 6 * oop.K * g_np(order)
    err_p = rel_l2(got[eh:], want[eh:])
    print(f"debye apply order={order} form={form} rel_l2={err:.3e} P block {err_p:.3e}")
    assert err < tol and err_p < tol


def g_np(order):
    return (order + 1) * (order + 2) * (order + 3) // 6


@pytest.mark.parametrize("form", ALL_FORMS)
@pytest.mark.parametrize("drive", ["free", "vector", "point_in_p"])
def test_debye_transient_matches_oracle(ctx, oracle, form, drive):
    """LSERK4 transient of the augmented system (the fused forms complete
    each stage with the Debye stage kernel) against the oracle, every kernel
    variant of the fused forms; the drive reaches the P block too."""
    from rapidfem_amd import Drive

    oop, g, elems = _debye_pair(ctx, oracle, form, 2)
    n = oop.n_dof
    eh = 6 * oop.K * g_np(2)
    y0 = np.cos(0.2 + 0.013 * np.arange(n))
    steps, sub, dt = 2, 3, 0.06
    h = dt / sub
    wave = np.cos(0.7 * np.arange(steps * sub))
    if drive == "free":
        d, od, gg = Drive.free(), None, None
    elif drive == "vector":
        b = np.sin(np.arange(n) * 0.01)
        d, od, gg = Drive.vector(b), ("vector", b), wave
    else:
        dof = eh + 7
        d, od, gg = Drive.point(dof), ("point", dof), wave
    want = [y0]
    y = y0
    for k in range(steps):
        y = oop.lserk(y, h, sub, od, None if gg is None else gg[k * sub:(k + 1) * sub])
        want.append(y)
    tol = {"plain_f32": 1e-3, "plain_f64": 1e-12, "fused_f64": 1e-12, "fused_f64c": 1e-7}[form]
    prec = "f32" if form == "plain_f32" else "f64"
    for v in range(g.n_variants()):
        if not g.variant_available(v):
            continue
        g.set_variant(v)
        traj = g.transient(ctx, y0, dt, steps, sub, d, gg, True, precision=prec)
        for k in range(steps + 1):
            err = rel_l2(traj[k], want[k])
            err_p = rel_l2(traj[k][eh:], want[k][eh:])
            assert err < tol and err_p < tol, (form, drive, v, k, err, err_p)
        print(f"debye transient form={form} drive={drive} v={v} final rel_l2 {err:.3e} P {err_p:.3e}")


@pytest.mark.parametrize("form", FORMS64)
def test_debye_kcl_matches_oracle(ctx, oracle, form):
    """KCL + PI controller on the augmented system: same accept/reject
    counts and final state as the oracle."""
    from rapidfem_amd import Controller

    oop, g, _ = _debye_pair(ctx, oracle, form, 2)
    y0 = np.cos(0.2 + 0.013 * np.arange(oop.n_dof))
    ctl = Controller(atol=1e-8, rtol=1e-6)
    traj, st = g.transient_kcl(ctx, y0, 0.05, 3, ctl=ctl, precision="f64")
    octl = oracle.Controller.default()
    octl.atol, octl.rtol = 1e-8, 1e-6
    y, acc, rej, hmin, hmax = oop.adaptive(y0, 0.05, 3, octl)
    err = rel_l2(traj[-1], y)
    print(f"debye kcl form={form}: gpu {st.accepted}/{st.rejected} oracle {acc}/{rej} rel_l2 {err:.3e}")
    assert (st.accepted, st.rejected) == (acc, rej)
    assert err < (1e-7 if form == "fused_f64c" else 1e-11)


def test_debye_kcl_compressed_registers_match_oracle(ctx, oracle):
    """Compressed KCL registers on the augmented system at N=4 with 81 Debye
    elements: n = 42,525 is odd, so the bf16 halves of y and its backup end
    on a padded word (the rejection restore copies whole words)."""
    from rapidfem_amd import Controller

    oop, g, elems = _debye_pair(ctx, oracle, "fused_f64c", 4)
    n = oop.n_dof
    assert n % 2 == 1 and len(elems) == 81
    g.set_kcl_registers("compressed")
    y0 = np.cos(0.2 + 0.013 * np.arange(n))
    ctl = Controller(atol=1e-9, rtol=1e-7)
    traj, st = g.transient_kcl(ctx, y0, 0.05, 2, ctl=ctl, precision="f64")
    octl = oracle.Controller.default()
    octl.atol, octl.rtol = 1e-9, 1e-7
    y, acc, rej, _, _ = oop.adaptive(y0, 0.05, 2, octl)
    err = rel_l2(traj[-1], y)
    err_p = rel_l2(traj[-1][6 * oop.K * 35:], y[6 * oop.K * 35:])
    print(f"debye kcl compressed N=4: gpu {st.accepted}/{st.rejected} oracle {acc}/{rej} rel_l2 {err:.3e} "
          f"P {err_p:.3e}", flush=True)
    assert (st.accepted, st.rejected) == (acc, rej)
    assert err < 1e-7 and err_p < 1e-7


def test_compressed_krylov_basis_on_an_odd_length(ctx, oracle):
    """The adaptive exponential with its finished basis columns stored as f32
    + bf16 on the augmented Debye system at N=4 (n = 42,525 is odd: the
    sweeps take their one-entry path and the bf16 halves of the columns are
    padded to whole words) against the oracle's fixed-dimension expmv over
    fine steps: within the tolerance, as with the FP64 basis."""
    from rapidfem_amd.proposer import ExpController

    oop, g, elems = _debye_pair(ctx, oracle, "fused_f64", 4)
    n = oop.n_dof
    assert n % 2 == 1
    y0 = np.cos(0.2 + 0.013 * np.arange(n))
    t_end, tol = 0.05, 1e-7
    yf = y0.copy()
    for _ in range(10):
        yf, _ = oop.expmv(yf, t_end / 10, 40, 0.0)
    ctl = ExpController(tol=tol, m=16)
    out64, st64, _ = g.expmv_adaptive(ctx, y0, t_end, ctl)
    g.set_krylov_basis("compressed")
    out, st, _ = g.expmv_adaptive(ctx, y0, t_end, ctl)
    e, e64 = rel_l2(out, yf), rel_l2(out64, yf)
    print(f"compressed Krylov basis, Debye N=4 n={n}: {len(st)} steps (FP64 basis {len(st64)}), vs fine "
          f"{e:.2e} (FP64 basis {e64:.2e})", flush=True)
    assert len(st) >= 2 and e < 2 * tol and e64 < 2 * tol


@pytest.mark.parametrize("form", ["plain_f64", "fused_f64"])
def test_debye_expmv_matches_oracle(ctx, oracle, form):
    """Krylov exp(tA)v of the augmented operator (the P block is ordinary
    state to the propagator, dispersive.rs:20-23)."""
    oop, g, _ = _debye_pair(ctx, oracle, form, 2)
    y0 = np.cos(0.2 + 0.013 * np.arange(oop.n_dof))
    want, wdim = oop.expmv(y0, 0.3, 30, 1e-10)
    got = g.expmv(ctx, y0, 0.3, 30, 1e-10)
    err = rel_l2(got, want)
    print(f"debye expmv form={form}: dim {g.last_dim}/{wdim} rel_l2 {err:.3e}")
    assert g.last_dim == wdim
    assert err < 1e-10


def test_debye_rejects_bad_element_lists(ctx, oracle):
    """rhs.rs:970-974: an element out of range or listed twice is an error."""
    from rapidfem_amd import GpuOperator, RfhipError

    xyz, tets = _mesh(oracle, 2)
    oop = oracle.Operator(xyz, tets, 1, 1.0)
    d, arrs, re_, ex = oracle_desc(oop)
    for bad in ([0, oop.K], [3, 3], [-1]):
        with pytest.raises(RfhipError) as ei:
            GpuOperator.from_desc(ctx, d, "fused_f64", dispersive=(bad, -1.0, 1.0, 0.5))
        assert ei.value.code == 1


# ---- port source assembled on the device (rhs.rs:1157-1199) ----------------

def _port_case(ctx, oracle, form):
    """Port faces on the z = 0 wall (tagged as ports in the operator),
    element-varying eps and mu, random mode profiles."""
    xyz, tets = oracle.box_mesh(3, 3, 2, 1.0, 1.0, 1.0, 0.1, 4, True)
    K = tets.shape[0]
    ex0 = oracle.Operator(xyz, tets, 2, 1.0).export()
    faces = [(e, f) for e in range(K) for f in range(4)
             if ex0["nbr"][e, f] < 0 and ex0["normal"][e, f, 2] < -0.99]
    port = np.full(4 * K, -1, np.int32)
    for e, f in faces:
        port[4 * e + f] = 0
    eps = 1.0 + 0.5 * (np.arange(K) % 3)[:, None] * np.ones((1, 3))
    mu = 1.0 + 0.25 * (np.arange(K) % 2)[:, None] * np.ones((1, 3))
    oop = oracle.Operator(xyz, tets, 2, 1.0, eps=eps, mu=mu, face_port=port)
    g = _gop(ctx, oop, form)
    prof = np.random.default_rng(11).standard_normal((len(faces), 6, 6))
    fa = np.array(faces)
    g.set_port(0, fa[:, 0], fa[:, 1], prof, np.ones((len(faces), 6)))
    return oop, g, faces, prof


@pytest.mark.parametrize("form", ALL_FORMS)
def test_port_source_on_device_matches_oracle(ctx, oracle, form):
    """rfhip_op_port_source_f64 against the oracle's port_source (itself
    pinned by the ghost-state apply): every entry, all forms."""
    oop, g, faces, prof = _port_case(ctx, oracle, form)
    want = oracle.port_source(oop, faces, prof)
    got = g.port_source(0)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"port source form={form}: max rel err {err:.3e}")
    assert err < (1e-6 if form == "plain_f32" else 1e-13)


@pytest.mark.parametrize("form", FORMS64)
def test_resident_port_source_drives_like_the_uploaded_pattern(ctx, oracle, form):
    """A transient driven by the device-resident port source equals the same
    run with the pattern passed from the host, and the oracle's run with the
    oracle's port_source; passing a host pattern afterwards replaces it."""
    from rapidfem_amd import Drive

    oop, g, faces, prof = _port_case(ctx, oracle, form)
    b = g.port_source(0)
    y0 = np.cos(0.2 + 0.013 * np.arange(oop.n_dof))
    steps, sub, dt = 3, 2, 0.04
    wave = np.sin(0.9 * np.arange(steps * sub))
    res = g.transient(ctx, y0, dt, steps, sub, Drive.port_source(), wave, True, precision="f64")
    up = g.transient(ctx, y0, dt, steps, sub, Drive.vector(b), wave, True, precision="f64")
    assert np.abs(res - up).max() == 0.0
    want = y0
    bo = oracle.port_source(oop, faces, prof)
    for k in range(steps):
        want = oop.lserk(want, dt / sub, sub, ("vector", bo), wave[k * sub:(k + 1) * sub])
    err = rel_l2(res[-1], want)
    print(f"resident port drive form={form}: rel_l2 {err:.3e}")
    assert err < (1e-7 if form == "fused_f64c" else 1e-11)
    # the host pattern given above replaced the resident one
    from rapidfem_amd import RfhipError
    with pytest.raises(RfhipError):
        g.transient(ctx, y0, dt, 1, 1, Drive.port_source(), wave[:1], True, precision="f64")


# ---- graph-captured resident stepping ----------------------------------------

@pytest.mark.parametrize("form", ALL_FORMS + ["fused_f64c_debye"])
def test_graph_stepping_equals_eager_stepping(ctx, oracle, form):
    """rfhip_op_state_lserk_graph replays a captured two-step graph; with
    eager work, reads of the state and a variant change between replays,
    every replay must equal the eager stepper bit for bit (odd step counts
    exercise the eager tail and both starting buffers)."""
    if form.endswith("_debye"):
        oop, g, _ = _debye_pair(ctx, oracle, "fused_f64c", 2)
        g2 = _debye_pair(ctx, oracle, "fused_f64c", 2)[1]
    else:
        oop, _ = _te101(oracle, 3, 2)
        g, g2 = _gop(ctx, oop, form), _gop(ctx, oop, form)
    y0 = np.cos(0.2 + 0.013 * np.arange(oop.n_dof))
    g.state_load(y0)
    g2.state_load(y0)
    for rep, (steps, variant) in enumerate([(4, None), (3, None), (5, None), (2, 1), (3, 1)]):
        if variant is not None and g.n_variants() > variant:
            g.set_variant(variant)
            g2.set_variant(variant)
        g.state_lserk_graph(0.01, steps)
        g2.state_lserk(0.01, steps)
        a, b = g.state_store(), g2.state_store()
        print(f"graph form={form} replay {rep}: max |diff| {np.abs(a - b).max():.3e}")
        assert np.array_equal(a, b), rep


# ---- sparse state-space export (rhs.rs:1575-1730) ---------------------------

def _sparse_oracle_case(oracle, order, n=2, jitter=True, alpha=0.6, debye_every=0, ports=True, materials=True):
    """Oracle operator with PEC walls, port faces on x = 0, per-element
    anisotropic materials with both conductivities, optional Debye elements;
    returns it with the GPU-side dispersive tuple."""
    xyz, tets = _mesh(oracle, n, jitter)
    K = tets.shape[0]
    port = np.full(4 * K, -1, np.int32)
    if ports:
        ex0 = oracle.Operator(xyz, tets, 1, 1.0).export()
        for e in range(K):
            for f in range(4):
                if ex0["nbr"][e, f] < 0 and ex0["normal"][e, f, 0] < -0.99:
                    port[4 * e + f] = 0
    kw = {}
    if materials:
        rng = np.random.default_rng(5)
        kw = dict(eps=1.0 + rng.random((K, 3)), mu=1.0 + rng.random((K, 3)),
                  sigma=np.where(np.arange(K) % 3 == 0, 0.4, 0.0),
                  sigma_m=np.where(np.arange(K) % 4 == 1, 0.3, 0.0))
    disp = None
    if debye_every:
        elems = np.arange(1, K, debye_every)
        ei, es, tau = DEBYE["eps_inf"], DEBYE["eps_s"], DEBYE["tau"]
        kw["dispersive"] = (elems, ei, es, tau)
        disp = (elems, -1.0 / tau, (es - ei) / tau, 1.0 / ei)
    oop = oracle.Operator(xyz, tets, order, alpha, face_port=port, **kw)
    return oop, disp


def _assert_csr_equal(got, want):
    rp, ci, va = want
    assert got.n == len(rp) - 1
    assert np.array_equal(got.row_ptr, rp)
    assert np.array_equal(got.col_idx, ci)
    assert np.array_equal(got.values, va), f"max |diff| {np.abs(got.values - va).max()}"


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("debye", [0, 3])
def test_sparse_export_matches_oracle_bit_for_bit(ctx, oracle, order, debye):
    """Every branch of the flux (PEC, port, interior), materials with both
    conductivities, upwind blend 0.6, and Debye elements: row_ptr, col_idx
    and values equal the oracle's unit-column probe exactly."""
    from rapidfem_amd import assemble_sparse

    oop, disp = _sparse_oracle_case(oracle, order, debye_every=debye)
    d, arrs, _, _ = oracle_desc(oop)
    print(f"sparse export order={order} debye={debye} K={oop.K}", flush=True)
    got = assemble_sparse(ctx, desc=d, dispersive=disp)
    _assert_csr_equal(got, oop.assemble_sparse())
    if debye:
        assert got.n == oop.n_dof and oop.n_disp > 0


@pytest.mark.parametrize("order,alpha", [(2, 0.0), (3, 1.0), (4, 0.0)])
def test_sparse_export_exact_zeros_on_an_unjittered_box(ctx, oracle, order, alpha):
    """Axis-aligned faces and a central flux make many products exactly
    zero; the stored pattern must drop exactly the entries the oracle drops."""
    from rapidfem_amd import assemble_sparse

    oop, _ = _sparse_oracle_case(oracle, order, n=2, jitter=False, alpha=alpha, ports=False, materials=False)
    d, arrs, _, _ = oracle_desc(oop)
    print(f"sparse export unjittered order={order} alpha={alpha}", flush=True)
    got = assemble_sparse(ctx, desc=d)
    _assert_csr_equal(got, oop.assemble_sparse())


def test_sparse_export_at_scale_matches_oracle_and_apply(ctx, oracle):
    """1296 tets at N=3 (n = 155,520, 43.6M nonzeros): equal to the oracle,
    within the 5 * stride * n stencil bound (rhs.rs:2288-2330), and CSR x
    equals the GPU operator's own FP64 apply (rhs.rs:2259-2286)."""
    import scipy.sparse as sp

    from rapidfem_amd import assemble_sparse

    oop, disp = _sparse_oracle_case(oracle, 3, n=6, debye_every=5)
    d, arrs, re_, ex = oracle_desc(oop)
    print(f"sparse export at scale K={oop.K} n={oop.n_dof}", flush=True)
    got = assemble_sparse(ctx, desc=d, dispersive=disp)
    _assert_csr_equal(got, oop.assemble_sparse())
    n = got.n
    assert got.nnz() <= 5 * 6 * 20 * n
    from rapidfem_amd import GpuOperator

    g = GpuOperator.from_desc(ctx, d, "fused_f64", dispersive=disp)
    v = np.cos(1.0 + np.arange(n) * 0.07)
    a = sp.csr_matrix((got.values, got.col_idx, got.row_ptr), shape=(n, n))
    assert rel_l2(a @ v, g.apply_f64(v)) < 1e-12
    assert got.count_ms > 0.0 and got.fill_ms > 0.0


def test_sparse_export_rejects_bad_input(ctx, oracle):
    from rapidfem_amd import RfhipError, assemble_sparse

    oop, _ = _sparse_oracle_case(oracle, 2)
    d, arrs, _, _ = oracle_desc(oop)
    with pytest.raises(RfhipError):
        assemble_sparse(ctx, desc=d, dispersive=([1, 1], -1.0, 1.0, 0.5))
    with pytest.raises(RfhipError):
        assemble_sparse(ctx, desc=d, dispersive=([oop.K], -1.0, 1.0, 0.5))
    perm = arrs["face_perm"].reshape(-1)
    q = int(np.flatnonzero((arrs["face_neighbor"].reshape(-1) >= 0) & (arrs["face_port"].reshape(-1) < 0))[0])
    nfp = perm.size // (4 * oop.K)
    old = perm[q * nfp]
    perm[q * nfp] = nfp  # out of range: refused on the host, before any launch
    try:
        with pytest.raises(RfhipError):
            assemble_sparse(ctx, desc=d)
    finally:
        perm[q * nfp] = old


def test_product_builder_sparse_export_matches_oracle(ctx, oracle):
    """The package's own host builder (MaxwellOperator with materials, a
    port and Debye elements) into the device export: the same CSR as the
    oracle's probe of the oracle's operator, bit for bit."""
    import rapidfem_amd as R

    mesh = R.structured_box_jittered(3, 3, 2, 1.0, 1.0, 0.7, 0.1, 9)
    K = mesh.n_tets
    rng = np.random.default_rng(2)
    eps = rng.uniform(1, 4, (K, 3))
    mu = rng.uniform(1, 2, (K, 3))
    sig = rng.uniform(0, 0.5, K)
    port = np.full((K, 4), -1, np.int32)
    port[:9, 1] = 0
    elems = [1, 4, 10]
    deb = R.DebyeMaterial(DEBYE["eps_inf"], DEBYE["eps_s"], DEBYE["tau"])
    eps[elems] = DEBYE["eps_inf"]
    op = R.MaxwellOperator(mesh, 3, 0.8, materials={"eps": eps, "mu": mu, "sigma": sig}, face_port=port,
                           dispersive=[(e, deb) for e in elems])
    oop = oracle.Operator(mesh.nodes, mesh.tets, 3, 0.8, eps=eps, mu=mu, sigma=sig, face_port=port.ravel(),
                          dispersive=(np.array(elems), DEBYE["eps_inf"], DEBYE["eps_s"], DEBYE["tau"]))
    print(f"product sparse export K={K} n={op.n_dof}", flush=True)
    got = R.assemble_sparse(ctx, op)
    _assert_csr_equal(got, oop.assemble_sparse())


def test_cfl_step_matches_oracle_on_the_cavity(ctx, oracle):
    """TdSession::cfl_dt (session.rs:514-569) over the GPU operator against the
    oracle's restatement on the 1,296-tet N=2 jittered cavity: the same
    power iteration from the same seeded vector and the same bracket and
    bisection give the same step (the GPU and CPU applies differ only by
    round-off, which moves rho in the last digits)."""
    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    op, _ = W.cavity_te101(6, 2)
    oop = oracle.Operator(op.mesh.nodes, op.mesh.tets, 2, 1.0)
    want = oop.cfl_dt()
    got = R.TdSession(op, 1.0, "fused_f64", ctx).cfl_dt()
    print(f"cfl_dt gpu {got:.15e} oracle {want:.15e}", flush=True)
    assert abs(got - want) <= 1e-10 * want


def test_batch_lserk_equals_individual_runs(ctx, oracle):
    """Operators of different meshes, orders and forms (two of them on one
    mesh, as two right-hand sides; one with Debye elements) stepped together
    from one graph end exactly where each operator's own run ends; an odd
    step count exercises the step after the graph replays."""
    import rapidfem_amd as R

    def make(order, n, form, seed, debye=False):
        xyz, tets = _mesh(oracle, n)
        K = tets.shape[0]
        kw, disp = {}, None
        if debye:
            elems = np.arange(0, K, 3)
            kw["dispersive"] = (elems, DEBYE["eps_inf"], DEBYE["eps_s"], DEBYE["tau"])
            a, gg = R.DebyeMaterial(DEBYE["eps_inf"], DEBYE["eps_s"], DEBYE["tau"]).relaxation_coeffs()
            disp = (elems, a, gg, 1.0 / DEBYE["eps_inf"])
        oop = oracle.Operator(xyz, tets, order, 1.0, **kw)
        d, arrs, re_, ex = oracle_desc(oop)
        pair = [R.GpuOperator.from_desc(ctx, d, form, dispersive=disp) for _ in range(2)]
        y0 = np.random.default_rng(seed).standard_normal(oop.n_dof) * 0.1
        return pair, y0

    cases = [make(2, 3, "fused_f64", 1), make(3, 2, "fused_f64c", 2), make(2, 3, "plain_f64", 3),
             make(2, 3, "fused_f64", 4), make(1, 3, "fused_f64", 5, debye=True)]
    dts = [0.01, 0.008, 0.01, 0.012, 0.01]
    batch = [pair[0] for pair, _ in cases]
    for (pair, y0) in cases:
        for g in pair:
            g.state_load(y0)
    print("batch of 5 operators, 7 steps", flush=True)
    R.batch_lserk(batch, dts, 7)
    for (pair, _), dt in zip(cases, dts):
        pair[1].state_lserk(dt, 7)
        assert np.array_equal(pair[0].state_store(), pair[1].state_store())
    with pytest.raises(R.RfhipError):
        R.batch_lserk([batch[0], batch[0]], [0.01, 0.01], 2)
