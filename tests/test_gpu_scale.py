"""Properties at the full workload sizes, where the oracle cannot follow in
seconds: the 324,000-tet N=3 pulse box (38.88M DOFs), the 998,250-tet N=4
jittered modes box (209.6M DOFs) and a 201,600-tet N=3 box for the Krylov
exponential (24.2M DOFs), uniform and graded toward a lossy slab.

Size-independent checks only: two independent GPU implementations of the
operator (the plain FP64 form, which keeps the reference kernel structure,
and the fused forms) agree; the operator is linear; the discrete energy
never grows under LSERK4 and the upwind flux dissipates more than the
central one; exp(-tA) exp(tA) v = v and exp(tA/2)^2 v = exp(tA) v.
Tolerances: 1e-12 relative L2 between FP64 forms, 1e-7 for the compressed
state form (f32 hi + bf16 lo registers hold 2^-32 of each value).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from rapidfem_amd import GpuContext

    return GpuContext()


@pytest.fixture(scope="module")
def pulse_box():
    from rapidfem_amd import workloads as W

    print("building the 324,000-tet N=3 pulse box", flush=True)
    return W.pulse_box(60, 30, 30, 3)


def test_pulse_box_fused_forms_equal_plain_f64_and_are_linear(ctx, pulse_box):
    import rapidfem_amd as R

    op, y0 = pulse_box
    n = op.n_dof
    rng = np.random.default_rng(21)
    y = y0 + 0.1 * rng.standard_normal(n)
    z = rng.standard_normal(n)
    plain = R.GpuOperator(ctx, op, "plain_f64")
    want = plain.apply_f64(y)
    del plain
    fused = R.GpuOperator(ctx, op, "fused_f64")
    got = fused.apply_f64(y)
    print(f"pulse box apply fused_f64 vs plain_f64 {rel_l2(got, want):.3e}", flush=True)
    assert rel_l2(got, want) < 1e-12
    lhs = fused.apply_f64(0.7 * y - 1.3 * z)
    rhs = 0.7 * got - 1.3 * fused.apply_f64(z)
    assert rel_l2(lhs, rhs) < 1e-12
    del fused
    comp = R.GpuOperator(ctx, op, "fused_f64c").apply_f64(y)
    print(f"pulse box apply fused_f64c vs plain_f64 {rel_l2(comp, want):.3e}", flush=True)
    assert rel_l2(comp, want) < 1e-7


def test_state_round_trip_through_the_staging_ring(ctx, pulse_box):
    """38.88M DOFs of f64 (311 MB: four full 64 MB pinned slots and a partial
    one) uploaded and downloaded bit for bit, twice in a row so the second
    transfer reuses slots the first one left."""
    import rapidfem_amd as R

    op, y0 = pulse_box
    rng = np.random.default_rng(8)
    g = R.GpuOperator(ctx, op, "fused_f64")
    for _ in range(2):
        y = rng.standard_normal(op.n_dof)
        g.state_load(y)
        back = g.state_store()
        assert back.dtype == np.float64 and back.shape == y.shape
        assert np.array_equal(back.view(np.uint64), y.view(np.uint64))


def test_pulse_box_energy_never_grows_and_upwind_dissipates_more(ctx, pulse_box):
    """LSERK4 below the CFL step on M~-skew (central) and M~-dissipative
    (upwind) operators: the energy 0.5 y^T M~ y is non-increasing step after
    step, and the upwind run ends with less of it than the central run."""
    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    op_up, y0 = pulse_box
    op_c, _ = W.pulse_box(60, 30, 30, 3, flux_alpha=0.0)
    dts = []
    for op in (op_up, op_c):
        s = R.TdSession(op, 1.0, "fused_f64", ctx)
        dts.append(s.cfl_dt())
        del s
    dt = min(dts)
    print(f"pulse box energy: dt {dt:.4e} (upwind cfl {dts[0]:.4e}, central {dts[1]:.4e})", flush=True)
    finals = []
    for op in (op_up, op_c):
        g = R.GpuOperator(ctx, op, "fused_f64c")
        g.state_load(y0)
        e = [g.state_energy()]
        for _ in range(8):
            g.state_lserk(dt, 5)
            e.append(g.state_energy())
        assert all(b <= a * (1.0 + 1e-9) for a, b in zip(e, e[1:])), e
        assert e[-1] > 0.5 * e[0]
        finals.append(e[-1] / e[0])
        del g
    print(f"energy kept after 40 steps: upwind {finals[0]:.9f}, central {finals[1]:.9f}", flush=True)
    assert finals[0] < finals[1]


def test_pulse_box_compressed_state_tracks_fp64_over_steps(ctx, pulse_box):
    """The chosen form's precision gate at full size: 40 LSERK4 steps at the
    CFL step against the FP64 fused form, relative L2 below 1e-6."""
    import rapidfem_amd as R

    op, y0 = pulse_box
    s = R.TdSession(op, 1.0, "fused_f64", ctx)
    dt = s.cfl_dt()
    ref = s.gop
    ref.state_load(y0)
    ref.state_lserk(dt, 40)
    want = ref.state_store()
    g = R.GpuOperator(ctx, op, "fused_f64c")
    g.autotune()
    g.state_load(y0)
    g.state_lserk(dt, 40)
    err = rel_l2(g.state_store(), want)
    print(f"pulse box 40 steps f64c (variant {g.variant}) vs f64: {err:.3e}", flush=True)
    assert err < 1e-6


def test_modes_box_n4_fused_equals_plain_f64(ctx):
    """998,250 tets at N=4 (55^3 x 6, jittered), 209.6M DOFs, 16 cavity
    modes: one apply of the fused FP64 form and of the compressed form
    against the plain FP64 form."""
    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    print("building the 998,250-tet N=4 modes box", flush=True)
    op, y0 = W.modes_box(55, 4)
    assert op.n_elem == 998_250
    plain = R.GpuOperator(ctx, op, "plain_f64")
    want = plain.apply_f64(y0)
    del plain
    got = R.GpuOperator(ctx, op, "fused_f64").apply_f64(y0)
    print(f"modes box apply fused_f64 vs plain_f64 {rel_l2(got, want):.3e}", flush=True)
    assert rel_l2(got, want) < 1e-12
    del got
    comp = R.GpuOperator(ctx, op, "fused_f64c").apply_f64(y0)
    print(f"modes box apply fused_f64c vs plain_f64 {rel_l2(comp, want):.3e}", flush=True)
    assert rel_l2(comp, want) < 1e-7


@pytest.mark.parametrize("workload", ["pulse", "slab"])
def test_krylov_exponential_inverse_and_semigroup_at_scale(ctx, workload):
    """201,600 tets at N=3 (24.2M DOFs), the uniform pulse box and the box
    graded 20:1 toward an eps_r 10, sigma 50 slab; Arnoldi dimension 40 over
    t = 3 / rho (rho the spectral radius from the CFL power iteration, so
    ||tA|| is about 3): exp(-tA) exp(tA) v = v and
    exp(tA/2) exp(tA/2) v = exp(tA) v."""
    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    build = W.pulse_box if workload == "pulse" else W.slab_box
    op, y0 = build(40, 30, 28, 3)
    assert op.n_elem == 201_600
    s = R.TdSession(op, 1.0, "fused_f64", ctx)
    s.cfl_dt()
    t = 3.0 / s.rho
    g = s.gop
    w = g.expmv(ctx, y0, t, 40)
    back = g.expmv(ctx, w, -t, 40)
    half = g.expmv(ctx, g.expmv(ctx, y0, 0.5 * t, 40), 0.5 * t, 40)
    e_inv, e_semi = rel_l2(back, y0), rel_l2(half, w)
    print(f"krylov at scale [{workload}]: t {t:.3e}, inverse {e_inv:.3e}, semigroup {e_semi:.3e}", flush=True)
    assert e_inv < 1e-9
    assert e_semi < 1e-9
    assert g.last_call_ms > 0.0
    assert rel_l2(w, y0) > 1e-3  # the field did move
