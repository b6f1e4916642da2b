"""The batching form with port excitations: several meshes, several
right-hand sides per mesh (one operator each, driven by its own port's
TE10-shaped source, port_source rhs.rs:1157-1199, under a Gaussian-modulated
pulse, session.rs:242-261), stepped together from one captured graph
(rfhip_batch_lserk_ports).

  oracle size   three small port boxes at orders 1-3 on three forms, two
                excitations each plus one free operator: every final state
                against the oracle's vector-driven LSERK4 with the same
                source and samples (1e-11 FP64 forms, 1e-7 compressed), and
                against each operator's own port-driven transient;
  C4 size       the 8 seeded boxes of 20k-300k tets at orders 2/3/4 with 4
                excitations each (32 right-hand sides, compressed-register
                form), 600 steps: one operator per mesh run alone ends bit
                for bit where the 32-operator batch leaves it; every state is
                finite and driven, and the four excitations of a mesh differ.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from rapidfem_amd import GpuContext

    return GpuContext()


def _rhs(ctx, op, ports, form, excite):
    import rapidfem_amd as R

    g = R.GpuOperator(ctx, op, form)
    for q, (el, lf, prof, w) in enumerate(ports):
        g.set_port(q, el, lf, prof, w)
    b = g.port_source(excite)
    return g, b


def test_port_batch_matches_oracle_and_single_runs(ctx, oracle):
    import rapidfem_amd as R
    from rapidfem_amd import Drive, GaussianPulse
    from rapidfem_amd import workloads as W

    steps, dt = 7, 0.01  # odd: three graph replays and an eager step
    cases = [((3, 3, 2), 2, "fused_f64", 1), ((2, 3, 3), 3, "fused_f64c", 2), ((3, 2, 2), 1, "plain_f64", 3)]
    batch, dts, waves, checks = [], [], [], []
    for cells, order, form, seed in cases:
        op, ports = W.port_box(*cells, order, seed)
        oop = oracle.Operator(op.mesh.nodes, op.mesh.tets, order, 1.0, eps=1.0 / op.inv_eps,
                              face_port=op.port.reshape(-1))
        for r in (0, 3):
            g, b = _rhs(ctx, op, ports, form, r)
            pulse = GaussianPulse(0.03, 0.02, 3.0 + 2.0 * r + seed)
            w = np.array([pulse(k * dt) for k in range(steps)])
            y0 = 0.01 * np.cos(0.3 + 0.011 * np.arange(op.n_dof) * (r + 1))
            g.state_load(y0)
            batch.append(g)
            dts.append(dt)
            waves.append(w)
            checks.append((g, oop, b, w, y0, form))
    # one free operator in the same graph
    opf, _ = W.port_box(2, 2, 2, 2, 9)
    gf = R.GpuOperator(ctx, opf, "fused_f64")
    yf0 = np.sin(0.2 + 0.01 * np.arange(opf.n_dof))
    gf.state_load(yf0)
    batch.append(gf)
    dts.append(dt)
    waves.append(None)
    print(f"port batch: {len(batch)} operators, {steps} steps", flush=True)
    R.batch_lserk_ports(batch, dts, steps, waves)
    for g, oop, b, w, y0, form in checks:
        got = g.state_store()
        want = oop.lserk(y0, dt, steps, ("vector", b), w)
        err = rel_l2(got, want)
        single = g.transient(ctx, y0, dt, steps, 1, Drive.port_source(), w, False, precision="f64")
        err_s = rel_l2(got, single)
        print(f"  {form} N={oop.order}: vs oracle {err:.2e}, vs own port-driven transient {err_s:.2e}, "
              f"driven part {rel_l2(want, oop.lserk(y0, dt, steps)):.2e}", flush=True)
        tol = 1e-7 if form == "fused_f64c" else 1e-11
        assert err < tol and err_s < tol
        assert rel_l2(want, oop.lserk(y0, dt, steps)) > 1e-4  # the source is not negligible
    gf2 = R.GpuOperator(ctx, opf, "fused_f64")
    gf2.state_load(yf0)
    gf2.state_lserk(dt, steps)
    assert np.array_equal(gf.state_store(), gf2.state_store())


@pytest.mark.timeout(900)
def test_c4_port_batch_32_rhs_equals_single_operator_runs(ctx):
    import rapidfem_amd as R
    from rapidfem_amd import GaussianPulse
    from rapidfem_amd import workloads as W

    steps = 600
    specs = W.port_batch_specs()
    ops, dts, waves, firsts, groups = [], [], [], [], []
    total = 0
    for b, (nx, ny, nz, order) in enumerate(specs):
        op, ports = W.port_box(nx, ny, nz, order, 11 + b)
        cfl = R.TdSession(op, 1.0, "fused_f64c", ctx).cfl_dt()
        pulses = W.port_waveforms(100 + b, 4)
        group = []
        for r in range(4):
            g, _ = _rhs(ctx, op, ports, "fused_f64c", r)
            g.state_load(np.zeros(op.n_dof))
            pulse = GaussianPulse(*pulses[r])
            ops.append(g)
            dts.append(cfl)
            waves.append(np.array([pulse(k * cfl) for k in range(steps)]))
            group.append(g)
        groups.append(group)
        firsts.append((op, ports, cfl, waves[-4], group[0].variant))
        total += 4 * op.n_dof
        print(f"mesh {b}: {nx}x{ny}x{nz} N={order}, {op.n_elem} tets, cfl {cfl:.3e}", flush=True)
    print(f"C4: {len(ops)} right-hand sides, {total} DOFs, {steps} steps", flush=True)
    R.batch_lserk_ports(ops, dts, steps, waves)
    for b, group in enumerate(groups):
        ys = [g.state_store() for g in group]
        norms = [float(np.linalg.norm(y)) for y in ys]
        assert all(np.isfinite(n) and n > 0 for n in norms), norms
        assert min(rel_l2(ys[0], y) for y in ys[1:]) > 1e-3
        op, ports, cfl, w, variant = firsts[b]
        alone, _ = _rhs(ctx, op, ports, "fused_f64c", 0)
        alone.set_variant(variant)
        alone.state_load(np.zeros(op.n_dof))
        R.batch_lserk_ports([alone], [cfl], steps, [w])
        same = np.array_equal(alone.state_store(), ys[0])
        print(f"  mesh {b}: |y| of the 4 excitations {np.round(norms, 4).tolist()}, alone == batched: {same}",
              flush=True)
        assert same
        del alone
