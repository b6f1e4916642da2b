"""Benchmark: DGTD DOF-stage updates/s of the LSERK4 stepper on one MI355X.

Workload: the 2x1x1 box of 60x30x30 cells x 6 tets (324,000 tets), order
N=3 (Np=20, 38.88M DOFs), eps_r=4 block at x in [0.8, 1.2], Gaussian-pulse
E_x, PEC walls, upwind flux, LSERK4 at the CFL step. A step is one LSERK4
step (5 stages) over the whole mesh with the state resident in HBM.

Reported: the chosen form (fused_f64c: FP64 arithmetic, f32+bf16 state and
f32 residual registers) and the plain FP64 form (the reference kernel
structure) on the same state; the chosen form's relative L2 deviation from
the FP64 fused form after the timed steps and from the CPU oracle's
trajectory over the CPU sample's steps; the dominant kernel's roofline from
HIP-event timing, bound by whichever roof its arithmetic intensity reaches;
the CPU oracle (C restatement of the reference CPU path) timed on a bounded
sample on all host threads and on one; and the largest single-GPU stepper
workload (998,250 tets at N=4) beside the headline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "DGTD DOF-stage updates/s + %HBM/MFMA peak, chosen vs plain at 1e-6 rel L2"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fused_stage_pulse_box_traffic.json")
TRAFFIC_FILE_C3 = os.path.join(ROOT, "profiles", "fused_stage_modes_box_n4_traffic.json")


def traffic_entry(path, form, n, variant):
    """HBM bytes per launch of a stage-kernel variant from a committed
    counter summary (FETCH_SIZE, WRITE_SIZE), or None."""
    try:
        tr = json.load(open(path))
    except (OSError, ValueError):
        return None
    if tr.get("form") != form or tr.get("n_dof") != n:
        return None
    return tr.get("variants", {}).get(str(variant), {}).get("bytes_per_launch")


# The arithmetic each form computes in, and how its state registers are stored.
DTYPE_LABEL = {
    "fused_f64": "f64",
    "fused_f64c": "f64 arithmetic; y stored f32+bf16, p f32",
    "fused_bf16x3": "f64 arithmetic, volume contraction bf16x3 MFMA with f32 accumulation; y f32+bf16, p f32",
    "fused_bf16x2": "f64 arithmetic, volume contraction bf16 hi+lo MFMA with f32 accumulation; y f32+bf16, p f32",
    "fused_bf16x4": "f64 arithmetic, volume contraction bf16x4 MFMA with f32 accumulation; y f32+bf16, p f32",
    "plain_f64": "f64",
    "plain_f32": "f32",
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=40)
    ap.add_argument("--cells", type=int, nargs=3, default=[60, 30, 30])
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--form", default="fused_f64c")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--c3-steps", type=int, default=10,
                    help="steps of the 998,250-tet N=4 line reported beside the headline (0 = skip)")
    ap.add_argument("--autotune", type=int, default=1, help="pick the fastest kernel variant on this mesh")
    ap.add_argument("--variant", type=int, default=-1,
                    help="pin this stage-kernel variant instead of autotuning (fused forms)")
    ap.add_argument("--probe-c3", action="store_true",
                    help="with --traffic-probe: the 998,250-tet N=4 cavity instead of the pulse box")
    ap.add_argument("--drift", action="store_true",
                    help="long-run error of every register / precision form against the FP64 fused form")
    ap.add_argument("--drift-steps", type=int, default=5000)
    ap.add_argument("--traffic-probe", action="store_true",
                    help="counter-profiling run: a 1 GiB streaming copy (torch.neg, 8 B read + 8 B written "
                         "per element) as the calibration kernel, then 10 LSERK4 steps of the chosen form at a "
                         "fixed dt (no CFL search), for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--sparse-export", action="store_true",
                    help="time the sparse state-space export (assemble_sparse) on the device instead of the "
                         "stepper: --sparse-cells box at --order, one JSON line in nonzeros/s")
    ap.add_argument("--krylov", action="store_true",
                    help="time the Krylov exponential (expmv, FP64 Arnoldi with CGS2) on the --krylov-cells box "
                         "instead of the stepper: one JSON line in Arnoldi DOF-steps/s")
    ap.add_argument("--krylov-cells", type=int, nargs=3, default=[40, 30, 28])
    ap.add_argument("--krylov-dim", type=int, default=40)
    ap.add_argument("--proposer", action="store_true",
                    help="the learned-proposal form: plain vs learned (h, m) adaptive Krylov exponential on the "
                         "--krylov-cells slab box, final error against a fine-step run, one JSON line")
    ap.add_argument("--cost-only", action="store_true",
                    help="with --proposer: only fit and print the per-step cost model used for training labels")
    ap.add_argument("--proposer-cfl", type=float, default=200.0, help="horizon in CFL steps")
    ap.add_argument("--proposer-tol", type=float, default=1e-6)
    ap.add_argument("--proposer-weights", default="")
    ap.add_argument("--m-sweep", action="store_true",
                    help="with --proposer: also time the plain controller at fixed dimensions 10..60")
    ap.add_argument("--kcl", action="store_true",
                    help="time the KCL adaptive stepper (RK4(3)5 + PI controller) on the stepper workload: one "
                         "JSON line in DOF-stage/s over accepted and rejected substeps")
    ap.add_argument("--record", action="store_true",
                    help="time trajectory recording on the stepper workload: every frame's full state "
                         "downloaded (the reference GPU path) against probe DOFs recorded on the device")
    ap.add_argument("--batch", action="store_true",
                    help="time the batching form: 8 seeded boxes of 20k-300k tets at orders 2/3/4, 4 right-hand "
                         "sides each (32 operators), stepped one after another and then together from one graph")
    ap.add_argument("--batch-steps", type=int, default=3000)
    ap.add_argument("--sparse-cells", type=int, nargs=3, default=[16, 16, 16])
    ap.add_argument("--sparse-reps", type=int, default=3)
    return ap.parse_args()


def apply_bytes_per_dof(order: int) -> float:
    """Fused FP64 apply: y read + dy written (16 B) + per-element operator
    data (92 B over 6 Np DOFs)."""
    np_ = (order + 1) * (order + 2) * (order + 3) // 6
    return 16.0 + 92.0 / (6 * np_)


def krylov_bench(args) -> int:
    """expmv (fixed Arnoldi dimension m, no a-posteriori stop) on the stiff
    Krylov workload: the 201,600-tet N=3 unit box graded 20:1 toward a thin
    slab (eps_r 10, sigma 50), pulse field, t = 3 / rho (rho from 20 power
    iterations of the apply, so ||tA|| is about 3). Unit: one Arnoldi
    DOF-step = one basis vector entry of one Arnoldi step (apply + two CGS2
    passes)."""
    import numpy as np

    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    nx, ny, nz = args.krylov_cells
    op, y0 = W.slab_box(nx, ny, nz, args.order)
    n, m = op.n_dof, args.krylov_dim
    ctx = R.GpuContext()
    g = R.GpuOperator(ctx, op, "fused_f64")
    variant, _ = g.autotune()
    v = y0 / np.linalg.norm(y0)
    rho = 1.0
    for _ in range(20):
        av = g.apply_f64(v)
        rho = float(np.linalg.norm(av))
        v = av / rho
    t = 3.0 / rho
    log(f"krylov: {op.n_elem} tets, N={args.order}, n={n}, m={m}, rho {rho:.4e}, t={t:.4e}")
    times, walls = [], []
    for r in range(4):
        t0 = time.perf_counter()
        g.expmv(ctx, y0, t, m)
        wall = (time.perf_counter() - t0) * 1e3
        log(f"  rep {r}: dim {g.last_dim}, device part {g.last_call_ms:.2f} ms, call {wall:.1f} ms")
        if r > 0:
            times.append(g.last_call_ms)
            walls.append(wall)
    ms = float(np.median(times))
    wall = float(np.median(walls))
    dim = g.last_dim
    # algorithmic HBM bytes of the whole call: per step j (cols = j + 1 basis
    # vectors) the apply, three sweeps of the two-pass Gram-Schmidt (the
    # projections read cols + 1 vectors; the first update fused with the
    # second projections reads cols + 1 and writes 1; the second update
    # reads cols + 1 and writes 1) and the normalisation (read w, write the
    # next basis vector); plus the start (norm, scale) and the final
    # combination (reads dim vectors, writes one)
    b = 3 * 8.0 * n + (dim + 1) * 8.0 * n
    for j in range(dim):
        cols = j + 1
        b += apply_bytes_per_dof(args.order) * n + (3 * cols + 5) * 8.0 * n
        if j + 1 < dim:
            b += 16.0 * n
    gbs = b / (ms * 1e-3) / 1e9
    out = {"metric": "Krylov exponential (expmv, CGS2 Arnoldi) DOF-steps/s", "value": n * dim / (ms * 1e-3),
           "unit": "Arnoldi DOF-steps/s", "n_gpus": 1, "reps": len(times), "higher_is_better": True,
           "dtype": "f64", "data": "synthetic", "ms_per_call": ms, "ms_per_arnoldi_step": ms / dim,
           "host_inclusive": {"ms_per_call": wall, "value": n * dim / (wall * 1e-3),
                              "note": "the whole call: v uploaded and the result downloaded as f64 "
                                      "(194 MB each way) with their layout conversions"},
           "config": {"workload": f"slab box {nx}x{ny}x{nz} x 6 tets graded 20:1 to an eps_r 10, sigma 50 slab, "
                                  f"N={args.order}, m={dim}, t=3/rho", "n_elem": op.n_elem, "n": n, "t": t,
                      "variant": variant},
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                        "algorithmic_bytes_per_call": b,
                        "kernel": "whole call (apply + k_multi_dot + k_axpy_dot + k_axpy_basis_norm + scale + lincomb), "
                                  "host round trip per CGS2 pass included"}}
    if not args.no_cpu:
        from oracle import oracle as O

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.set_threads(threads)
        del g
        oop = O.Operator(op.mesh.nodes, op.mesh.tets, args.order, 1.0, eps=1.0 / op.inv_eps,
                         sigma=op.sigma_eps[:, 0] / op.inv_eps[:, 0])
        t0 = time.perf_counter()
        _, cdim = oop.expmv(y0, t, m)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * cdim / tc, "unit": "Arnoldi DOF-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"oracle expmv (propagator.rs:186-326 restated, OpenMP apply), one call "
                                         f"on the same slab box and t, m={cdim}, {tc:.1f} s"}
    print(json.dumps(out))
    return 0


def slab_c5(args):
    """The stiff workload of the learned-proposal form: the 201,600-tet N=3
    slab box graded 20:1 toward an eps_r 10, sigma 50 slab, pulse field."""
    from rapidfem_amd import workloads as W

    nx, ny, nz = args.krylov_cells
    return W.slab_box(nx, ny, nz, args.order)


def krylov_cost_model(g, ctx, y0, cfl, dims=(10, 20, 30, 40, 50, 60)):
    """Per-step wall time of the adaptive Krylov step at fixed dimensions on
    this operator, fitted as c0 + c1 m + c2 m (m + 1) / 2 (ms; per step, per
    operator apply, per basis column swept) by non-negative least squares
    over every step of a 40-CFL-step run at each dimension (first steps
    left out)."""
    import numpy as np
    from scipy.optimize import nnls

    from rapidfem_amd.proposer import ExpController

    rows, ms = [], []
    for m in dims:
        ctl = ExpController(tol=1e-6, m=m, h0=cfl, growth_limit=1.5, h_scale=cfl)
        _, steps, _ = g.expmv_adaptive(ctx, y0, 40 * cfl, ctl, log_cap=256)
        for s in steps[1:]:
            rows.append([1.0, s["m"], 0.5 * s["m"] * (s["m"] + 1)])
            ms.append(s["ms"])
    coef, _ = nnls(np.array(rows), np.array(ms))
    return [float(c) for c in coef], len(ms)


def proposer_bench(args) -> int:
    """The learned-proposal form: exp(t_end A) v on the stiff slab workload
    with the plain error-controlled controller (m = 40, the session's
    krylov_dim) and with the trained (h, m) proposer, at the same tolerance;
    both final states against a fine-step FP64 propagation (the plain
    controller at a thousand times tighter tolerance). value = plain wall
    time / learned wall time."""
    import numpy as np

    import rapidfem_amd as R
    from rapidfem_amd.proposer import WEIGHTS, ExpController, Proposer

    op, y0 = slab_c5(args)
    ctx = R.GpuContext()
    g = R.GpuOperator(ctx, op, "fused_f64")
    variant, _ = g.autotune()
    cfl = R.TdSession(op, 1.0, "fused_f64", ctx).cfl_dt()
    t_end = args.proposer_cfl * cfl
    tol = args.proposer_tol
    log(f"proposer: {op.n_elem} tets, n={op.n_dof}, cfl {cfl:.4e}, t_end {t_end:.4e} ({args.proposer_cfl} CFL steps), "
        f"tol {tol:g}")
    cost, nfit = krylov_cost_model(g, ctx, y0, cfl)
    log(f"  cost model (ms): c0 {cost[0]:.3f} + c1 {cost[1]:.4f} m + c2 {cost[2]:.5f} m(m+1)/2 from {nfit} steps")
    if args.cost_only:
        print(json.dumps({"metric": "adaptive Krylov step cost model", "cost_model_ms": cost, "steps_fitted": nfit,
                          "config": {"workload": f"slab box {op.n_elem} tets N={args.order}"}}))
        return 0
    fine_ctl = ExpController(tol=tol * 1e-3, m=40, h_scale=cfl)
    yf, sf, _ = g.expmv_adaptive(ctx, y0, t_end, fine_ctl)
    log(f"  fine run: {len(sf)} steps (m 40, tol {tol * 1e-3:g}), {sf['ms'].sum():.0f} ms")
    sweep = {}
    if args.m_sweep:
        # the plain controller at each fixed dimension: what choosing m can gain on this workload
        for m in (10, 16, 20, 24, 28, 32, 40, 48, 60):
            ctl = ExpController(tol=tol, m=m, h_scale=cfl)
            y, st, _ = g.expmv_adaptive(ctx, y0, t_end, ctl, log_cap=4096)
            sweep[m] = {"device_ms": round(g.last_call_ms, 1), "steps": int(len(st)),
                        "rel_l2_vs_fine": float(np.linalg.norm(y - yf) / np.linalg.norm(yf))}
            log(f"  plain m={m}: {len(st)} steps, device {g.last_call_ms:.0f} ms")
    res = {}
    prop = Proposer(ctx)
    prop.load(args.proposer_weights or WEIGHTS)
    for name, p in (("plain", None), ("learned", prop)):
        ctl = ExpController(tol=tol, m=40, m_min=10, m_max=60, h_scale=cfl)
        g.expmv_adaptive(ctx, y0, t_end, ctl, proposer=p, log_cap=4096)  # warm-up
        t0 = time.perf_counter()
        y, st, _ = g.expmv_adaptive(ctx, y0, t_end, ctl, proposer=p, log_cap=4096)
        wall = (time.perf_counter() - t0) * 1e3
        err = float(np.linalg.norm(y - yf) / np.linalg.norm(yf))
        res[name] = {"device_ms": g.last_call_ms, "call_ms": wall, "steps": int(len(st)),
                     "applies": int(st["m"].sum()), "first_proposals_rejected": int((st["n_judged"] > 1).sum()),
                     "m_mean": float(st["m"].mean()), "m_first": [int(m) for m in st["m"][:40]], "rel_l2_vs_fine": err,
                     "h_over_cfl": [round(float(h / cfl), 2) for h in st["h"][:40]]}
        log(f"  {name}: {len(st)} steps, {int(st['m'].sum())} applies, device {g.last_call_ms:.0f} ms, "
            f"rel L2 vs fine {err:.2e}")
    # the same two runs with the finished basis columns stored as f32 + bf16
    g.set_krylov_basis("compressed")
    comp = {}
    for name, p in (("plain", None), ("learned", prop)):
        ctl = ExpController(tol=tol, m=40, m_min=10, m_max=60, h_scale=cfl)
        g.expmv_adaptive(ctx, y0, t_end, ctl, proposer=p, log_cap=4096)  # warm-up
        y, st, _ = g.expmv_adaptive(ctx, y0, t_end, ctl, proposer=p, log_cap=4096)
        err = float(np.linalg.norm(y - yf) / np.linalg.norm(yf))
        comp[name] = {"device_ms": g.last_call_ms, "steps": int(len(st)), "applies": int(st["m"].sum()),
                      "m_mean": float(st["m"].mean()), "rel_l2_vs_fine": err,
                      "speedup_vs_f64_basis": round(res[name]["device_ms"] / g.last_call_ms, 3)}
        log(f"  {name}, compressed basis: {len(st)} steps, {int(st['m'].sum())} applies, device "
            f"{g.last_call_ms:.0f} ms ({comp[name]['speedup_vs_f64_basis']:.2f}x the FP64 basis), "
            f"rel L2 vs fine {err:.2e}")
    g.set_krylov_basis("f64")
    gain = res["plain"]["device_ms"] / res["learned"]["device_ms"]
    out = {"metric": "Krylov-exp wall time, learned (h, m) proposer vs plain controller at equal tolerance",
           "value": round(gain, 3), "unit": "x (plain ms / learned ms)", "n_gpus": 1, "higher_is_better": True,
           "dtype": "f64 (Arnoldi, CGS2, Pade-13), f32 network", "data": "synthetic", "tol": tol,
           "plain": res["plain"], "learned": res["learned"], "fine": {"steps": int(len(sf)), "tol": tol * 1e-3},
           "compressed_basis": comp, "cost_model_ms": cost, "plain_fixed_m": sweep,
           # the plain controller at the dimension that was best in hindsight on this workload
           "vs_best_fixed_m": (None if not sweep else
                               {"m": min(sweep, key=lambda m: sweep[m]["device_ms"]),
                                "gain": round(min(v["device_ms"] for v in sweep.values())
                                              / res["learned"]["device_ms"], 3)}),
           "config": {"workload": f"slab box {args.krylov_cells[0]}x{args.krylov_cells[1]}x{args.krylov_cells[2]} x 6 "
                                  f"tets graded 20:1 to an eps_r 10, sigma 50 slab, N={args.order}, pulse field, "
                                  f"t_end {args.proposer_cfl} CFL steps",
                      "n_elem": op.n_elem, "n": op.n_dof, "variant": variant}}
    print(json.dumps(out))
    return 0


def drift_bench(args) -> int:
    """Long-run precision of the register and arithmetic forms: --drift-steps
    LSERK4 steps at the CFL step of the stepper workload (C2, 5000 steps by
    default) from the same pulse, each form's final state against the FP64
    fused form's, with each form's ms per step. The reference's f32 GPU
    structure is plain_f32."""
    import numpy as np

    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    nx, ny, nz = args.cells
    op, y0 = W.pulse_box(nx, ny, nz, args.order)
    ctx = R.GpuContext()
    sess = R.TdSession(op, 1.0, "fused_f64", ctx)
    dt = sess.cfl_dt()
    steps = args.drift_steps
    ref = sess.gop
    ref.autotune()
    ref.state_load(y0)
    t_ref, _ = ref.state_lserk_timed(dt, steps)
    want = ref.state_store()
    log(f"drift: {op.n_elem} tets N={args.order}, {steps} steps at dt {dt:.4e}; fused_f64 {t_ref / steps:.3f} ms/step")
    forms = {}
    for form in ("fused_f64c", "plain_f64", "plain_f32", "fused_bf16x4", "fused_bf16x3", "fused_bf16x2"):
        g = R.GpuOperator(ctx, op, form)
        if form.startswith("fused"):
            g.autotune()
        g.state_load(y0)
        tot, _ = g.state_lserk_timed(dt, steps)
        y = g.state_store()
        err = float(np.linalg.norm(y - want) / np.linalg.norm(want))
        forms[form] = {"rel_l2_vs_fp64": err, "ms_per_step": tot / steps}
        log(f"  {form}: rel L2 vs FP64 after {steps} steps {err:.3e}, {tot / steps:.3f} ms/step")
        del g
    out = {"metric": "relative L2 error against the FP64 fused form after a long LSERK4 run, per form",
           "value": forms["fused_f64c"]["rel_l2_vs_fp64"], "unit": "rel L2 (chosen form)", "n_gpus": 1,
           "higher_is_better": False, "steps": steps, "dt": dt, "forms": forms, "fused_f64_ms_per_step": t_ref / steps,
           "data": "synthetic", "config": {"workload": f"pulse_box {nx}x{ny}x{nz} cells x 6 tets, N={args.order}",
                                           "n_elem": op.n_elem, "n_dof": op.n_dof}}
    print(json.dumps(out))
    return 0


def kcl_bench(args) -> int:
    """KCL RK4(3)5[2R+]C with the PI controller on the stepper workload: two
    frames of 8 CFL steps each, the device part of the call timed (frame
    downloads excluded). Unit: DOF-stage over all substeps, rejected ones
    included (5 stages each)."""
    import numpy as np

    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    nx, ny, nz = args.cells
    op, y0 = W.pulse_box(nx, ny, nz, args.order)
    n = op.n_dof
    ctx = R.GpuContext()
    sess = R.TdSession(op, 1.0, "fused_f64", ctx)
    dt = sess.cfl_dt()
    frame = 8.0 * dt
    g = sess.gop
    variant, tune_ms = g.autotune()
    log(f"kcl: {op.n_elem} tets, N={args.order}, n={n}, cfl dt {dt:.4e}, frame {frame:.4e}, "
        f"stage kernel variant {variant}")
    def reps(gop, label):
        runs, last = [], None
        for r in range(3):
            t0 = time.perf_counter()
            traj, st = gop.transient_kcl(ctx, y0, frame, 2, precision="f64")
            wall = (time.perf_counter() - t0) * 1e3
            ms = gop.last_call_ms
            sub = st.accepted + st.rejected
            log(f"  {label} rep {r}: {st.accepted} accepted + {st.rejected} rejected substeps, device part "
                f"{ms:.2f} ms, call {wall:.1f} ms")
            if r > 0:
                runs.append((ms, sub, st.accepted, st.rejected, wall))
            last = traj[-1]
        return runs, last

    runs, y_f64 = reps(g, "f64 registers")
    # the same registers with the update as a pass of its own after each apply
    g.set_kcl_registers("f64_separate")
    runs_sep, y_sep = reps(g, "f64 registers, separate update")
    g.set_kcl_registers("f64")
    ms_sep = float(np.median([x[0] for x in runs_sep]))
    f64_separate = {"ms_per_call": ms_sep, "ms_per_substep": ms_sep / max(runs_sep[-1][1], 1),
                    "value": n * 5 * runs_sep[-1][1] / (ms_sep * 1e-3),
                    "rel_l2_vs_fused_update": float(np.linalg.norm(y_sep - y_f64) / np.linalg.norm(y_f64))}
    # the compressed-register form with compressed KCL registers
    gc = R.GpuOperator(ctx, op, "fused_f64c")
    vc, _ = gc.autotune()
    def summary(label, runs_c, y_c):
        ms_c = float(np.median([x[0] for x in runs_c]))
        sub_c = runs_c[-1][1]
        r = {"form": "fused_f64c", "variant": vc, "ms_per_call": ms_c, "substeps": sub_c,
             "accepted": runs_c[-1][2], "rejected": runs_c[-1][3], "ms_per_substep": ms_c / max(sub_c, 1),
             "value": n * 5 * sub_c / (ms_c * 1e-3),
             "rel_l2_vs_f64_registers": float(np.linalg.norm(y_c - y_f64) / np.linalg.norm(y_f64))}
        log(f"  {label}: {r['ms_per_substep']:.3f} ms per substep, "
            f"rel L2 vs FP64 registers {r['rel_l2_vs_f64_registers']:.2e}")
        return r

    # the update as a pass of its own after each apply
    gc.set_kcl_registers("compressed_separate")
    runs_s, y_s = reps(gc, "compressed registers, separate update")
    separate = summary("compressed registers, separate update", runs_s, y_s)
    # the update in the stage kernel's epilogue
    gc.set_kcl_registers("compressed")
    runs_c, y_c = reps(gc, "compressed registers, fused update")
    compressed = summary("compressed registers, fused update", runs_c, y_c)
    compressed["update"] = "in the stage kernel's epilogue (k_fused11, no k stored)"
    compressed["separate_update"] = separate
    compressed["rel_l2_vs_separate_update"] = float(np.linalg.norm(y_c - y_s) / np.linalg.norm(y_s))
    compressed["speedup_vs_separate_update"] = separate["ms_per_substep"] / compressed["ms_per_substep"]
    log(f"  fused update {compressed['speedup_vs_separate_update']:.2f}x the separate pass, rel L2 between them "
        f"{compressed['rel_l2_vs_separate_update']:.2e}")
    del gc
    ms = float(np.median([x[0] for x in runs]))
    wall = float(np.median([x[4] for x in runs]))
    sub, acc, rej = runs[-1][1], runs[-1][2], runs[-1][3]
    value = n * 5 * sub / (ms * 1e-3)
    out = {"metric": "KCL adaptive stepper DOF-stage/s (RK4(3)5[2R+]C + PI controller)", "value": value,
           "unit": "DOF-stage/s", "n_gpus": 1, "reps": len(runs), "higher_is_better": True, "dtype": "f64",
           "data": "synthetic", "ms_per_call": ms, "substeps": sub, "accepted": acc, "rejected": rej,
           "ms_per_substep": ms / max(sub, 1),
           "update": "in the stage kernel's epilogue where the operator allows it (k_fused11)",
           "separate_update": f64_separate, "compressed_registers": compressed,
           "host_inclusive": {"ms_per_call": wall, "value": n * 5 * sub / (wall * 1e-3),
                              "note": "the whole call: state upload, layout conversions and the 3 frames "
                                      "downloaded as f64 (311 MB each)"},
           "config": {"workload": f"pulse box {nx}x{ny}x{nz} x 6 tets, N={args.order}, 2 frames of 8 CFL steps",
                      "n_elem": op.n_elem, "n": n, "form": "fused_f64", "variant": variant}}
    if not args.no_cpu:
        from oracle import oracle as O

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.set_threads(threads)
        oop = O.Operator(op.mesh.nodes, op.mesh.tets, args.order, op.flux_alpha, eps=1.0 / op.inv_eps)
        t0 = time.perf_counter()
        _, oacc, orej, _, _ = oop.adaptive(y0, frame, 1)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * 5 * (oacc + orej) / tc, "unit": "DOF-stage/s", "cores": threads,
                               "kind": "port",
                               "sample": f"oracle adaptive KCL (explicit_adaptive.rs:84-152 + session.rs:781-827 "
                                         f"restated, OpenMP) on the same mesh, 1 frame ({oacc} + {orej} substeps), "
                                         f"{tc:.1f} s"}
    print(json.dumps(out))
    return 0


def batch_bench(args) -> int:
    """The batching form on one GPU: 8 jittered boxes with seeded sizes
    (20k-300k tets, log-uniform) and a seeded dielectric block, orders 2, 3, 4
    in turn, four port walls each with a TE10-shaped incident profile; each
    mesh carries 4 right-hand sides (one operator each, 32 operators), the
    r-th driven by port r's source (rhs.rs:1157-1199) under a Gaussian-
    modulated pulse with a centre frequency seeded in [3, 8]
    (session.rs:242-261), f64c registers, LSERK4 at each mesh's CFL step. One
    autotune per mesh, shared by its right-hand sides. Timed: --batch-steps
    steps of all operators from one graph (rfhip_batch_lserk_ports), and a
    shorter run of each operator alone for comparison. Unit: DOF-stage over
    all operators."""
    import numpy as np

    import rapidfem_amd as R
    from rapidfem_amd import GaussianPulse
    from rapidfem_amd import workloads as W

    ctx = R.GpuContext()
    ops, dts, waves, dofs, meshes = [], [], [], 0, []
    steps = args.batch_steps
    for b, (nx, ny, nz, order) in enumerate(W.port_batch_specs()):
        op, ports = W.port_box(nx, ny, nz, order, 11 + b)
        cfl = R.TdSession(op, 1.0, "fused_f64c", ctx).cfl_dt()
        pulses = W.port_waveforms(100 + b, 4)
        variant = None
        for r in range(4):
            g = R.GpuOperator(ctx, op, "fused_f64c")
            for q, (el, lf, prof, w) in enumerate(ports):
                g.set_port(q, el, lf, prof, w)
            g.port_source(r, download=False)
            if variant is None:
                variant, _ = g.autotune()
            g.set_variant(variant)
            g.state_load(np.zeros(op.n_dof))
            pulse = GaussianPulse(*pulses[r])
            ops.append(g)
            dts.append(cfl)
            waves.append(np.array([pulse(k * cfl) for k in range(steps)]))
        dofs += 4 * op.n_dof
        meshes.append(f"{nx}x{ny}x{nz} N={order} ({op.n_elem} tets, v{variant}, f0 "
                      f"{', '.join(f'{p[2]:.2f}' for p in pulses)})")
        log(f"mesh {b}: {meshes[-1]}")
    R.batch_lserk_ports(ops, dts, 2, [w[:2] for w in waves])  # warm-up (graph capture, kernels)
    for g in ops:
        g.state_load(np.zeros(g.n_dof()))
    t0 = time.perf_counter()
    R.batch_lserk_ports(ops, dts, steps, waves)
    t_batch = time.perf_counter() - t0
    norms = [float(np.linalg.norm(g.state_store())) for g in ops[::4]]
    seq_steps = min(steps, 100)
    t0 = time.perf_counter()
    for g, dt, w in zip(ops, dts, waves):
        R.batch_lserk_ports([g], [dt], seq_steps, [w[:seq_steps]])
    t_seq = time.perf_counter() - t0
    log(f"batched {t_batch * 1e3:.1f} ms for {steps} steps of {len(ops)} operators; one after another "
        f"{t_seq * 1e3:.1f} ms for {seq_steps} steps")
    out = {"metric": "batching form DOF-stage/s (8 meshes x 4 port excitations, one graph)",
           "value": dofs * 5 * steps / t_batch, "unit": "DOF-stage/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "f64 arithmetic, y f32+bf16 / p f32 registers", "data": "synthetic", "steps": steps,
           "ms_batched": t_batch * 1e3, "ms_per_step": t_batch * 1e3 / steps,
           "sequential": {"steps": seq_steps, "ms": t_seq * 1e3, "value": dofs * 5 * seq_steps / t_seq},
           "final_norms_first_rhs": norms,
           "config": {"workload": "8 seeded jittered port boxes (20k-300k tets), orders 2/3/4, 4 TE10 port "
                                  "excitations each (Gaussian pulses, f0 in [3, 8]), f64c registers, LSERK4 at the "
                                  "CFL step", "meshes": meshes, "operators": len(ops), "total_dofs": dofs}}
    print(json.dumps(out))
    return 0


def record_bench(args) -> int:
    """20 frames of one LSERK4 step each at dt = 0.002 (below the CFL step
    2.32e-3) on the stepper workload, three ways: every frame's full state
    downloaded (Record::States through the reference GPU path's per-frame
    download, session.rs:914-921), 64 probe DOFs recorded on the device and
    downloaded once (Record::Dofs, rfhip_op_transient_probes_f64), and the
    final state only. Wall time of each call; value = recorded frames/s of
    the probe path."""
    import numpy as np

    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    nx, ny, nz = args.cells
    op, y0 = W.pulse_box(nx, ny, nz, args.order)
    n = op.n_dof
    ctx = R.GpuContext()
    g = R.GpuOperator(ctx, op, args.form)
    variant, _ = g.autotune()
    frames, dt = 20, 0.002
    probes = np.linspace(0, n - 1, 64).astype(np.int64)
    g.transient(ctx, y0, dt, 2, 1, record=False, precision="f64")  # warm-up
    res = {}
    for name in ("full", "probes", "final"):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            if name == "full":
                out = g.transient(ctx, y0, dt, frames, 1, record=True, precision="f64")
            elif name == "probes":
                out, _ = g.transient_probes(ctx, y0, dt, frames, probes)
            else:
                out = g.transient(ctx, y0, dt, frames, 1, record=False, precision="f64")
            ts.append(time.perf_counter() - t0)
            del out
        res[name] = float(np.median(ts))
        log(f"record {name}: {res[name] * 1e3:.1f} ms for {frames} frames")
    value = frames / res["probes"]
    out = {"metric": "recorded frames/s, probe DOFs on the device vs full-state download per frame",
           "value": value, "unit": "frames/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
           "data": "synthetic", "frames": frames,
           "ms_per_frame": {k: v * 1e3 / frames for k, v in res.items()},
           "full_state_bytes_per_frame": 8 * n,
           "speedup_probes_vs_full": res["full"] / res["probes"],
           "config": {"workload": f"pulse box {nx}x{ny}x{nz} x 6 tets, N={args.order}, dt 0.002, 1 step per frame",
                      "n": n, "form": args.form, "variant": variant, "probes": int(probes.shape[0])}}
    print(json.dumps(out))
    return 0


def sparse_export(args) -> int:
    """Device assemble_sparse on the jittered box (PEC walls, upwind flux):
    value = stored nonzeros per second of the two row passes (count + fill,
    HIP events), the CPU leg the oracle's unit-column probe on an 8x8x8 box
    of the same order."""
    import numpy as np

    import rapidfem_amd as R

    nx, ny, nz = args.sparse_cells
    mesh = R.structured_box_jittered(nx, ny, nz, 1.0, 1.0, 1.0, 0.1, 7)
    op = R.MaxwellOperator(mesh, args.order, 1.0)
    ctx = R.GpuContext()
    log(f"sparse export: {op.n_elem} tets, N={args.order}, n={op.n_dof}")
    runs = []
    for r in range(args.sparse_reps + 1):
        t0 = time.perf_counter()
        c = R.assemble_sparse(ctx, op, download=False)
        wall = time.perf_counter() - t0
        log(f"  rep {r}: nnz {c.nnz()} count {c.count_ms:.2f} ms fill {c.fill_ms:.2f} ms call {wall * 1e3:.1f} ms")
        if r > 0:
            runs.append((c.count_ms + c.fill_ms, c.count_ms, c.fill_ms, wall))
    ms, cms, fms, wall = (float(np.median([x[k] for x in runs])) for k in range(4))
    nnz, n = c.nnz(), c.n
    del c
    t0 = time.perf_counter()
    ch = R.assemble_sparse(ctx, op, download=True)
    wall_dl = time.perf_counter() - t0
    host_bytes = ch.row_ptr.nbytes + ch.col_idx.nbytes + ch.values.nbytes
    log(f"  with the CSR downloaded: {wall_dl * 1e3:.1f} ms, {host_bytes / 1e9:.2f} GB to the host")
    del ch
    fill_bytes = 16.0 * nnz + 8.0 * n  # (col_idx, value) per nonzero + row_ptr read
    out = {"metric": "sparse state-space export (assemble_sparse) stored nonzeros/s", "value": nnz / (ms * 1e-3),
           "unit": "nnz/s", "n_gpus": 1, "reps": args.sparse_reps, "higher_is_better": True, "dtype": "f64",
           "data": "synthetic", "ms_two_passes": ms, "ms_count": cms, "ms_fill": fms, "ms_call": wall * 1e3,
           "host_inclusive": {"ms_per_call": wall_dl * 1e3, "value": nnz / wall_dl, "host_bytes": host_bytes,
                              "note": "one call with the CSR (row_ptr, col_idx, values) downloaded to host "
                                      "arrays, as MaxwellOperator::assemble_sparse returns it"},
           "config": {"workload": f"jittered box {nx}x{ny}x{nz} x 6 tets, N={args.order}, PEC, upwind",
                      "n_elem": op.n_elem, "n": n, "nnz": nnz},
           "roofline": {"bound": "hbm", "achieved": fill_bytes / (fms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": fill_bytes / (fms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "k_sparse_rows fill pass, 16 B per nonzero + 8 B per row"}}
    if not args.no_cpu:
        from oracle import oracle as O

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.set_threads(threads)
        sm = R.structured_box_jittered(8, 8, 8, 1.0, 1.0, 1.0, 0.1, 7)
        oop = O.Operator(sm.nodes, sm.tets, args.order, 1.0)
        t0 = time.perf_counter()
        _, _, va = oop.assemble_sparse()
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": len(va) / t, "unit": "nnz/s", "cores": threads, "kind": "port",
                               "sample": f"oracle unit-column probe (rhs.rs:1621-1730 restated, OpenMP) on the "
                                         f"8x8x8 box ({oop.K} tets, N={args.order}, {len(va)} nonzeros), "
                                         f"{t:.1f} s, two passes (count, fill)"}
    print(json.dumps(out))
    return 0


def traffic_probe(args) -> int:
    import torch

    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    n_copy = 1 << 27  # 2^27 f64 = 1 GiB read + 1 GiB written per launch
    a = torch.ones(n_copy, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    del a, b
    if args.probe_c3:
        op, y0 = W.modes_box(55, 4)  # the 998,250-tet N=4 cavity of the C3 line
        order = 4
    else:
        nx, ny, nz = args.cells
        op, y0 = W.pulse_box(nx, ny, nz, args.order)
        order = args.order
    ctx = R.GpuContext()
    g = R.GpuOperator(ctx, op, args.form)
    if args.variant >= 0:
        g.set_variant(args.variant)
    elif args.autotune:
        g.autotune()
    g.state_load(y0)
    g.state_lserk(0.002, 10)
    print(json.dumps({"probe": "traffic", "copy_bytes_read": 8 * n_copy, "copy_bytes_written": 8 * n_copy,
                      "form": args.form, "variant": g.variant, "n_dof": op.n_dof, "stage_launches": 50,
                      "workload": f"{op.n_elem} tets N={order}",
                      "algorithmic_bytes_per_launch": W.bytes_per_dof_stage(order, args.form) * op.n_dof}))
    return 0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(op, y0, dt, steps):
    """Oracle LSERK4 on the full mesh: `steps` steps on all host threads,
    then one step on one thread. Returns the baseline record and the
    oracle's state after the `steps` steps (the parity anchor of the
    benchmark's GPU forms)."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    t0 = time.perf_counter()
    oop = O.Operator(op.mesh.nodes, op.mesh.tets, op.order, op.flux_alpha,
                     eps=1.0 / op.inv_eps)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    y_or = oop.lserk(y0, dt, steps)
    t = time.perf_counter() - t0
    rate = oop.n_dof * 5 * steps / t
    O.set_threads(1)
    t0 = time.perf_counter()
    oop.lserk(y0, dt, 1)
    t1 = time.perf_counter() - t0
    O.set_threads(threads)
    rate1 = oop.n_dof * 5 / t1
    rec = {"value": rate, "unit": "DOF-stage/s", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "sample": f"oracle (C restatement of the reference CPU path, OpenMP) LSERK4, {steps} steps on the "
                     f"full {op.n_elem}-tet mesh, {t:.1f} s (operator build {t_build:.1f} s excluded)",
           "one_core": {"value": rate1, "cores": 1,
                        "sample": f"the same oracle, 1 LSERK4 step on the full mesh on one thread, {t1:.1f} s"}}
    return rec, y_or


def roofline_record(order, form, n, stage_ms, traffic=None):
    """Roofline of one fused stage launch over n DOFs: algorithmic bytes and
    flops per DOF-stage over the launch's mean duration. The binding roof is
    the one whose ridge the kernel's intensity passes: FP64 compute (vector
    and matrix FP64 share one 78.6 TFLOP/s peak on MI355X) when flops/bytes
    exceeds 78.6e12 / 8e12 = 9.8 flop/B, else HBM."""
    from rapidfem_amd import workloads as W

    bpd = W.bytes_per_dof_stage(order, form)
    fpd = W.flops_per_dof_stage(order)
    gbs = bpd * n / (stage_ms * 1e-3) / 1e9
    tfl = fpd * n / (stage_ms * 1e-3) / 1e12
    intensity = fpd / bpd
    ridge = FP64_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9)
    hbm = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
           "bytes_per_dof_stage": round(bpd, 3)}
    fp64 = {"achieved": round(tfl, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tfl / FP64_PEAK_TFLOPS, 4), "flops_per_dof_stage": round(fpd, 1)}
    top = dict(fp64 if intensity > ridge else hbm)
    top.update({"bound": "mfma" if intensity > ridge else "hbm", "traffic": traffic,
                "intensity_flop_per_byte": round(intensity, 2), "ridge_flop_per_byte": round(ridge, 2),
                "algorithmic_bytes_per_launch": bpd * n, "algorithmic_flops_per_launch": fpd * n,
                "compute_pipe": "FP64 (VALU and MFMA share the 78.6 TFLOP/s peak)", "hbm": hbm, "fp64": fp64})
    return top


def c3_line(ctx, args):
    """The largest single-GPU stepper workload beside the headline: the
    998,250-tet jittered N=4 box with 16 cavity modes (209.6M DOFs). The
    chosen form and the plain FP64 form step the same state at a fixed
    stable dt (throughput does not depend on dt; the CFL-step precision run
    of this workload is a GPU test); rel L2 of the chosen form against the
    FP64 fused form after the timed steps."""
    import numpy as np

    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    t0 = time.perf_counter()
    op, y0 = W.modes_box(55, 4)
    n = op.n_dof
    log(f"c3: {op.n_elem} tets N=4, n={n}, built in {time.perf_counter() - t0:.1f}s")
    dt, steps = 2.0e-4, args.c3_steps
    ch = R.GpuOperator(ctx, op, args.form)
    variant, _ = ch.autotune()
    ch.state_load(y0)
    ch.state_lserk(dt, 2)
    tot, stage = ch.state_lserk_timed(dt, steps)
    y_ch = ch.state_store()
    del ch
    ref = R.GpuOperator(ctx, op, "fused_f64")
    rvar, _ = ref.autotune()
    ref.state_load(y0)
    ref.state_lserk(dt, 2)
    rtot, _ = ref.state_lserk_timed(dt, steps)
    err = float(np.linalg.norm(ref.state_store() - y_ch) / np.linalg.norm(y_ch))
    del ref
    pl = R.GpuOperator(ctx, op, "plain_f64")
    pl.state_load(y0)
    pl.state_lserk(dt, 1)
    ptot, papply = pl.state_lserk_timed(dt, max(4, steps // 2))
    pms = ptot / max(4, steps // 2)
    del pl
    value = n * 5 * steps / (tot * 1e-3)
    log(f"c3: {args.form} v{variant} {tot / steps:.3f} ms/step ({value:.4e} DOF-stage/s), stage {stage:.4f} ms; "
        f"fused_f64 v{rvar} {rtot / steps:.3f} ms/step; plain_f64 {pms:.3f} ms/step; rel L2 {err:.3e}")
    return {"workload": "modes_box 55^3 cells x 6 tets (jittered unit box), N=4, 16 PEC cavity modes, upwind, "
                        f"LSERK4 at dt {dt:g}",
            "n_elem": op.n_elem, "n_dof": n, "order": 4, "form": args.form, "variant": variant, "steps": steps,
            "value": value, "ms_per_step": tot / steps,
            "fused_f64": {"variant": rvar, "ms_per_step": rtot / steps},
            "plain": {"form": "plain_f64", "ms_per_step": pms, "value": n * 5 / (pms * 1e-3),
                      "apply_kernel_ms": papply},
            "speedup_vs_plain": pms / (tot / steps), "rel_l2_vs_f64": err, "rel_l2_steps": steps + 2,
            "roofline": roofline_record(4, args.form, n, stage, traffic_entry(TRAFFIC_FILE_C3, args.form, n, variant))}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(the element-slab sharding across GPUs is designed, not built)"}))
        return 2
    if args.traffic_probe:
        return traffic_probe(args)
    if args.sparse_export:
        return sparse_export(args)
    if args.proposer:
        return proposer_bench(args)
    if args.krylov:
        return krylov_bench(args)
    if args.record:
        return record_bench(args)
    if args.batch:
        return batch_bench(args)
    if args.kcl:
        return kcl_bench(args)
    if args.drift:
        return drift_bench(args)
    import numpy as np

    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    nx, ny, nz = args.cells
    t0 = time.perf_counter()
    op, y0 = W.pulse_box(nx, ny, nz, args.order)
    log(f"operator build {time.perf_counter() - t0:.1f}s: K={op.n_elem} N={args.order} n_dof={op.n_dof}")
    ctx = R.GpuContext()
    log(f"device {ctx.device_name}")
    sess = R.TdSession(op, 1.0, "fused_f64", ctx)
    t0 = time.perf_counter()
    dt = sess.cfl_dt()
    log(f"cfl_dt {dt:.6e} ({time.perf_counter() - t0:.1f}s, rho={sess.rho:.4e})")
    ref = sess.gop
    ref_variant, _ = ref.autotune()
    log(f"fused_f64 autotune: variant {ref_variant}")
    n = op.n_dof
    K, Wm = args.steps, args.warmup

    # Reference FP64 trajectory for the error check: warmup + K steps.
    ref.state_load(y0)
    ref.state_lserk(dt, Wm + K)
    y_ref = ref.state_store()

    chosen = R.GpuOperator(ctx, op, args.form)
    variant = 0
    if args.variant >= 0:
        chosen.set_variant(args.variant)
        variant = chosen.variant
        log(f"variant {variant} pinned")
    elif args.autotune:
        variant, tune_ms = chosen.autotune()
        log(f"autotune: variant {variant} of {chosen.n_variants()} ({tune_ms:.4f} ms/stage)")
    chosen.state_load(y0)
    chosen.state_lserk(dt, Wm)
    t0 = time.perf_counter()
    total_ms, stage_ms = chosen.state_lserk_timed(dt, K)
    wall = time.perf_counter() - t0
    y_ch = chosen.state_store()
    err = float(np.linalg.norm(y_ch - y_ref) / np.linalg.norm(y_ref))
    ms_per_step = total_ms / K
    value = n * 5 * K / (total_ms * 1e-3)
    log(f"{args.form}: {value:.4e} DOF-stage/s, {ms_per_step:.3f} ms/step (wall {wall:.2f}s), "
        f"stage kernel {stage_ms:.4f} ms, rel L2 vs fused_f64 after {Wm + K} steps {err:.3e}")

    # Plain FP64 form (reference structure), same state and dt.
    plain = R.GpuOperator(ctx, op, "plain_f64")
    plain.state_load(y0)
    plain.state_lserk(dt, 2)
    ptot, papply = plain.state_lserk_timed(dt, args.plain_steps)
    pvalue = n * 5 * args.plain_steps / (ptot * 1e-3)
    log(f"plain_f64: {pvalue:.4e} DOF-stage/s, {ptot / args.plain_steps:.3f} ms/step, apply kernel {papply:.4f} ms")
    del plain

    # Fused FP64 form for reference.
    ref.state_load(y0)
    ftot, fstage = ref.state_lserk_timed(dt, max(10, K // 4))
    fvalue = n * 5 * max(10, K // 4) / (ftot * 1e-3)
    log(f"fused_f64: {fvalue:.4e} DOF-stage/s, stage kernel {fstage:.4f} ms")

    # Roofline of the dominant kernel (the fused stage kernel).
    traffic = traffic_entry(TRAFFIC_FILE, args.form, n, variant)
    roofline = roofline_record(args.order, args.form, n, stage_ms, traffic)
    cpu = None
    vs_oracle = None
    if not args.no_cpu:
        try:
            cpu, y_or = cpu_baseline(op, y0, dt, args.cpu_steps)
            log(f"cpu baseline {cpu['value']:.4e} DOF-stage/s on {cpu['cores']} threads, "
                f"{cpu['one_core']['value']:.4e} on one ({cpu['cpu_model']})")
            vs_oracle = {"steps": args.cpu_steps}
            for name, g in (("chosen", chosen), ("fused_f64", ref)):
                g.state_load(y0)
                g.state_lserk(dt, args.cpu_steps)
                vs_oracle[name] = float(np.linalg.norm(g.state_store() - y_or) / np.linalg.norm(y_or))
            log(f"rel L2 vs the oracle after {args.cpu_steps} steps: {args.form} {vs_oracle['chosen']:.3e}, "
                f"fused_f64 {vs_oracle['fused_f64']:.3e}")
            del y_or
        except Exception as e:  # the oracle is optional on a box without gcc
            log(f"cpu baseline failed: {e}")
    c3 = None
    if args.c3_steps > 0:
        del chosen, ref, sess
        c3 = c3_line(ctx, args)
    out = {
        "metric": METRIC, "value": value, "unit": "DOF-stage/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE_LABEL.get(args.form, "f64"), "data": "synthetic",
        "config": {"workload": f"pulse_box {nx}x{ny}x{nz} cells x 6 tets (2x1x1 box), N={args.order}, eps_r=4 "
                               f"block, Gaussian E_x, PEC, upwind, LSERK4 at cfl_dt",
                   "n_elem": op.n_elem, "n_dof": n, "order": args.order, "dt": dt, "form": args.form,
                   "variant": variant,
                   "registers": "f64" if args.form == "fused_f64" else "y f32+bf16, p f32"},
        "plain": {"form": "plain_f64", "value": pvalue, "ms_per_step": ptot / args.plain_steps},
        "fused_f64": {"value": fvalue, "variant": ref_variant},
        "speedup_vs_plain": value / pvalue,
        "rel_l2_vs_f64": err, "rel_l2_steps": Wm + K, "rel_l2_vs_oracle": vs_oracle,
        "roofline": roofline, "cpu_baseline": cpu, "c3": c3,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
