"""Trains the adaptive Krylov exponential's (h, m) proposer on the plain
controller's own accept/reject logs, on the GPU, with the library's HIP
training kernels (no reference counterpart; the reference's exponential
route takes a fixed step and dimension, crates/rapidfem-td/src/
session.rs:700-726).

Data: the plain controller (GpuOperator.expmv_adaptive without a proposer)
runs at Krylov dimension m_log = 60 on seeded stiff workloads: slab boxes
graded 10:1 to 30:1 toward a lossy slab (eps_r 4..16, sigma 10..100, N=3)
in three size classes (about 1k-5k, 10k-40k and 130k-250k tets, the last
the size of the target workload), several initial pulses each, tolerances
1e-9..1e-4, horizons 20..400 CFL steps, the CFL step as the features' unit of
time. Every step's features and Hessenberg matrix are logged. Labels: for
each step, the step h_j each grid dimension j admits under the run's
tolerance and the j with the most time advanced per unit of cost
(rfhip_krylov_labels), cost per step c0 + c1 j + c2 j (j+1) / 2 in
milliseconds, from the workload the proposer will run on (measure it with
`bench.py --proposer --cost-only`).

Usage:
  python -m rapidfem_amd.train_proposer --rows 50000 --cost C0 C1 C2 \
      [--out rapidfem_amd/proposer_weights.npz] [--epochs 60] [--max-seconds S]
Prints one JSON line with the data and training summary.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import math
import sys
import time

import numpy as np

from . import workloads as W
from .gpu import GpuContext, GpuOperator
from .proposer import WEIGHTS, ExpController, Proposer, features_and_targets, labels
from .session import TdSession

M_LOG = 60
LABEL_GRID = (10, 12, 16, 20, 24, 28, 32, 40, 48, 56, 60)


def log(msg: str) -> None:
    print(f"[train_proposer] {msg}", file=sys.stderr, flush=True)


# size classes of the training workloads: (nx, ny, nz) ranges, initial fields
# per mesh, share of the collection time
SIZES = {
    "small": (((8, 15), (4, 8), (4, 8)), 4, 0.35),
    "medium": (((16, 27), (10, 17), (10, 17)), 3, 0.35),
    "large": (((36, 45), (26, 33), (24, 31)), 2, 0.30),
}


def instance(rng: np.random.Generator, size: str = "small"):
    """One seeded stiff workload: a graded slab box of the size class, N=3."""
    (xr, yr, zr), _, _ = SIZES[size]
    while True:  # never the 40x30x28 box the gain is measured on
        nx, ny, nz = (int(rng.integers(*r)) for r in (xr, yr, zr))
        if (nx, ny, nz) != (40, 30, 28):
            break
    op, y0 = W.slab_box(nx, ny, nz, 3, seed=int(rng.integers(1 << 30)), ratio=float(rng.uniform(10, 30)),
                        eps_r=float(rng.uniform(4, 16)), sigma=float(rng.uniform(10, 100)))
    return op, y0


def initial_fields(op, y0, rng, count):
    """The workload's pulse plus pulses at random centres and frequencies."""
    x = op.node_coords()
    out = [y0]
    for _ in range(count - 1):
        c = rng.uniform(0.2, 0.8, 3)
        s = rng.uniform(0.04, 0.15)
        k = rng.uniform(1.0, 8.0)
        comp = int(rng.integers(0, 3))
        r2 = ((x - c) ** 2).sum(-1)
        y = np.zeros((op.n_elem, x.shape[1], 6))
        y[..., comp] = np.exp(-r2 / (2 * s * s)) * np.sin(2 * math.pi * k * x[..., (comp + 1) % 3])
        out.append(y.reshape(-1))
    return out


def collect(ctx, rows: int, cost, seed: int = 0, nthreads: int = 15, max_seconds: float = 1e9, shares=None,
            fields=None):
    """(features, targets, summary) of at least `rows` labelled steps, or of
    what max_seconds of collection gave."""
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    n_rows = n_steps = n_runs = n_rejected = 0
    per_size = {}
    t0 = time.time()
    pool = cf.ThreadPoolExecutor(max_workers=1)
    pending = []  # (future, logged steps)

    def drain(block):
        nonlocal n_rows
        keep = []
        for f, k in pending:
            if block or f.done():
                x, t = f.result()
                xs.append(x)
                ts.append(t)
                n_rows += len(x)
            else:
                keep.append((f, k))
        pending[:] = keep

    def label_job(steps, hess, rate, h_scale):
        hb, mb, _ = labels(steps, hess, rate, cost, m_grid=LABEL_GRID, nthreads=nthreads)
        return features_and_targets(steps, hb, mb, h_scale)

    # the size classes take turns within their shares of the time budget
    shares = dict(zip(SIZES, shares)) if shares else {k: v[2] for k, v in SIZES.items()}
    fields = dict(zip(SIZES, fields)) if fields else {k: v[1] for k, v in SIZES.items()}
    budget = {k: max(shares[k], 1e-9) * min(max_seconds, 1e7) for k in SIZES}
    used = {k: 0.0 for k in SIZES}
    while True:
        drain(False)
        if n_rows + sum(k for _, k in pending) >= rows or time.time() - t0 > max_seconds:
            drain(True)
            if n_rows >= rows or time.time() - t0 > max_seconds:
                break
        size = min(SIZES, key=lambda k: used[k] / budget[k])
        t_inst = time.time()
        op, y0 = instance(rng, size)
        g = GpuOperator(ctx, op, "fused_f64")
        cfl = TdSession(op, 1.0, "fused_f64", ctx).cfl_dt()
        for y in initial_fields(op, y0, rng, fields[size]):
            t_end = cfl * float(np.exp(rng.uniform(np.log(20.0), np.log(400.0))))
            tol = float(10.0 ** rng.uniform(-9, -4))
            ctl = ExpController(tol=tol, m=M_LOG, m_min=10, m_max=M_LOG, h_scale=cfl)
            _, steps, hess = g.expmv_adaptive(ctx, y, t_end, ctl, log_cap=4096, hess_dim=M_LOG)
            rate = tol * float(np.linalg.norm(y)) / t_end
            pending.append((pool.submit(label_job, steps, hess, rate, cfl), len(steps)))
            n_steps += len(steps)
            n_rejected += int((steps["n_judged"] > 1).sum())
            n_runs += 1
            per_size[size] = per_size.get(size, 0) + len(steps)
        used[size] += time.time() - t_inst
        del g
        log(f"{n_runs} runs ({size} {op.n_elem} tets), {n_steps} logged steps, {n_rows} labelled rows, "
            f"{time.time() - t0:.0f} s")
    pool.shutdown()
    x = np.concatenate(xs) if xs else np.zeros((0, 12), np.float32)
    t = np.concatenate(ts) if ts else np.zeros((0, 2), np.float32)
    return x, t, {"runs": n_runs, "logged_steps": n_steps, "steps_per_size": per_size,
                  "rejected_first_proposals": n_rejected, "rows": int(len(x)), "collect_s": round(time.time() - t0, 1)}


def size_weights(x: np.ndarray) -> np.ndarray:
    """Per-row weights that give each system-size class (from the n_dof
    feature: below 1e6, below 6.3e6, above) the same total weight, mean 1."""
    lg = x[:, 9].astype(np.float64) * 8.0  # log10 n_dof
    cls = np.digitize(lg, [6.0, 6.8])
    counts = np.bincount(cls, minlength=3).astype(np.float64)
    present = counts > 0
    w = (len(x) / present.sum()) / counts[cls]
    return w.astype(np.float32)


def train(ctx, x, t, epochs: int = 60, seed: int = 0, balance: bool = True):
    prop = Proposer(ctx, seed=seed)
    prop.fit_normalisation(x, t)
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(x))
    n_val = max(1, len(x) // 10)
    val, tr = idx[:n_val], idx[n_val:]
    l0, _ = prop.loss_grad(x[val], t[val])
    t0 = time.time()
    w = size_weights(x[tr]) if balance else None
    hist = prop.train(x[tr], t[tr], w, epochs=epochs, batch=256, lr=2e-3, seed=seed + 1)
    l1, _ = prop.loss_grad(x[val], t[val])
    pred = prop.forward(x[val])
    m_err = float(np.mean(np.abs(np.rint(pred[:, 1]) - t[val, 1])))
    h_err = float(np.sqrt(np.mean((pred[:, 0] - t[val, 0]) ** 2)))
    return prop, {"train_s": round(time.time() - t0, 2), "epochs": epochs, "val_loss_before": l0,
                  "val_loss_after": l1, "train_loss_first_epoch": float(hist[0]), "train_loss_last_epoch": float(hist[-1]),
                  "val_mean_abs_m_err": m_err, "val_rms_log_h_err": h_err}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=50_000)
    ap.add_argument("--cost", type=float, nargs=3, required=True, metavar=("C0", "C1", "C2"))
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--data-out", default="", help="also save the labelled rows (npz)")
    ap.add_argument("--data-in", nargs="*", default=[], help="labelled rows saved by earlier runs, trained on too")
    ap.add_argument("--max-seconds", type=float, default=1e9, help="stop collecting after this long")
    ap.add_argument("--shares", type=float, nargs=3, default=None, metavar=("SMALL", "MEDIUM", "LARGE"),
                    help="shares of the collection time per size class")
    ap.add_argument("--fields", type=int, nargs=3, default=None, metavar=("SMALL", "MEDIUM", "LARGE"),
                    help="initial fields per mesh of each size class")
    ap.add_argument("--no-balance", action="store_true", help="train without the size-class weights")
    a = ap.parse_args(argv)
    ctx = GpuContext()
    xin, tin = [], []
    for path in a.data_in:
        with np.load(path, allow_pickle=False) as z:
            xin.append(z["x"])
            tin.append(z["t"])
    n_in = int(sum(len(v) for v in xin))
    x, t, summ = collect(ctx, max(a.rows - n_in, 0), a.cost, a.seed, max_seconds=a.max_seconds, shares=a.shares,
                         fields=a.fields)
    log(f"collected {summ}")
    if a.data_out:
        np.savez_compressed(a.data_out, x=x, t=t)
    if xin:
        x = np.concatenate(xin + [x])
        t = np.concatenate(tin + [t])
        summ["rows_from_earlier_runs"] = n_in
        summ["rows"] = int(len(x))
    prop, tsum = train(ctx, x, t, a.epochs, a.seed, balance=not a.no_balance)
    prop.save(a.out)
    summ.update(tsum)
    summ["cost_model_ms"] = list(a.cost)
    summ["m_star_histogram"] = {int(k): int(v) for k, v in zip(*np.unique(t[:, 1].astype(int), return_counts=True))}
    print(json.dumps(summ), flush=True)


if __name__ == "__main__":
    main()
