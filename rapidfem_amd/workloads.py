"""Synthetic workloads: the meshes, materials and initial fields the
benchmark and the parity tests run on.

  cavity_te101  6x6x6 jittered Kuhn box (1296 tets, +-10% h), N=2, TE101
                E_y = sin(pi x) sin(pi z), vacuum, PEC, upwind.
  pulse_box     2x1x1 box, nx x ny x nz cells x 6 tets (60x30x30 = 324,000
                tets at full size), N=3, eps_r=4 on tets whose centroid has
                x in [0.8, 1.2], E_x = exp(-|r-c|^2/(2 s^2)) sin(2 pi 5 x),
                s=0.05, c=(0.4,0.5,0.5), PEC, upwind.
  modes_box     n^3 jittered unit box (55^3*6 = 998,250 tets at full size),
                N=4, sum of 16 PEC cavity modes with N(0,1) amplitudes.
  port_box      jittered box with a dielectric block and four port planes
                (the x and y walls) carrying TE10-shaped incident profiles:
                one mesh of the batching workload (port_batch: 8 seeded
                boxes of 20k-300k tets at orders 2/3/4, 4 excitations each).
  slab_box      unit box graded along x toward a thin slab at a seeded
                position (cell widths 20:1 between far field and slab),
                nx x ny x nz cells x 6 tets (40x30x28 = 201,600 tets at full
                size), N=3, eps_r=10 and sigma=50 in the slab, the pulse
                E_x of pulse_box centred off the slab: the stiff operator of
                the Krylov workload.
"""
from __future__ import annotations

import math

import numpy as np

from .maxwell import (Mesh, MaxwellOperator, boundary_faces_on_plane, port_face_data, structured_box,
                      structured_box_jittered)


def te101_field(op: MaxwellOperator) -> np.ndarray:
    x = op.node_coords()
    y = np.zeros((op.n_elem, op.re.n_nodes, 6))
    y[..., 1] = np.sin(math.pi * x[..., 0]) * np.sin(math.pi * x[..., 2])
    return y.reshape(-1)


def cavity_te101(n=6, order=2, amplitude=0.1, seed=1, flux_alpha=1.0) -> tuple[MaxwellOperator, np.ndarray]:
    mesh = structured_box_jittered(n, n, n, 1.0, 1.0, 1.0, amplitude, seed)
    op = MaxwellOperator(mesh, order, flux_alpha)
    return op, te101_field(op)


def pulse_box(nx=60, ny=30, nz=30, order=3, eps_r=4.0, flux_alpha=1.0) -> tuple[MaxwellOperator, np.ndarray]:
    mesh = structured_box(nx, ny, nz, 2.0, 1.0, 1.0)
    cx = mesh.nodes[mesh.tets].mean(axis=1)[:, 0]
    eps = np.ones((mesh.n_tets, 3))
    eps[(cx >= 0.8) & (cx <= 1.2)] = eps_r
    op = MaxwellOperator(mesh, order, flux_alpha, materials={"eps": eps})
    x = op.node_coords()
    r2 = (x[..., 0] - 0.4) ** 2 + (x[..., 1] - 0.5) ** 2 + (x[..., 2] - 0.5) ** 2
    y = np.zeros((op.n_elem, op.re.n_nodes, 6))
    y[..., 0] = np.exp(-r2 / (2 * 0.05 ** 2)) * np.sin(2 * math.pi * 5 * x[..., 0])
    return op, y.reshape(-1)


def slab_box(nx=40, ny=30, nz=28, order=3, seed=5, ratio=20.0, width=0.02, flux_alpha=1.0, eps_r=10.0,
             sigma=50.0) -> tuple[MaxwellOperator, np.ndarray]:
    rng = np.random.default_rng(seed)
    x0 = float(rng.uniform(0.35, 0.65))
    # cell widths along x: 1 far from the slab, 1/ratio at it (Gaussian
    # profile three slab widths wide), scaled to the unit length
    xc = (np.arange(nx) + 0.5) / nx
    w = 1.0 / (1.0 + (ratio - 1.0) * np.exp(-(((xc - x0) / (3.0 * width)) ** 2)))
    xs = np.concatenate([[0.0], np.cumsum(w)])
    xs /= xs[-1]
    base = structured_box(nx, ny, nz, 1.0, 1.0, 1.0)
    nodes = base.nodes.copy()
    i = np.rint(nodes[:, 0] * nx).astype(np.int64)
    nodes[:, 0] = xs[i]
    mesh = Mesh(nodes, base.tets)
    cx = nodes[mesh.tets].mean(axis=1)[:, 0]
    slab = np.abs(cx - x0) <= 0.5 * width + 0.5 * np.diff(xs).min()
    eps = np.ones((mesh.n_tets, 3))
    eps[slab] = eps_r
    sig = np.where(slab, sigma, 0.0)
    op = MaxwellOperator(mesh, order, flux_alpha, materials={"eps": eps, "sigma": sig})
    x = op.node_coords()
    xc0 = x0 - 0.2 if x0 > 0.5 else x0 + 0.2
    r2 = (x[..., 0] - xc0) ** 2 + (x[..., 1] - 0.5) ** 2 + (x[..., 2] - 0.5) ** 2
    y = np.zeros((op.n_elem, op.re.n_nodes, 6))
    y[..., 0] = np.exp(-r2 / (2 * 0.05 ** 2)) * np.sin(2 * math.pi * 5 * x[..., 0])
    return op, y.reshape(-1)


# Port walls of port_box: (axis, side) for x = 0, x = Lx, y = 0, y = Ly.
PORT_WALLS = ((0, 0), (0, 1), (1, 0), (1, 1))


def te10_profile(axis: int, side: int, lengths):
    """Incident TE10-shaped profile on the wall x[axis] = 0 (side 0) or
    = lengths[axis] (side 1): e_t along the wall's second transverse axis v,
    sin(pi x_u / L_u) across the first one u, and h_t = k x e_t with k the
    unit normal into the box (a unit-impedance wave entering the domain)."""
    u, v = (axis + 1) % 3, (axis + 2) % 3
    k = np.zeros(3)
    k[axis] = 1.0 if side == 0 else -1.0

    def prof(x):
        e = np.zeros_like(x)
        e[:, v] = np.sin(math.pi * x[:, u] / lengths[u])
        return e, np.cross(k, e)

    return prof


def port_box(nx, ny, nz, order, seed, lengths=(1.0, 1.0, 1.0), eps_r=None, flux_alpha=1.0):
    """Jittered box (+-10% h, walls kept planar) with a dielectric block
    (eps_r seeded in [2, 6] unless given) in a seeded sub-box, PEC walls
    except the four port walls of PORT_WALLS, each tagged as its own port
    (ids 0..3, rhs.rs:862-880). Returns (operator, ports) with
    ports[r] = (elem, local_face, profiles, weights) for GpuOperator.set_port."""
    rng = np.random.default_rng(seed)
    mesh = structured_box_jittered(nx, ny, nz, *lengths, 0.1, seed)
    cen = mesh.nodes[mesh.tets].mean(axis=1)
    lo = rng.uniform(0.2, 0.5, 3) * np.asarray(lengths)
    hi = lo + rng.uniform(0.2, 0.4, 3) * np.asarray(lengths)
    inside = np.all((cen >= lo) & (cen <= hi), axis=1)
    eps = np.ones((mesh.n_tets, 3))
    eps[inside] = float(rng.uniform(2.0, 6.0)) if eps_r is None else eps_r
    probe = MaxwellOperator(mesh, order, flux_alpha, materials={"eps": eps})
    walls = [boundary_faces_on_plane(probe, axis, 0.0 if side == 0 else lengths[axis]) for axis, side in PORT_WALLS]
    face_port = np.full((mesh.n_tets, 4), -1, np.int32)
    for r, faces in enumerate(walls):
        for e, f in faces:
            face_port[e, f] = r
    op = MaxwellOperator(mesh, order, flux_alpha, materials={"eps": eps}, face_port=face_port)
    ports = [port_face_data(op, faces, te10_profile(axis, side, lengths))
             for faces, (axis, side) in zip(walls, PORT_WALLS)]
    return op, ports


def port_batch_specs(seed=2024, n_meshes=8, tets=(20_000, 300_000)):
    """Seeded sizes of the batching workload: (nx, ny, nz, order) per mesh,
    tet counts log-uniform in `tets`, orders 2, 3, 4 in turn."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n_meshes):
        target = float(np.exp(rng.uniform(np.log(tets[0]), np.log(tets[1]))))
        c = (target / 6.0) ** (1.0 / 3.0)
        asp = rng.uniform(0.8, 1.25)
        nx = max(2, int(round(c * asp)))
        ny = max(2, int(round(c / asp)))
        nz = max(2, int(round(target / (6.0 * nx * ny))))
        out.append((nx, ny, nz, (2, 3, 4)[b % 3]))
    return out


def port_waveforms(seed, n, t0_scale=4.0):
    """n Gaussian-modulated pulses (session.rs:242-261) with centre
    frequencies seeded in [3, 8]: (t0, tau, f0) with tau = 1.5 / f0 and
    t0 = t0_scale * tau."""
    rng = np.random.default_rng(seed)
    f0 = rng.uniform(3.0, 8.0, n)
    tau = 1.5 / f0
    return [(float(t0_scale * t), float(t), float(f)) for t, f in zip(tau, f0)]


def cavity_modes_field(op: MaxwellOperator, n_modes=16, seed=0, lengths=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Sum of PEC box cavity modes (standing E, H = 0) with N(0,1) amplitudes.

    Mode (m, n, p): E = (A cos(kx x) sin(ky y) sin(kz z), B sin cos sin,
    C sin sin cos) with k = pi (m/lx, n/ly, p/lz) and (A, B, C) . k = 0, so
    E is divergence-free and its tangential part vanishes on every wall.
    """
    rng = np.random.default_rng(seed)
    lx, ly, lz = lengths
    cands = [(m, n, p) for m in range(0, 4) for n in range(0, 4) for p in range(0, 4)
             if (m > 0) + (n > 0) + (p > 0) >= 2]
    picks = rng.choice(len(cands), size=n_modes, replace=False)
    modes = []
    for idx in picks:
        m, n, p = cands[idx]
        k = np.array([math.pi * m / lx, math.pi * n / ly, math.pi * p / lz])
        pol = rng.normal(size=3)
        pol -= k * (pol @ k) / (k @ k)
        pol /= np.linalg.norm(pol)
        modes.append(((m, n, p), rng.normal() * pol))
    x = op.node_coords().reshape(-1, 3)
    y = np.zeros((x.shape[0], 6))
    step = 1 << 20
    for a in range(0, x.shape[0], step):
        xc = x[a:a + step]
        # cos/sin(pi q x_d / l_d) for q = 0..3 on this chunk, per axis
        trig = []
        for d, ld in enumerate((lx, ly, lz)):
            ang = (math.pi / ld) * xc[:, d]
            trig.append([(np.cos(q * ang), np.sin(q * ang)) for q in range(4)])
        for (m, n, p), amp in modes:
            (cx, sx), (cy, sy), (cz, sz) = trig[0][m], trig[1][n], trig[2][p]
            y[a:a + step, 0] += amp[0] * cx * sy * sz
            y[a:a + step, 1] += amp[1] * sx * cy * sz
            y[a:a + step, 2] += amp[2] * sx * sy * cz
    return y.reshape(-1)


def modes_box(n=55, order=4, amplitude=0.1, seed=3, flux_alpha=1.0) -> tuple[MaxwellOperator, np.ndarray]:
    mesh = structured_box_jittered(n, n, n, 1.0, 1.0, 1.0, amplitude, seed)
    op = MaxwellOperator(mesh, order, flux_alpha)
    return op, cavity_modes_field(op, 16, seed)


# Algorithmic work per DOF-stage of the fused stage kernel: flops (volume
# 36 Np^2, metric+curl 114 Np, flux 78*4Nfp, LIFT 48 Np Nfp, materials
# 18 Np, per element, plus 5 for the update, per DOF) and HBM bytes
# (read y, write y, read p, write p + per-element geometry).
def flops_per_dof_stage(order: int) -> float:
    np_ = (order + 1) * (order + 2) * (order + 3) // 6
    nfp = (order + 1) * (order + 2) // 2
    per_elem = 36 * np_ ** 2 + 114 * np_ + 78 * 4 * nfp + 48 * np_ * nfp + 18 * np_
    return per_elem / (6 * np_) + 5.0


def bytes_per_dof_stage(order: int, form: str) -> float:
    np_ = (order + 1) * (order + 2) * (order + 3) // 6
    if form in ("fused_f64c", "fused_bf16x4", "fused_bf16x3", "fused_bf16x2"):
        state = 6 + 6 + 4 + 4  # y (f32+bf16) read+write, p (f32) read+write
        geom = 9 * 8 + 4 * 4 + 4  # J^-1 f64, 4 face codes, material index
    elif form == "fused_f64":
        state = 4 * 8
        geom = 9 * 8 + 4 * 4 + 4
    elif form == "plain_f64":
        state = 7 * 8  # apply: read y, write k; stage: read p, k, y, write p, y
        geom = (9 + 12 + 12 + 4) * 8 + (4 + 4 + 4) * 4 + 4 * ((order + 1) * (order + 2) // 2) * 4
    else:  # plain_f32
        state = 7 * 4
        geom = (9 + 12 + 12 + 4) * 4 + (4 + 4 + 4) * 4 + 4 * ((order + 1) * (order + 2) // 2) * 4
    return state + geom / (6 * np_)
