"""CPU tests of the FP64 oracle: the reference's golden vectors, exact
rational tables, closed forms and invariants of the DG Maxwell operator and
its time integrators.

Each test names the reference test or source lines it mirrors. A plausible
slip in the oracle (a dropped flux term, a sign, a transposed operand, a
wrong node order) fails at least one of them.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy.linalg import expm as scipy_expm

from conftest import GOLDEN

EXACT = json.load(open(os.path.join(GOLDEN, "exact_basis.json")))["orders"]


def dense_a(op):
    """Dense A from unit vectors: column j = A e_j (rhs.rs assemble_dense)."""
    n = op.n_dof
    a = np.empty((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        a[:, j] = op.apply(e)
        e[j] = 0.0
    return a


def energy_mass(oracle, op):
    """Dense M~ = blockdiag(|det J_e| w_c M_ref) of the energy
    0.5 y^T M~ y (rhs.rs:1509-1544 with eps on E, mu on H); the test below
    checks it against field_energy."""
    ex = op.export()
    r = oracle.refelem(op.order)
    n = r["np"]
    mt = np.zeros((op.n_dof, op.n_dof))
    eps = 1.0 / ex["inv_eps"]
    mu = 1.0 / ex["inv_mu"]
    for e in range(op.K):
        w = np.concatenate([eps[e], mu[e]]) * abs(ex["det"][e])
        for c in range(6):
            ix = (e * n + np.arange(n)) * 6 + c
            mt[np.ix_(ix, ix)] = w[c] * r["mass"]
    return mt


def wave_6x6():
    """lserk4_convergence_test.rs / propagator_expm_test.rs operator
    A = [[0, -L], [L, 0]], L = tridiag(-1, 2, -1) on 3 nodes."""
    lap = np.array([[2.0, -1, 0], [-1, 2, -1], [0, -1, 2]])
    a = np.zeros((6, 6))
    a[:3, 3:] = -lap
    a[3:, :3] = lap
    return a


# ---------------------------------------------------------------- a1 basis

@pytest.mark.parametrize("p", [1, 2, 3])
def test_basis_matches_reference_goldens(oracle, p):
    """dg_basis_golden_test.rs: mass, Dr, Ds, Dt at p = 1..3 (1e-9)."""
    g = json.load(open(os.path.join(GOLDEN, "dg_basis_golden.json")))
    tol = g["tol"]
    gold = g["orders"][f"p{p}"]
    r = oracle.refelem(p)
    for key in ("mass", "dr", "ds", "dt"):
        want = np.array(gold[key])
        got = r[key].ravel()
        err = np.abs(got - want).max() / max(np.abs(want).max(), 1e-300)
        assert err < tol, (p, key, err)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_basis_matches_exact_rationals(oracle, p):
    """p = 4 and LIFT at every order are pinned by the exact-rational
    construction (tests/golden/make_exact_basis.py), which itself reproduces
    the reference goldens at p <= 3 (test below)."""
    ex = EXACT[f"p{p}"]
    r = oracle.refelem(p)
    assert (r["face_nodes"] == np.array(ex["face_nodes"])).all()
    for key in ("mass", "dr", "ds", "dt", "lift"):
        want = np.array(ex[key])
        err = np.abs(r[key].ravel() - want).max() / np.abs(want).max()
        assert err < 1e-9, (p, key, err)


def test_exact_tables_reproduce_reference_goldens():
    g = json.load(open(os.path.join(GOLDEN, "dg_basis_golden.json")))["orders"]
    for p in (1, 2, 3):
        for key in ("mass", "dr", "ds", "dt"):
            assert np.abs(np.array(EXACT[f"p{p}"][key]) - np.array(g[f"p{p}"][key])).max() < 1e-14


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_basis_invariants(oracle, p):
    """dg_basis.rs:395-556: M symmetric, sum M = 1/6 (tet volume), D rows sum
    to 0, D differentiates polynomials of degree <= p exactly, and the face
    rows of M*LIFT integrate face traces."""
    r = oracle.refelem(p)
    m, x = r["mass"], r["nodes"]
    assert np.abs(m - m.T).max() < 1e-9 * np.abs(m).max()
    assert abs(m.sum() - 1.0 / 6.0) < 1e-12
    rng = np.random.default_rng(p)
    for d in (r["dr"], r["ds"], r["dt"]):
        assert np.abs(d.sum(axis=1)).max() < 1e-11
    # f = sum c_abc r^a s^b t^c of degree p
    coeffs = {(a, b, c): rng.standard_normal() for a in range(p + 1) for b in range(p + 1 - a)
              for c in range(p + 1 - a - b)}
    f = sum(v * x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c for (a, b, c), v in coeffs.items())
    grads = [sum(v * a * x[:, 0] ** max(a - 1, 0) * x[:, 1] ** b * x[:, 2] ** c for (a, b, c), v in coeffs.items()),
             sum(v * b * x[:, 0] ** a * x[:, 1] ** max(b - 1, 0) * x[:, 2] ** c for (a, b, c), v in coeffs.items()),
             sum(v * c * x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** max(c - 1, 0) for (a, b, c), v in coeffs.items())]
    for d, g in zip((r["dr"], r["ds"], r["dt"]), grads):
        assert np.abs(d @ f - g).max() < 1e-9 * max(1.0, np.abs(g).max())
    # 1^T M LIFT[:, f*Nfp + m] = integral over the face of the face basis
    # function m = its face quadrature weight; the weights of one face sum to
    # the unit right triangle's area 1/2.
    nfp = r["nfp"]
    ml = m @ r["lift"]
    w = r["fweights"]
    assert np.abs(ml.sum(axis=0) - w.ravel()).max() < 1e-10
    assert np.abs(w.sum(axis=1) - 0.5).max() < 1e-12
    # a constant face trace lifts onto M^{-1} E 1, whose volume integral is
    # the face area again
    for f in range(4):
        assert abs(ml[:, f * nfp:(f + 1) * nfp].sum() - 0.5) < 1e-10


# -------------------------------------------------------------- a2/a6 curl

def test_curl_matches_reference_golden(oracle):
    """rhs_curl_golden_test.rs: p = 2 on a sheared tet (1e-9)."""
    g = json.load(open(os.path.join(GOLDEN, "rhs_curl_golden.json")))
    got = oracle.element_curl(g["p"], np.array(g["tet"]), np.array(g["field"]))
    want = np.array(g["curl"])
    assert np.abs(got - want).max() / np.abs(want).max() < g["tol"]


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_curl_exact_on_polynomials(oracle, p):
    """rhs.rs:1779-1827: the discrete curl of a polynomial field of degree
    <= p is exact on any straight-sided tet."""
    rng = np.random.default_rng(10 + p)
    tet = np.array([[0.1, 0.2, 0.0], [1.3, 0.1, 0.2], [0.2, 1.1, -0.1], [0.3, 0.4, 0.9]])
    r = oracle.refelem(p)
    jac = (tet[1:] - tet[0]).T
    xyz = r["nodes"] @ jac.T + tet[0]
    # F = (P, Q, R) with P, Q, R linear combinations of x^a y^b z^c, a+b+c <= p
    mon = [(a, b, c) for a in range(p + 1) for b in range(p + 1 - a) for c in range(p + 1 - a - b)]
    cf = rng.standard_normal((3, len(mon)))

    def ev(comp, da=(0, 0, 0)):
        out = np.zeros(len(xyz))
        for v, (a, b, c) in zip(cf[comp], mon):
            e = [a, b, c]
            coef = v
            for ax in range(3):
                if da[ax]:
                    coef *= e[ax]
                    e[ax] -= 1
            if min(e) < 0:
                continue
            out += coef * xyz[:, 0] ** e[0] * xyz[:, 1] ** e[1] * xyz[:, 2] ** e[2]
        return out

    field = np.stack([ev(0), ev(1), ev(2)], axis=1).ravel()
    dx, dy, dz = (1, 0, 0), (0, 1, 0), (0, 0, 1)
    curl = np.stack([ev(2, dy) - ev(1, dz), ev(0, dz) - ev(2, dx), ev(1, dx) - ev(0, dy)], axis=1).ravel()
    got = oracle.element_curl(p, tet, field)
    assert np.abs(got - curl).max() < 1e-9 * max(1.0, np.abs(curl).max())


# ------------------------------------------------------------ a3/a4 meshes

def test_box_mesh_kuhn_split(oracle):
    """mesh_gen.rs:26-77: six tets per cell, sorted vertex ids, total volume
    = box volume, every tet positively sized."""
    xyz, tets = oracle.box_mesh(3, 2, 4, 2.0, 1.0, 1.5)
    assert tets.shape == (3 * 2 * 4 * 6, 4)
    assert (np.diff(tets, axis=1) > 0).all()
    v = xyz[tets]
    vol = np.abs(np.einsum("kij,kij->k", np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])[:, None, :],
                           (v[:, 3] - v[:, 0])[:, None, :])) / 6.0
    assert (vol > 0).all()
    assert abs(vol.sum() - 3.0) < 1e-12
    # node id (k*nyp + j)*nxp + i
    assert np.allclose(xyz[(1 * 3 + 2) * 4 + 3], [2.0, 1.0, 0.375])


def test_jittered_box_moves_only_interior_nodes(oracle):
    """mesh_gen.rs:86-139: boundary components stay on the box faces, the
    offset is at most amplitude*h per axis, and the hash is deterministic."""
    a, _ = oracle.box_mesh(4, 4, 4, 1.0, 1.0, 1.0)
    b, _ = oracle.box_mesh(4, 4, 4, 1.0, 1.0, 1.0, amplitude=0.25, seed=7, jittered=True)
    c, _ = oracle.box_mesh(4, 4, 4, 1.0, 1.0, 1.0, amplitude=0.25, seed=7, jittered=True)
    d, _ = oracle.box_mesh(4, 4, 4, 1.0, 1.0, 1.0, amplitude=0.25, seed=8, jittered=True)
    assert (b == c).all() and not (b == d).all()
    on_bnd = (a == 0.0) | (a == 1.0)
    assert (b[on_bnd] == a[on_bnd]).all()
    off = np.abs(b - a)
    assert off.max() <= 0.25 * 0.25 + 1e-15 and off[~on_bnd].max() > 0.02


def test_operator_face_topology(oracle):
    """core/topology.rs + rhs.rs:805-880 on a jittered box: neighbour links
    are reciprocal, perm maps each face node onto the node at the same
    physical point, normals are opposite across a face, fscale = 2*area/|det|."""
    xyz, tets = oracle.box_mesh(3, 3, 3, 1.0, 1.0, 1.0, amplitude=0.2, seed=3, jittered=True)
    op = oracle.Operator(xyz, tets, 3)
    ex = op.export()
    r = oracle.refelem(3)
    K = op.K
    nbr, lf, perm = ex["nbr"], ex["nbr_lf"], ex["perm"]
    n_bnd = 0
    for e in range(K):
        v = xyz[tets[e]]
        jac = (v[1:] - v[0]).T
        for f in range(4):
            fn = r["face_nodes"][f]
            if nbr[e, f] < 0:
                n_bnd += 1
                continue
            en, g = nbr[e, f], lf[e, f]
            assert nbr[en, g] == e and lf[en, g] == f
            assert np.allclose(ex["normal"][e, f], -ex["normal"][en, g], atol=1e-14)
            assert sorted(perm[e, f]) == list(range(len(fn)))
            vn = xyz[tets[en]]
            jn = (vn[1:] - vn[0]).T
            p_own = r["nodes"][fn] @ jac.T + v[0]
            p_nbr = r["nodes"][r["face_nodes"][g][perm[e, f]]] @ jn.T + vn[0]
            assert np.abs(p_own - p_nbr).max() < 1e-12
    # boundary faces: 2 triangles per cell face on the box surface
    assert n_bnd == 6 * 9 * 2
    # fscale = 2*area/|det J|
    e, f = 5, 3
    v = xyz[tets[e]]
    tri = v[[[0, 1, 2], [0, 2, 3], [0, 3, 1], [1, 2, 3]][f]]
    area = 0.5 * np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
    assert abs(ex["fscale"][e, f] - 2 * area / abs(ex["det"][e])) < 1e-12
    nrm = ex["normal"][e, f]
    opp = v[[3, 1, 2, 0][f]]
    assert np.dot(nrm, tri[0] - opp) > 0 and abs(np.linalg.norm(nrm) - 1) < 1e-14


# ------------------------------------------------------- a7/a8 the operator

def test_central_flux_is_skew_in_energy_inner_product(oracle):
    """rhs.rs:1830-1868: with alpha = 0, M~ A is skew-symmetric (1e-9),
    including heterogeneous anisotropic materials (rhs.rs:1900, :2031)."""
    xyz, tets = oracle.box_mesh(1, 1, 1)
    K = len(tets)
    eps = np.array([[2.0 + i, 3.0, 1.5] for i in range(K)])
    mu = np.array([[1.0, 1.2 + 0.3 * i, 2.0] for i in range(K)])
    for kw in ({}, dict(eps=eps, mu=mu)):
        op = oracle.Operator(xyz, tets, 2, flux_alpha=0.0, **kw)
        a = dense_a(op)
        mt = energy_mass(oracle, op)
        y = oracle.normal_vector(op.n_dof, 4)
        assert abs(0.5 * y @ mt @ y - op.field_energy(y)) < 1e-12 * op.field_energy(y)
        ma = mt @ a
        scale = np.abs(ma).max()
        assert np.abs(ma + ma.T).max() < 1e-9 * max(scale, 1.0)
        # upwind adds a negative semi-definite symmetric part
        up = oracle.Operator(xyz, tets, 2, flux_alpha=1.0, **kw)
        mu_ = energy_mass(oracle, up) @ dense_a(up)
        sym = 0.5 * (mu_ + mu_.T)
        assert np.linalg.eigvalsh(sym).max() < 1e-9 * scale
        assert np.linalg.eigvalsh(sym).min() < -1e-3 * scale


def test_field_energy_matches_dense_quadratic_form(oracle):
    """rhs.rs:1509-1544 field_energy = 0.5 y^T M~ y with
    M~ = blockdiag(|det J| w_c M_ref)."""
    xyz, tets = oracle.box_mesh(2, 1, 1, 2.0, 1.0, 1.0, amplitude=0.1, seed=2, jittered=True)
    K = len(tets)
    eps = np.array([[1.0 + 0.5 * i, 2.0, 3.0] for i in range(K)])
    mu = np.array([[1.5, 1.0, 1.0 + 0.1 * i] for i in range(K)])
    op = oracle.Operator(xyz, tets, 2, eps=eps, mu=mu)
    ex = op.export()
    r = oracle.refelem(2)
    n = r["np"]
    y = oracle.normal_vector(op.n_dof, 5)
    want = 0.0
    for e in range(K):
        blk = y[e * n * 6:(e + 1) * n * 6].reshape(n, 6)
        w = np.concatenate([eps[e], mu[e]])
        for c in range(6):
            want += 0.5 * abs(ex["det"][e]) * w[c] * blk[:, c] @ r["mass"] @ blk[:, c]
    assert abs(op.field_energy(y) - want) < 1e-12 * want


def test_cavity_fundamental_is_pi_sqrt2(oracle):
    """rhs.rs:2113-2150: unit PEC cube, 2x2x2 cells, N = 2, upwind: the
    least-damped non-static eigenvalue has |Im| = pi*sqrt(2) within 5% and
    Re < 0."""
    xyz, tets = oracle.box_mesh(2, 2, 2)
    op = oracle.Operator(xyz, tets, 2, flux_alpha=1.0)
    ev = np.linalg.eigvals(dense_a(op))
    ev = ev[np.abs(ev.imag) > 1.0]
    fund = ev[np.argmax(ev.real)]
    want = math.pi * math.sqrt(2.0)
    assert fund.real < 0.0
    assert abs(abs(fund.imag) - want) / want < 0.05


def test_conductivity_damps_at_analytic_rate(oracle):
    """rhs.rs:1997-2030: uniform sigma shifts the least-damped mode's real
    part by -sigma/(2 eps) (10%)."""
    xyz, tets = oracle.box_mesh(1, 1, 1)

    def least_damped(sig):
        op = oracle.Operator(xyz, tets, 2, flux_alpha=1.0, sigma=np.full(len(tets), sig))
        ev = np.linalg.eigvals(dense_a(op))
        return ev[np.abs(ev.imag) > 1.0].real.max()

    delta = least_damped(0.2) - least_damped(0.0)
    assert abs(delta + 0.1) / 0.1 < 0.1


def test_dense_operator_rules_brute_force(oracle):
    """Brute force on a two-tet mesh: every entry of A recomputed from the
    flux rules of rhs.rs:1341-1454 written out per face node."""
    xyz = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]])
    tets = np.array([[0, 1, 2, 3], [1, 2, 3, 4]])
    p = 1
    alpha = 0.7
    eps = np.array([[1.0, 2.0, 3.0], [2.0, 1.0, 1.5]])
    mu = np.array([[1.0, 1.5, 1.0], [0.5, 1.0, 2.0]])
    op = oracle.Operator(xyz, tets, p, flux_alpha=alpha, eps=eps, mu=mu)
    a = dense_a(op)
    r = oracle.refelem(p)
    ex = op.export()
    n, nfp = r["np"], r["nfp"]
    want = np.zeros_like(a)

    def idx(e, i, c):
        return (e * n + i) * 6 + c

    def cross(u, v):
        return np.cross(u, v)

    for e in range(2):
        ji = ex["jinv"][e]
        ddx = [sum(ji[k][d] * D for k, D in enumerate((r["dr"], r["ds"], r["dt"]))) for d in range(3)]
        for i in range(n):
            for j in range(n):
                # dE = curl H / eps, dH = -curl E / mu
                for (cE, dd, cH, sgn) in ((0, 1, 5, 1), (0, 2, 4, -1), (1, 2, 3, 1), (1, 0, 5, -1),
                                          (2, 0, 4, 1), (2, 1, 3, -1)):
                    want[idx(e, i, cE), idx(e, j, cH)] += sgn * ddx[dd][i, j] / eps[e, cE]
                    want[idx(e, i, cH), idx(e, j, cE)] += sgn * ddx[dd][i, j] / mu[e, cH - 3]
        for f in range(4):
            nv = ex["normal"][e, f]
            fs = ex["fscale"][e, f]
            en = ex["nbr"][e, f]
            for m in range(nfp):
                vi = r["face_nodes"][f][m]
                # jump [E] = E- - E+, [H] likewise, as linear maps of y
                for c in range(3):
                    # unit jumps: response of the flux to a unit [E]_c / [H]_c
                    uE = np.zeros(3)
                    uE[c] = 1.0
                    fe_E = alpha * cross(nv, cross(nv, uE))
                    fh_E = cross(nv, uE)
                    fe_H = -cross(nv, uE)
                    fh_H = alpha * cross(nv, cross(nv, uE))
                    if en < 0:
                        # PEC ghost: [E] = 2 E_t = 2(E - (n.E)n), [H] = 2 (n.H) n
                        jE = 2.0 * (uE - nv * nv[c])
                        jH = 2.0 * nv * nv[c]
                        fe_E = -cross(nv, np.zeros(3)) + alpha * cross(nv, cross(nv, jE))
                        fh_E = cross(nv, jE)
                        fe_H = -cross(nv, jH)
                        fh_H = alpha * cross(nv, cross(nv, jH))
                    srcs = [(idx(e, vi, c), 1.0, c)]
                    if en >= 0:
                        vj = r["face_nodes"][ex["nbr_lf"][e, f]][ex["perm"][e, f][m]]
                        srcs.append((idx(en, vj, c), -1.0, c))
                    for i in range(n):
                        w = 0.5 * fs * r["lift"][i, f * nfp + m]
                        for col, s, cc in srcs:
                            for d in range(3):
                                want[idx(e, i, d), col] += s * w * fe_E[d] / eps[e, d]
                                want[idx(e, i, 3 + d), col] += s * w * fh_E[d] / mu[e, d]
                                want[idx(e, i, d), col + 3] += s * w * fe_H[d] / eps[e, d]
                                want[idx(e, i, 3 + d), col + 3] += s * w * fh_H[d] / mu[e, d]
    assert np.abs(a - want).max() < 1e-11 * np.abs(want).max()


# --------------------------------------------------------------- a9 LSERK4

def test_lserk4_matches_exponential_and_is_fourth_order(oracle):
    """lserk4_convergence_test.rs: 6x6 skew operator, T = 2, 4096 steps
    < 1e-10 rel, observed order in [3.6, 4.2] over 16..512 steps."""
    a = wave_6x6()
    y0 = (np.arange(6) + 1.0) / 7.0
    exact = scipy_expm(2.0 * a) @ y0
    fine = oracle.lserk_dense(a, y0, 2.0 / 4096, 4096)
    assert np.linalg.norm(fine - exact) / np.linalg.norm(exact) < 1e-10
    errs = [np.linalg.norm(oracle.lserk_dense(a, y0, 2.0 / s, s) - exact) for s in (16, 32, 64, 128, 256, 512)]
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert ((orders > 3.6) & (orders < 4.2)).all(), orders


def test_lserk4_driven_point_and_vector_sources(oracle):
    """explicit.rs:44-56 Source::add_to: a point source adds g to one DOF, a
    vector source adds g*b; with A = 0 the step integrates the held source
    exactly: y += dt * g * b."""
    xyz, tets = oracle.box_mesh(1, 1, 1)
    op = oracle.Operator(xyz, tets, 1)
    y = np.zeros(op.n_dof)
    # one tiny step from rest: y = dt*g*b + O(dt^2 |A|)
    out = op.lserk(y, 1e-6, 1, drive=("point", 17), g=np.array([2.0]))
    assert abs(out[17] - 2e-6) < 1e-10 and np.abs(np.delete(out, 17)).max() < 1e-10
    b = oracle.normal_vector(op.n_dof, 3)
    out2 = op.lserk(y, 1e-6, 1, drive=("vector", b), g=np.array([0.5]))
    assert np.abs(out2 - 0.5e-6 * b).max() < 1e-10 * np.abs(b).max()


def test_lserk_on_mesh_matches_dense_exponential(oracle):
    """The matrix-free LSERK4 transient on a mesh operator against
    scipy expm of the dense operator."""
    xyz, tets = oracle.box_mesh(1, 1, 1, amplitude=0.0)
    op = oracle.Operator(xyz, tets, 2)
    a = dense_a(op)
    y0 = oracle.normal_vector(op.n_dof, 1)
    dt, steps = 0.002, 50
    got = op.lserk(y0, dt, steps)
    want = scipy_expm(dt * steps * a) @ y0
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-7


# ------------------------------------------------------------ a12 KCL + PI

def _rotation(omega=1.3):
    return np.array([[0.0, -omega], [omega, 0.0]])


def test_kcl_is_fourth_order():
    """explicit_adaptive.rs:160-187: rotation ODE, error ratio 20 -> 40 steps
    > 12."""
    from oracle import oracle as O

    a = _rotation()
    t_end = 1.7

    def integrate(n):
        y = np.array([1.0, 0.4])
        for _ in range(n):
            y, _ = O.kcl_step_dense(a, y, t_end / n)
        return y

    exact = scipy_expm(t_end * a) @ np.array([1.0, 0.4])
    rate = np.linalg.norm(integrate(20) - exact) / np.linalg.norm(integrate(40) - exact)
    assert rate > 12.0


def test_kcl_embedded_error_rate():
    """explicit_adaptive.rs:189-216: embedded error drops ~16x per halving."""
    from oracle import oracle as O

    a = _rotation()
    e1 = np.linalg.norm(O.kcl_step_dense(a, np.array([1.0, 0.4]), 0.05)[1])
    e2 = np.linalg.norm(O.kcl_step_dense(a, np.array([1.0, 0.4]), 0.025)[1])
    assert e1 / e2 > 10.0


def test_kcl_matches_exponential_below_cfl(oracle):
    """explicit_adaptive.rs:219-254: one KCL step at dt = 1e-3 on the 2x2x2
    N = 2 operator agrees with expmv (m = 40) to 1e-9."""
    xyz, tets = oracle.box_mesh(2, 2, 2)
    op = oracle.Operator(xyz, tets, 2, flux_alpha=1.0)
    y0 = np.sin(0.3 + 0.017 * np.arange(op.n_dof))
    y_rk, _ = op.kcl_step(y0, 1e-3)
    y_exp, _ = op.expmv(y0, 1e-3, 40)
    assert np.linalg.norm(y_rk - y_exp) < 1e-9 * np.linalg.norm(y_exp)


def test_controller_closed_forms(oracle):
    """session.rs:141-186: defaults, RMS error norm and the clamped PI
    factor."""
    c = oracle.Controller.default()
    assert (c.atol, c.rtol, c.safety, c.growth_limit, c.shrink_limit) == (1e-8, 1e-4, 0.9, 5.0, 0.2)
    assert abs(c.pi_alpha - 0.175) < 1e-16 and abs(c.pi_beta - 0.1) < 1e-16
    y0 = np.array([1.0, -2.0, 0.0])
    y1 = np.array([1.5, -1.0, 0.0])
    err = np.array([3e-4, -4e-4, 1e-8])
    sc = c.atol + c.rtol * np.maximum(np.abs(y0), np.abs(y1))
    want = math.sqrt(np.mean((err / sc) ** 2))
    assert abs(oracle.err_norm(c, y0, y1, err) - want) < 1e-14 * want
    assert oracle.err_norm(c, y0, np.array([1.0, np.inf, 0.0]), err) == math.inf
    assert abs(oracle.factor(c, 1.0, 0.0, False) - 0.9) < 1e-15
    assert abs(oracle.factor(c, 0.5, 0.25, False) - 0.9 * 0.5 ** -0.175 * 0.25 ** 0.1) < 1e-14
    assert oracle.factor(c, 1e-30, 1.0, False) == 5.0  # growth clamp (err floored at 1e-12)
    assert oracle.factor(c, 1e6, 1.0, True) == 0.2     # shrink clamp on reject
    assert abs(oracle.factor(c, 2.0, 0.5, True) - 0.9 * 2.0 ** -0.175) < 1e-15


def test_adaptive_run_counts_and_accuracy(oracle):
    """session.rs:781-827 adaptive frames: the adaptive run tracks the
    exponential within the controller tolerance and accepts >= one step per
    frame."""
    xyz, tets = oracle.box_mesh(2, 2, 2)
    op = oracle.Operator(xyz, tets, 2)
    y0 = oracle.normal_vector(op.n_dof, 7)
    y, acc, rej, hmin, hmax = op.adaptive(y0, 0.05, 6)
    ref, _ = op.expmv(y0, 0.3, 40, tol=1e-12)
    assert acc >= 6 and rej >= 0 and 0 < hmin <= hmax
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-4


# ----------------------------------------------------- a10 CFL, session bits

def test_normal_vector_statistics(oracle):
    """session.rs:415-431: deterministic per seed, standard normal."""
    a = oracle.normal_vector(200000, 0)
    assert (a == oracle.normal_vector(200000, 0)).all()
    assert not (a == oracle.normal_vector(200000, 1)).all()
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1.0) < 0.01


@pytest.mark.parametrize("cells,order,alpha", [((1, 1, 1), 2, 0.0), ((1, 1, 1), 3, 0.0), ((2, 1, 1), 2, 0.0),
                                                ((1, 1, 1), 2, 1.0), ((1, 1, 1), 3, 1.0)])
def test_cfl_step_is_the_lserk4_stability_boundary(oracle, cells, order, alpha):
    """The value of cfl_dt, pinned by the mathematics it approximates.
    session.rs:514-569 brackets and bisects the largest z for which 64
    LSERK4 steps of h = z / rho grow the dominant mode by at most 1e-3 per
    step, and returns 0.8 z / rho. That z is the boundary of
    |R(w)| <= 1 + 1e-3 for the LSERK4 amplification factor R along the
    direction of the dominant eigenvalue: the imaginary axis for the central
    flux (M~-skew operator), the negative real axis for the upwind one. R is
    taken from the oracle's dense LSERK4 on the 2x2 rotation generator (or
    the 1x1 decay), the spectral radius from a dense eigensolve, and the
    bisection leaves z at most its resolution (12/128) below the boundary.
    A wrong safety factor, bracket or growth test moves h off that window."""
    xyz, tets = oracle.box_mesh(*cells)
    op = oracle.Operator(xyz, tets, order, alpha)
    lam = np.linalg.eigvals(dense_a(op))
    rho = np.abs(lam).max()
    dom = lam[np.argmax(np.abs(lam))]
    if alpha == 0.0:
        assert abs(dom.real) < 1e-8 * rho  # central flux: imaginary spectrum
        gen = lambda y: np.array([[0.0, -y], [y, 0.0]])  # noqa: E731
        start = np.array([1.0, 0.0])
    else:
        assert abs(dom.imag) < 1e-8 * rho and dom.real < 0  # upwind: the top mode decays
        gen = lambda y: np.array([[-y]])  # noqa: E731
        start = np.array([1.0])
    lo, hi = 0.5, 8.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        grow = np.linalg.norm(oracle.lserk_dense(gen(mid), start, 1.0, 1))
        lo, hi = (mid, hi) if grow <= 1.0 + 1e-3 else (lo, mid)
    z = op.cfl_dt() * rho / 0.8
    assert lo - 12.0 / 128 <= z <= lo + 1e-3, (z, lo)


def test_cfl_step_is_stable(oracle):
    """session.rs:1352-1364: the CFL step is positive, finite and stable over
    500 steps (norm growth <= 1%); four times the step is not."""
    xyz, tets = oracle.box_mesh(2, 2, 2)
    op = oracle.Operator(xyz, tets, 2)
    h = op.cfl_dt()
    assert 0 < h < 1 and math.isfinite(h)
    y = oracle.normal_vector(op.n_dof, 1)
    out = op.lserk(y, h, 500)
    assert np.linalg.norm(out) <= 1.01 * np.linalg.norm(y)
    # spectral bound: the LSERK4 stability interval on the imaginary axis
    # is ~2.8/rho; 0.8*z/rho with z <= 15 keeps h below the full real range
    rho = np.abs(np.linalg.eigvals(dense_a(op))).max()
    assert h * rho < 0.8 * 15.0 + 1e-9
    blown = op.lserk(y, 4.0 * h, 200)
    assert not np.isfinite(blown).all() or np.linalg.norm(blown) > 10 * np.linalg.norm(y)


# ------------------------------------------------------------ a13-a15 Krylov

def test_expm_matches_reference_golden(oracle):
    """propagator_expm_test.rs: 6x6 wave operator, sympy 30-digit values of
    exp(hA) v0 at h = 0.1, 0.3, 1.0 through the Krylov path with m = 12
    (1e-10)."""
    g = json.load(open(os.path.join(GOLDEN, "propagator_expm_golden.json")))
    a = np.array(g["a"]).reshape(g["n"], g["n"])
    v0 = np.array(g["v0"])
    for case in g["cases"]:
        want = np.array(case["ref"])
        got, _ = oracle.expmv_dense(a, v0, case["h"], g["m"])
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < g["tol"]
        full = oracle.expm(case["h"] * a) @ v0
        assert np.linalg.norm(full - want) / np.linalg.norm(want) < g["tol"]


def test_expm_closed_forms(oracle):
    """propagator.rs:417-444: diagonal, nilpotent and rotation generators,
    plus scipy's expm on a random matrix with a large norm (squaring)."""
    d = np.diag([0.5, -1.0, 2.0])
    assert np.abs(oracle.expm(d) - np.diag(np.exp([0.5, -1.0, 2.0]))).max() < 1e-14 * math.e ** 2
    nil = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]])
    assert np.abs(oracle.expm(nil) - np.array([[1, 1, 0.5], [0, 1, 1], [0, 0, 1]])).max() < 1e-15
    th = 3.7
    rot = oracle.expm(np.array([[0.0, -th], [th, 0.0]]))
    assert np.abs(rot - np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])).max() < 1e-13
    rng = np.random.default_rng(1)
    m = rng.standard_normal((9, 9)) * 3
    want = scipy_expm(m)
    assert np.abs(oracle.expm(m) - want).max() < 1e-10 * np.abs(want).max()


def test_expmv_tolerance_stop_and_breakdown(oracle):
    """propagator.rs:186-326: with tol > 0 the a-posteriori stop ends early
    on a smooth start; an invariant subspace ends the Arnoldi loop at its
    dimension (happy breakdown)."""
    a = wave_6x6()
    v = (np.arange(6) + 1.0) / 7.0
    _, dim_fixed = oracle.expmv_dense(a, v, 0.1, 6)
    out, dim = oracle.expmv_dense(a, v, 0.1, 6, tol=1e-2)
    assert dim_fixed == 6 and dim < 6
    assert np.linalg.norm(out - scipy_expm(0.1 * a) @ v) < 1e-2 * np.linalg.norm(v)
    # e_0 of a diagonal generator spans a 1-D invariant subspace
    diag = np.diag([1.0, 2.0, 3.0, 4.0])
    out, dim = oracle.expmv_dense(diag, np.array([1.0, 0, 0, 0]), 0.5, 4)
    assert dim == 1 and abs(out[0] - math.exp(0.5)) < 1e-14


def test_etd_step_analytic(oracle):
    """propagator.rs:536-559: diagonal A, y(h) = e^{hA} y + (e^{hA} - 1)/A b."""
    lam = np.array([-1.0, -0.5, 0.3, -2.0])
    a = np.diag(lam)
    y = np.array([1.0, 2.0, -1.0, 0.5])
    b = np.array([0.3, -0.2, 1.0, 0.0])
    h = 0.7
    want = np.exp(h * lam) * y + (np.exp(h * lam) - 1.0) / lam * b
    got = oracle.etd_dense(a, y, b, h, 5)
    assert np.abs(got - want).max() < 1e-13


def test_expmv_on_mesh_matches_dense(oracle):
    """The matrix-free Krylov path on a mesh operator against scipy expm."""
    xyz, tets = oracle.box_mesh(1, 1, 1, amplitude=0.0)
    op = oracle.Operator(xyz, tets, 1)
    a = dense_a(op)
    v = oracle.normal_vector(op.n_dof, 2)
    got, _ = op.expmv(v, 0.05, 30)
    assert np.linalg.norm(got - scipy_expm(0.05 * a) @ v) < 1e-10 * np.linalg.norm(v)
    b = oracle.normal_vector(op.n_dof, 3)
    h = 0.05
    aug = np.zeros((op.n_dof + 1, op.n_dof + 1))
    aug[:-1, :-1] = a
    aug[:-1, -1] = b
    want = (scipy_expm(h * aug) @ np.concatenate([v, [1.0]]))[:-1]
    got = op.etd_step(v, b, h, 30)
    assert np.linalg.norm(got - want) < 1e-10 * np.linalg.norm(want)


# ---- Debye dispersive block (rhs.rs:469-500, 963-985, 1295-1324, 1456-1472;
# dispersive.rs:50-53) -------------------------------------------------------

def _debye_op(oracle, order, alpha, every=1, eps_inf=2.0, eps_s=6.0, tau=0.3, cells=(1, 1, 1)):
    xyz, tets = oracle.box_mesh(*cells)
    K = tets.shape[0]
    elems = np.arange(0, K, every)
    op = oracle.Operator(xyz, tets, order, alpha, eps=eps_inf,
                         dispersive=(elems, eps_inf, eps_s, tau))
    return op, elems


def _dense(op):
    n = op.n_dof
    a = np.zeros((n, n))
    for j in range(n):
        ej = np.zeros(n)
        ej[j] = 1.0
        a[:, j] = op.apply(ej)
    return a


def test_debye_dof_count_and_validation(oracle):
    """rhs.rs:3077-3121: no Debye material leaves n_dof = 6 Np K; a Debye
    material on every element appends exactly 3 Np per element; the apply
    stays finite. Out-of-range or repeated elements are refused
    (rhs.rs:970-974)."""
    xyz, tets = oracle.box_mesh(2, 2, 2)
    K = tets.shape[0]
    plain = oracle.Operator(xyz, tets, 2, 1.0)
    assert plain.n_disp == 0 and plain.n_dof == 6 * 10 * K
    op, _ = _debye_op(oracle, 2, 1.0, eps_inf=2.0, eps_s=5.0, tau=0.4)
    assert op.n_disp == 6 and op.n_dof == 6 * 10 * 6 + 3 * 10 * 6
    y = np.sin(0.3 + 0.017 * np.arange(op.n_dof))
    assert np.isfinite(op.apply(y)).all()
    xyz1, tets1 = oracle.box_mesh(1, 1, 1)
    with pytest.raises(ValueError):
        oracle.Operator(xyz1, tets1, 1, 1.0, dispersive=([0, 6], 2.0, 5.0, 0.4))
    with pytest.raises(ValueError):
        oracle.Operator(xyz1, tets1, 1, 1.0, dispersive=([1, 1], 2.0, 5.0, 0.4))


def test_debye_polarisation_rows_and_analytic_permittivity(oracle):
    """rhs.rs:3164-3250: every appended P row holds exactly a = -1/tau on its
    diagonal and g = (eps_s - eps_inf)/tau on the E DOF of the same element,
    node and component, and eps(w) = eps_inf + g/(iw - a) equals the Debye
    permittivity eps_inf + (eps_s - eps_inf)/(1 + i w tau) (dispersive.rs:41-47)."""
    eps_inf, eps_s, tau = 2.0, 6.0, 0.3
    op, elems = _debye_op(oracle, 2, 0.0, eps_inf=eps_inf, eps_s=eps_s, tau=tau)
    a = _dense(op)
    np_, eh = 10, 6 * 10 * op.K
    want_a, want_g = -1.0 / tau, (eps_s - eps_inf) / tau
    for slot, e in enumerate(elems):
        for node in range(np_):
            for c in range(3):
                row = eh + (slot * np_ + node) * 3 + c
                nz = np.flatnonzero(a[row])
                ecol = (e * np_ + node) * 6 + c
                assert sorted(nz) == sorted([row, ecol])
                assert abs(a[row, row] - want_a) < 1e-12
                assert abs(a[row, ecol] - want_g) < 1e-12
    for w in (0.2, 1.0, 3.3, 12.0, 20.0):
        op_eps = eps_inf + want_g / (1j * w - want_a)
        ref = eps_inf + (eps_s - eps_inf) / (1.0 + 1j * w * tau)
        assert abs(op_eps - ref) < 1e-12 * abs(ref)


def test_debye_e_rows_carry_the_polarisation_current(oracle):
    """rhs.rs:1456-1472: on a Debye element the E rows differ from the
    non-dispersive operator (same eps = eps_inf) only by -Pdot/eps_inf: the
    own-E diagonal gains -g/eps_inf and the matching P column holds
    -a/eps_inf; rows of non-dispersive elements are unchanged."""
    eps_inf, eps_s, tau = 2.5, 7.0, 0.3
    xyz, tets = oracle.box_mesh(2, 1, 1)
    K = tets.shape[0]
    elems = np.arange(0, K, 2)
    op = oracle.Operator(xyz, tets, 1, 1.0, eps=eps_inf, dispersive=(elems, eps_inf, eps_s, tau))
    plain = oracle.Operator(xyz, tets, 1, 1.0, eps=eps_inf)
    a, a0 = _dense(op), _dense(plain)
    np_, eh = 4, 6 * 4 * K
    g, aa = (eps_s - eps_inf) / tau, -1.0 / tau
    d = a[:eh, :eh] - a0
    slot_of = {int(e): s for s, e in enumerate(elems)}
    for e in range(K):
        for node in range(np_):
            for c in range(6):
                row = (e * np_ + node) * 6 + c
                if c < 3 and e in slot_of:
                    want = np.zeros(eh)
                    want[row] = -g / eps_inf
                    assert np.abs(d[row] - want).max() < 1e-12
                    pcol = eh + (slot_of[e] * np_ + node) * 3 + c
                    prow = a[row, eh:].copy()
                    assert abs(prow[pcol - eh] + aa / eps_inf) < 1e-12
                    prow[pcol - eh] = 0.0
                    assert np.abs(prow).max() == 0.0
                else:
                    assert np.abs(d[row]).max() < 1e-12
                    assert np.abs(a[row, eh:]).max() == 0.0


def test_debye_energy_dissipation_identity(oracle):
    """With the central flux and PEC walls the curl part is skew in the
    energy inner product, so for the Debye-augmented system the energy
    0.5 eps_inf |E|_M^2 + 0.5 mu |H|_M^2 + 0.5 |P|_M^2 / (eps_s - eps_inf)
    changes at exactly -(eps_s - eps_inf)/tau * |E - P/(eps_s - eps_inf)|_M^2
    summed over the Debye elements (from tau Pdot + P = (eps_s - eps_inf) E and
    eps_inf Edot = curl H - Pdot). A wrong sign or a missing 1/eps_inf on
    either coupling breaks it."""
    eps_inf, eps_s, tau = 2.0, 5.0, 0.4
    de = eps_s - eps_inf
    xyz, tets = oracle.box_mesh(2, 1, 1, 1.0, 1.0, 1.0, 0.1, 3, True)
    K = tets.shape[0]
    elems = np.arange(1, K, 3)
    op = oracle.Operator(xyz, tets, 2, 0.0, eps=eps_inf, dispersive=(elems, eps_inf, eps_s, tau))
    ex = op.export()
    r = oracle.refelem(2)
    np_, eh = r["np"], 6 * r["np"] * K
    w = energy_mass(oracle, op)
    for s, e in enumerate(elems):
        ix = eh + (s * np_ + np.arange(np_)) * 3
        for c in range(3):
            w[np.ix_(ix + c, ix + c)] = abs(ex["det"][e]) * r["mass"] / de
    rng = np.random.default_rng(7)
    for _ in range(3):
        y = rng.standard_normal(op.n_dof)
        rate = y @ w @ op.apply(y)
        want = 0.0
        for s, e in enumerate(elems):
            m = abs(ex["det"][e]) * r["mass"]
            for c in range(3):
                ev = y[(e * np_ + np.arange(np_)) * 6 + c]
                pv = y[eh + (s * np_ + np.arange(np_)) * 3 + c]
                d = ev - pv / de
                want -= de / tau * d @ m @ d
        assert abs(rate - want) < 1e-9 * abs(want)
        assert want < 0.0


def test_port_source_is_the_incident_field_flux(oracle):
    """rhs.rs:1150-1199: the port flux uses the ghost state (E_inc, H_inc);
    its own-trace part is the absorbing operator A and the incident part is
    port_source. So for any state y the operator with the incident field as
    the port faces' ghost equals A y + b. The ghost apply runs the C flux
    code, independent of the numpy port_source restatement; a sign, cross
    product or scaling slip in either breaks it. Materials differ per element
    so the 1/eps, 1/mu scaling is exercised."""
    xyz, tets = oracle.box_mesh(2, 2, 2, 1.0, 1.0, 1.0, 0.1, 4, True)
    K = tets.shape[0]
    order = 2
    nfp = 6
    port = np.full(4 * K, -1, np.int32)
    # every boundary face on z = 0 is a port face
    tmp = oracle.Operator(xyz, tets, order, 1.0)
    ex0 = tmp.export()
    faces = [(e, f) for e in range(K) for f in range(4)
             if ex0["nbr"][e, f] < 0 and ex0["normal"][e, f, 2] < -0.99]
    assert len(faces) == 8
    for e, f in faces:
        port[4 * e + f] = 0
    eps = 1.0 + 0.5 * (np.arange(K) % 3)[:, None] * np.ones((1, 3))
    mu = 1.0 + 0.25 * (np.arange(K) % 2)[:, None] * np.ones((1, 3))
    op = oracle.Operator(xyz, tets, order, 0.7, eps=eps, mu=mu, face_port=port)
    rng = np.random.default_rng(3)
    prof = rng.standard_normal((len(faces), nfp, 6))
    ghost = np.zeros((K, 4, nfp, 6))
    for q, (e, f) in enumerate(faces):
        ghost[e, f] = prof[q]
    # a ghost on non-port faces is ignored
    ghost[0, 0] += 5.0 if port[0] < 0 else 0.0
    b = oracle.port_source(op, faces, prof)
    assert np.abs(b).max() > 0.1
    for y in (np.zeros(op.n_dof), rng.standard_normal(op.n_dof)):
        lhs = op.apply_ghost(y, ghost) - op.apply(y)
        assert np.abs(lhs - b).max() < 1e-12 * np.abs(b).max()
    # linear in the profile; rows outside the port elements are zero
    assert np.allclose(oracle.port_source(op, faces, 2.0 * prof), 2.0 * b, rtol=0, atol=1e-13)
    touched = {e for e, _ in faces}
    for e in range(K):
        if e not in touched:
            assert not np.any(b[e * 60:(e + 1) * 60])


# ---- sparse state-space export (rhs.rs:1575-1730) ---------------------------

def _csr_dense(rp, ci, va, n):
    a = np.zeros((n, n))
    for r in range(n):
        for q in range(rp[r], rp[r + 1]):
            a[r, ci[q]] += va[q]
    return a


def _sparse_case_ops(oracle):
    """Operators covering every flux branch: PEC boundary, port faces,
    jittered interior faces, per-element anisotropic materials with both
    conductivities, central and partly upwind flux, Debye elements."""
    xyz, tets = oracle.box_mesh(2, 1, 1, 2.0, 1.0, 1.0, 0.1, 7, True)
    K = tets.shape[0]
    tmp = oracle.Operator(xyz, tets, 1, 1.0).export()
    port = np.full(4 * K, -1, np.int32)
    for e in range(K):
        for f in range(4):
            if tmp["nbr"][e, f] < 0 and tmp["normal"][e, f, 0] < -0.99:
                port[4 * e + f] = 0
    assert (port >= 0).sum() == 2
    rng = np.random.default_rng(11)
    eps = 1.0 + rng.random((K, 3))
    mu = 1.0 + rng.random((K, 3))
    sig = np.where(np.arange(K) % 3 == 0, 0.4, 0.0)
    sigm = np.where(np.arange(K) % 4 == 1, 0.3, 0.0)
    out = []
    for order, alpha in ((1, 0.0), (2, 0.6)):
        out.append(oracle.Operator(xyz, tets, order, alpha, eps=eps, mu=mu, sigma=sig,
                                   sigma_m=sigm, face_port=port))
    xyz1, tets1 = oracle.box_mesh(1, 1, 1)
    out.append(oracle.Operator(xyz1, tets1, 2, 1.0, eps=2.0,
                               dispersive=(np.arange(0, 6, 2), 2.0, 6.0, 0.3)))
    return out


def test_sparse_assembly_is_the_dense_operator_exactly(oracle):
    """Brute force: the CSR expanded is the dense A of unit-vector applies
    bit for bit, its pattern is exactly the nonzeros of A (probe_column keeps
    v != 0.0), rows are sorted by column, and every P row has exactly its two
    entries (rhs.rs:1704-1727). A wrong stencil, a dropped P column, an
    unsorted row or a lost neighbour block fails one of these."""
    for op in _sparse_case_ops(oracle):
        n = op.n_dof
        rp, ci, va = op.assemble_sparse()
        assert rp[0] == 0 and rp[-1] == len(ci) == len(va) and np.all(np.diff(rp) >= 0)
        for r in range(n):
            row = ci[rp[r]:rp[r + 1]]
            assert np.all(np.diff(row) > 0), f"row {r} not strictly ascending"
        a = dense_a(op)
        got = _csr_dense(rp, ci, va, n)
        assert np.array_equal(got, a)
        eh = n - 3 * op.n_disp * orc_np_of(op)
        mask = np.zeros((n, n), bool)
        for r in range(n):
            mask[r, ci[rp[r]:rp[r + 1]]] = True
        assert np.array_equal(mask[:eh], a[:eh] != 0.0)
        assert np.all(np.diff(rp)[eh:] == 2)


def orc_np_of(op):
    from oracle import oracle as O
    return O.lib().orc_np(op.order)


def test_sparse_assembly_matches_matrix_free_apply(oracle):
    """rhs.rs:2259-2286: structured_box(2,2,2), N=2, upwind: CSR x equals the
    matrix-free apply to 1e-10 and the operator is genuinely sparse."""
    xyz, tets = oracle.box_mesh(2, 2, 2)
    op = oracle.Operator(xyz, tets, 2, 1.0)
    n = op.n_dof
    rp, ci, va = op.assemble_sparse()
    v = np.cos(1.0 + np.arange(n) * 0.07)
    import scipy.sparse as sp
    a = sp.csr_matrix((va, ci, rp), shape=(n, n))
    mf = op.apply(v)
    assert np.linalg.norm(a @ v - mf) < 1e-10 * np.linalg.norm(mf)
    assert len(va) < n * n // 4


def test_sparse_assembly_scales_without_densifying(oracle):
    """rhs.rs:2288-2330: structured_box(4,4,4), N=2: nnz within the
    5 * stride * n stencil bound and CSR x equals the apply."""
    xyz, tets = oracle.box_mesh(4, 4, 4)
    op = oracle.Operator(xyz, tets, 2, 1.0)
    n = op.n_dof
    rp, ci, va = op.assemble_sparse()
    assert len(va) <= 5 * 60 * n
    v = np.sin(0.3 + np.arange(n) * 0.013)
    import scipy.sparse as sp
    a = sp.csr_matrix((va, ci, rp), shape=(n, n))
    mf = op.apply(v)
    assert np.linalg.norm(a @ v - mf) < 1e-10 * np.linalg.norm(mf)


def test_sparse_assembly_debye_rows_give_the_analytic_permittivity(oracle):
    """rhs.rs:3133-3165 (half the elements dispersive, CSR x = apply) and
    rhs.rs:3167-3255: every P row carries exactly g at an E column and a on
    its diagonal, and eps_inf + g/(i w - a) is the Debye permittivity
    eps_inf + (eps_s - eps_inf)/(1 + i w tau)."""
    import scipy.sparse as sp
    xyz, tets = oracle.box_mesh(2, 2, 2)
    K = tets.shape[0]
    op = oracle.Operator(xyz, tets, 2, 1.0, eps=2.5,
                         dispersive=(np.arange(0, K, 2), 2.5, 7.0, 0.3))
    n = op.n_dof
    rp, ci, va = op.assemble_sparse()
    v = np.cos(1.0 + np.arange(n) * 0.07)
    mf = op.apply(v)
    a = sp.csr_matrix((va, ci, rp), shape=(n, n))
    assert np.linalg.norm(a @ v - mf) < 1e-10 * max(np.linalg.norm(mf), 1.0)

    xyz1, tets1 = oracle.box_mesh(1, 1, 1)
    eps_inf, eps_s, tau = 2.0, 6.0, 0.3
    op1 = oracle.Operator(xyz1, tets1, 2, 0.0, eps=eps_inf,
                          dispersive=(np.arange(6), eps_inf, eps_s, tau))
    rp, ci, va = op1.assemble_sparse()
    eh = 6 * 10 * 6
    assert op1.n_dof == eh + 3 * 10 * 6
    avals, gvals = [], []
    for r in range(eh, op1.n_dof):
        assert rp[r + 1] - rp[r] == 2
        for q in range(rp[r], rp[r + 1]):
            (avals if ci[q] == r else gvals).append(va[q])
            if ci[q] != r:
                assert ci[q] < eh
    want_a, want_g = -1.0 / tau, (eps_s - eps_inf) / tau
    assert np.allclose(avals, want_a, rtol=0, atol=1e-12)
    assert np.allclose(gvals, want_g, rtol=0, atol=1e-12)
    for w in (0.2, 1.0, 3.3, 12.0, 20.0):
        op_eps = eps_inf + want_g / (1j * w - want_a)
        exact = eps_inf + (eps_s - eps_inf) / (1 + 1j * w * tau)
        assert abs(op_eps - exact) < 1e-12 * abs(exact)
