"""ctypes wrapper of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package. The oracle is a
plain-C restatement of rapidfem-td's CPU path (crates/rapidfem-td/src/*.rs
in the reference); see dgtd_oracle.c for the per-function citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load() -> C.CDLL:
    if not os.path.exists(_SO):
        build()
    lib = C.CDLL(_SO)
    vp = C.c_void_p
    i64 = C.c_int64
    d = C.c_double
    sig = {
        "orc_np": (C.c_int, [C.c_int]),
        "orc_nfp": (C.c_int, [C.c_int]),
        "orc_refelem": (C.c_int, [C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp]),
        "orc_box_nodes": (None, [C.c_int, C.c_int, C.c_int, d, d, d, d, C.c_uint64, C.c_int, _dp]),
        "orc_kuhn_tets": (None, [C.c_int, C.c_int, C.c_int, _lp]),
        "orc_jitter_offset": (None, [i64, i64, i64, C.c_uint64, _dp]),
        "orc_element_curl": (None, [C.c_int, _dp, _dp, _dp]),
        "orc_geom_factors": (None, [_dp, _dp, _dp, _dp]),
        "orc_op_new": (vp, [i64, _dp, i64, _lp, C.c_int, d, vp, vp, vp, vp, vp]),
        "orc_op_free": (None, [vp]),
        "orc_op_n_dof": (i64, [vp]),
        "orc_op_n_elem": (i64, [vp]),
        "orc_op_order": (C.c_int, [vp]),
        "orc_op_export": (None, [vp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp]),
        "orc_op_set_dispersive": (C.c_int, [vp, i64, _lp, _dp, _dp, _dp]),
        "orc_op_n_dispersive": (i64, [vp]),
        "orc_op_set_threads": (None, [C.c_int]),
        "orc_apply": (None, [vp, _dp, _dp]),
        "orc_apply_ghost": (None, [vp, _dp, _dp, _dp]),
        "orc_assemble_sparse": (i64, [vp, _lp, vp, vp]),
        "orc_field_energy": (d, [vp, _dp]),
        "orc_lserk_run": (None, [vp, _dp, d, i64, C.c_int, i64, vp, vp]),
        "orc_lserk_run_dense": (None, [_dp, C.c_int, _dp, d, i64]),
        "orc_kcl_step": (None, [vp, _dp, _dp, d, C.c_int, i64, vp, d]),
        "orc_kcl_step_dense": (None, [_dp, C.c_int, _dp, _dp, d]),
        "orc_err_norm": (d, [vp, _dp, _dp, _dp, i64]),
        "orc_factor": (d, [vp, d, d, C.c_int]),
        "orc_adaptive_run": (C.c_int, [vp, _dp, d, i64, vp, C.c_int, i64, vp, vp, _lp, _dp]),
        "orc_controller_default": (None, [vp]),
        "orc_normal_vector": (None, [i64, C.c_uint64, _dp]),
        "orc_cfl_dt": (d, [vp]),
        "orc_expm": (None, [_dp, C.c_int, _dp]),
        "orc_expmv_dense": (C.c_int, [_dp, C.c_int, _dp, d, C.c_int, d, _dp]),
        "orc_expmv_op": (C.c_int, [vp, _dp, d, C.c_int, d, _dp]),
        "orc_etd_dense": (None, [_dp, C.c_int, _dp, _dp, d, C.c_int, d, _dp]),
        "orc_etd_op": (None, [vp, _dp, _dp, d, C.c_int, d, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Controller(C.Structure):
    """session.rs:130-140 Controller (all f64)."""

    _fields_ = [(n, C.c_double) for n in (
        "atol", "rtol", "safety", "growth_limit", "shrink_limit", "pi_alpha",
        "pi_beta", "min_step_factor")]

    @classmethod
    def default(cls) -> "Controller":
        c = cls()
        lib().orc_controller_default(C.byref(c))
        return c


def refelem(p: int) -> dict:
    """dg_basis.rs ReferenceElement::new(p) as numpy arrays."""
    L = lib()
    n, nfp = L.orc_np(p), L.orc_nfp(p)
    out = dict(
        nodes=np.zeros((n, 3)), mass=np.zeros((n, n)), dr=np.zeros((n, n)),
        ds=np.zeros((n, n)), dt=np.zeros((n, n)),
        face_nodes=np.zeros((4, nfp), np.int32), lift=np.zeros((n, 4 * nfp)),
        fweights=np.zeros((4, nfp)))
    rc = L.orc_refelem(p, out["nodes"], out["mass"], out["dr"], out["ds"], out["dt"],
                       out["face_nodes"], out["lift"], out["fweights"])
    assert rc == 0
    out.update(np=n, nfp=nfp, p=p)
    return out


def box_mesh(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, amplitude=0.0, seed=0, jittered=False):
    """mesh_gen.rs structured_box / structured_box_jittered: (xyz, tets)."""
    L = lib()
    xyz = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    L.orc_box_nodes(nx, ny, nz, lx, ly, lz, amplitude, seed, 1 if jittered else 0, xyz)
    tets = np.zeros((nx * ny * nz * 6, 4), np.int64)
    L.orc_kuhn_tets(nx, ny, nz, tets)
    return xyz, tets


def element_curl(p: int, tet: np.ndarray, field: np.ndarray) -> np.ndarray:
    out = np.zeros_like(np.ascontiguousarray(field, np.float64))
    lib().orc_element_curl(p, np.ascontiguousarray(tet, np.float64),
                           np.ascontiguousarray(field, np.float64), out)
    return out


class Operator:
    """rhs.rs MaxwellOperator restated in C.

    dispersive: optional (elems, eps_inf, eps_static, tau), one entry per
    Debye element (rhs.rs:782-788 OperatorOptions::dispersive); the state
    then carries the appended polarisation block (rhs.rs:490-500)."""

    def __init__(self, xyz, tets, order, flux_alpha=1.0, eps=None, mu=None,
                 sigma=None, sigma_m=None, face_port=None, dispersive=None):
        L = lib()
        self.xyz = np.ascontiguousarray(xyz, np.float64)
        self.tets = np.ascontiguousarray(tets, np.int64)
        K = self.tets.shape[0]
        keep = []

        def opt(a, dt, shape):
            if a is None:
                return None
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt), shape))
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        h = L.orc_op_new(self.xyz.shape[0], self.xyz, K, self.tets, order, float(flux_alpha),
                         opt(eps, np.float64, (K, 3)), opt(mu, np.float64, (K, 3)),
                         opt(sigma, np.float64, (K,)), opt(sigma_m, np.float64, (K,)),
                         opt(face_port, np.int32, (4 * K,)))
        if not h:
            raise RuntimeError("oracle operator build failed: unmatched face node")
        self._h = C.c_void_p(h)
        self.order = order
        self.K = K
        if dispersive is not None:
            de, einf, es, tau = dispersive
            de = np.ascontiguousarray(de, np.int64)
            n = de.shape[0]
            rc = L.orc_op_set_dispersive(self._h, n, de,
                                         np.ascontiguousarray(np.broadcast_to(einf, (n,)), np.float64),
                                         np.ascontiguousarray(np.broadcast_to(es, (n,)), np.float64),
                                         np.ascontiguousarray(np.broadcast_to(tau, (n,)), np.float64))
            if rc != 0:
                raise ValueError("dispersive element out of range or listed twice")
        self.n_disp = int(L.orc_op_n_dispersive(self._h))
        self.n_dof = int(L.orc_op_n_dof(self._h))
        self.flux_alpha = float(flux_alpha)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_op_free(self._h)
            self._h = None

    def export(self) -> dict:
        L = lib()
        K, nfp = self.K, L.orc_nfp(self.order)
        o = dict(jinv=np.zeros((K, 3, 3)), det=np.zeros(K), normal=np.zeros((K, 4, 3)),
                 fscale=np.zeros((K, 4)), nbr=np.zeros((K, 4), np.int32),
                 nbr_lf=np.zeros((K, 4), np.int32), port=np.zeros((K, 4), np.int32),
                 perm=np.zeros((K, 4, nfp), np.int32), inv_eps=np.zeros((K, 3)),
                 inv_mu=np.zeros((K, 3)), sig_eps=np.zeros((K, 3)), sig_mu=np.zeros((K, 3)))
        L.orc_op_export(self._h, o["jinv"], o["det"], o["normal"], o["fscale"], o["nbr"],
                        o["nbr_lf"], o["port"], o["perm"], o["inv_eps"], o["inv_mu"],
                        o["sig_eps"], o["sig_mu"])
        return o

    def apply(self, y: np.ndarray) -> np.ndarray:
        y = np.ascontiguousarray(y, np.float64)
        assert y.shape == (self.n_dof,)
        dy = np.empty_like(y)
        lib().orc_apply(self._h, y, dy)
        return dy

    def apply_ghost(self, y: np.ndarray, ghost: np.ndarray) -> np.ndarray:
        """A y with the incident field `ghost` ((K, 4, Nfp, 6): E_inc, H_inc at
        every face node, read on port faces only) as the port faces' ghost
        state; the absorbing apply is the zero-ghost case."""
        y = np.ascontiguousarray(y, np.float64)
        ghost = np.ascontiguousarray(ghost, np.float64)
        assert y.shape == (self.n_dof,) and ghost.size == self.K * 4 * lib().orc_nfp(self.order) * 6
        dy = np.empty_like(y)
        lib().orc_apply_ghost(self._h, y, ghost.reshape(-1), dy)
        return dy

    def assemble_sparse(self):
        """The operator as CSR (row_ptr, col_idx, values), probed column by
        column as MaxwellOperator::assemble_sparse does (rhs.rs:1621-1730)."""
        rp = np.empty(self.n_dof + 1, np.int64)
        nnz = lib().orc_assemble_sparse(self._h, rp, None, None)
        ci = np.empty(nnz, np.int64)
        va = np.empty(nnz, np.float64)
        lib().orc_assemble_sparse(self._h, rp, ci.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p))
        return rp, ci, va

    def field_energy(self, y) -> float:
        return float(lib().orc_field_energy(self._h, np.ascontiguousarray(y, np.float64)))

    def lserk(self, y, dt, nsteps, drive=None, g=None):
        """LSERK4 transient in place. drive: None, ('point', dof), ('vector', b)."""
        y = np.array(y, np.float64, copy=True)
        kind, dof, vec = _drive(drive, self.n_dof)
        gg = None if g is None else np.ascontiguousarray(g, np.float64)
        lib().orc_lserk_run(self._h, y, float(dt), int(nsteps), kind, dof, _ptr(vec), _ptr(gg))
        return y

    def kcl_step(self, y, dt, drive=None, g=0.0):
        y = np.array(y, np.float64, copy=True)
        err = np.zeros_like(y)
        kind, dof, vec = _drive(drive, self.n_dof)
        lib().orc_kcl_step(self._h, y, err, float(dt), kind, dof, _ptr(vec), float(g))
        return y, err

    def adaptive(self, y, dt, steps, ctl=None, drive=None, g=None):
        y = np.array(y, np.float64, copy=True)
        ctl = ctl or Controller.default()
        kind, dof, vec = _drive(drive, self.n_dof)
        gg = None if g is None else np.ascontiguousarray(g, np.float64)
        stats = np.zeros(2, np.int64)
        hmm = np.zeros(2)
        rc = lib().orc_adaptive_run(self._h, y, float(dt), int(steps), C.byref(ctl), kind, dof,
                                    _ptr(vec), _ptr(gg), stats, hmm)
        if rc != 0:
            raise RuntimeError("adaptive stepper: step size collapsed")
        return y, int(stats[0]), int(stats[1]), float(hmm[0]), float(hmm[1])

    def cfl_dt(self) -> float:
        return float(lib().orc_cfl_dt(self._h))

    def expmv(self, v, t, max_dim, tol=0.0):
        v = np.ascontiguousarray(v, np.float64)
        out = np.zeros_like(v)
        dim = lib().orc_expmv_op(self._h, v, float(t), int(max_dim), float(tol), out)
        return out, dim

    def etd_step(self, y, b, h, max_dim, tol=0.0):
        y = np.ascontiguousarray(y, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        out = np.zeros_like(y)
        lib().orc_etd_op(self._h, y, b, float(h), int(max_dim), float(tol), out)
        return out


def _drive(drive, n):
    if drive is None:
        return 0, 0, None
    kind, val = drive
    if kind == "point":
        return 1, int(val), None
    if kind == "vector":
        v = np.ascontiguousarray(val, np.float64)
        assert v.shape == (n,)
        return 2, 0, v
    raise ValueError(kind)


def set_threads(n: int) -> None:
    lib().orc_op_set_threads(int(n))


def expm(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float64)
    out = np.zeros_like(a)
    lib().orc_expm(a, a.shape[0], out)
    return out


def expmv_dense(a, v, t, max_dim, tol=0.0):
    a = np.ascontiguousarray(a, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    out = np.zeros_like(v)
    dim = lib().orc_expmv_dense(a, a.shape[0], v, float(t), int(max_dim), float(tol), out)
    return out, dim


def etd_dense(a, y, b, h, max_dim, tol=0.0):
    a = np.ascontiguousarray(a, np.float64)
    out = np.zeros(a.shape[0])
    lib().orc_etd_dense(a, a.shape[0], np.ascontiguousarray(y, np.float64),
                        np.ascontiguousarray(b, np.float64), float(h), int(max_dim), float(tol), out)
    return out


def lserk_dense(a, y, dt, nsteps):
    a = np.ascontiguousarray(a, np.float64)
    y = np.array(y, np.float64, copy=True)
    lib().orc_lserk_run_dense(a, a.shape[0], y, float(dt), int(nsteps))
    return y


def kcl_step_dense(a, y, dt):
    a = np.ascontiguousarray(a, np.float64)
    y = np.array(y, np.float64, copy=True)
    err = np.zeros_like(y)
    lib().orc_kcl_step_dense(a, a.shape[0], y, err, float(dt))
    return y, err


def normal_vector(n, seed=0):
    out = np.zeros(n)
    lib().orc_normal_vector(int(n), int(seed), out)
    return out


def err_norm(ctl, y_old, y_new, err):
    return float(lib().orc_err_norm(C.byref(ctl), np.ascontiguousarray(y_old, np.float64),
                                    np.ascontiguousarray(y_new, np.float64),
                                    np.ascontiguousarray(err, np.float64), len(err)))


def factor(ctl, e, prev, reject):
    return float(lib().orc_factor(C.byref(ctl), float(e), float(prev), 1 if reject else 0))


def port_source(op: "Operator", faces, profiles) -> np.ndarray:
    """rhs.rs:1157-1199 MaxwellOperator::port_source: the drive pattern b of a
    port's incident mode. faces: (n, 2) (element, local face); profiles:
    (n, Nfp, 6) e_t then h_t at the face nodes (face-node order). With the
    ghost state (E_inc, H_inc) = (e_t, h_t) the port flux jumps are [E] = -e_t,
    [H] = -h_t; their characteristic (alpha = 1) flux, lifted and scaled by
    the element's 1/eps, 1/mu, is b (numpy, for parity tests)."""
    ex = op.export()
    r = refelem(op.order)
    n_p, nfp = r["np"], r["nfp"]
    b = np.zeros(op.n_dof)
    prof = np.asarray(profiles, np.float64).reshape(len(faces), nfp, 6)
    for q, (e, f) in enumerate(faces):
        nv = ex["normal"][e, f]
        coef = 0.5 * ex["fscale"][e, f]
        ie, im = ex["inv_eps"][e], ex["inv_mu"][e]
        base = (e * n_p + np.arange(n_p)) * 6
        for m in range(nfp):
            et, ht = prof[q, m, :3], prof[q, m, 3:]
            nxe, nxh = np.cross(nv, et), np.cross(nv, ht)
            s_de = nxh - np.cross(nv, nxe)
            s_dh = -nxe - np.cross(nv, nxh)
            w = coef * r["lift"][:, f * nfp + m]
            b[base[:, None] + np.arange(3)] += ie * w[:, None] * s_de
            b[base[:, None] + 3 + np.arange(3)] += im * w[:, None] * s_dh
    return b


def port_projections(y, dof, prof, w):
    """rhs.rs:1211-1242 port_modal_projections for one port given as
    per-face-node entries: dof[q] = AoS index of the node's first field
    value, prof[q] = (e_t, h_t), w[q] = 2*area*face-node weight. Returns
    (P_e, P_h)."""
    y = np.asarray(y, np.float64)
    dof = np.asarray(dof, np.int64).reshape(-1)
    prof = np.asarray(prof, np.float64).reshape(-1, 6)
    w = np.asarray(w, np.float64).reshape(-1)
    ef = y[dof[:, None] + np.arange(3)]
    hf = y[dof[:, None] + 3 + np.arange(3)]
    e_dot = float(np.sum(w * np.sum(ef * prof[:, :3], axis=1)))
    h_dot = float(np.sum(w * np.sum(hf * prof[:, 3:], axis=1)))
    e_norm = float(np.sum(w * np.sum(prof[:, :3] ** 2, axis=1)))
    h_norm = float(np.sum(w * np.sum(prof[:, 3:] ** 2, axis=1)))
    return (e_dot / e_norm if e_norm > 0 else 0.0, h_dot / h_norm if h_norm > 0 else 0.0)
