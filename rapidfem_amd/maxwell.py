"""Host-side mirror of rapidfem-td's operator construction.

`ReferenceElement`, `structured_box`/`structured_box_jittered` and
`MaxwellOperator` mirror dg_basis.rs, mesh_gen.rs and rhs.rs of the
reference (crates/rapidfem-td/src/) as far as the GPU backend needs them:
they produce exactly the arrays GpuOperator::new reads
(gpu/operator.rs:216-292). The arithmetic is done by the C++ builder in
librfhip.so (include/rfhost.h); there is no CPU operator apply here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


def n_nodes(p: int) -> int:
    return int(L.lib.rfhost_np(p))


def n_face_nodes(p: int) -> int:
    return int(L.lib.rfhost_nfp(p))


@dataclass
class ReferenceElement:
    """dg_basis.rs:28-57. Row-major matrices as numpy arrays."""

    order: int
    nodes: np.ndarray
    mass: np.ndarray
    diff_r: np.ndarray
    diff_s: np.ndarray
    diff_t: np.ndarray
    face_nodes: np.ndarray
    lift: np.ndarray
    face_node_weights: np.ndarray

    @property
    def n_nodes(self) -> int:
        return self.nodes.shape[0]

    @property
    def n_face_nodes(self) -> int:
        return self.face_nodes.shape[1]

    @classmethod
    def new(cls, p: int) -> "ReferenceElement":
        """ReferenceElement::new(p), dg_basis.rs:61-119."""
        if p < 1:
            raise ValueError("DG reference element needs order >= 1")
        n, nfp = n_nodes(p), n_face_nodes(p)
        nodes, mass = np.zeros((n, 3)), np.zeros((n, n))
        dr, ds, dt = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        fn = np.zeros((4, nfp), np.int32)
        lift, fw = np.zeros((n, 4 * nfp)), np.zeros((4, nfp))
        rc = L.lib.rfhost_refelem(p, nodes, mass, dr, ds, dt, fn, lift, fw)
        if rc != 0:
            raise ValueError(f"reference element build failed ({rc})")
        return cls(p, nodes, mass, dr, ds, dt, fn, lift, fw)


@dataclass
class Mesh:
    """Node coordinates and tets (vertex ids ascending per tet)."""

    nodes: np.ndarray  # (n_nodes, 3) f64
    tets: np.ndarray   # (n_tets, 4) int64

    @property
    def n_tets(self) -> int:
        return self.tets.shape[0]


def _box(nx, ny, nz, lx, ly, lz, amplitude, seed, jittered) -> Mesh:
    if min(nx, ny, nz) < 1:
        raise ValueError("box needs at least one cell per axis")
    xyz = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    tets = np.zeros((6 * nx * ny * nz, 4), np.int64)
    L.lib.rfhost_box_mesh(nx, ny, nz, lx, ly, lz, amplitude, seed, 1 if jittered else 0, xyz, tets)
    return Mesh(xyz, tets)


def structured_box(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0) -> Mesh:
    """mesh_gen.rs:54-77: Kuhn 6-tet split of an nx*ny*nz box."""
    return _box(nx, ny, nz, lx, ly, lz, 0.0, 0, False)


def structured_box_jittered(nx, ny, nz, lx, ly, lz, amplitude, seed) -> Mesh:
    """mesh_gen.rs:86-121: interior nodes displaced by amplitude*h*jitter_offset."""
    return _box(nx, ny, nz, lx, ly, lz, amplitude, seed, True)


@dataclass
class ElemMaterial:
    """rhs.rs:430-467 ElemMaterial (diagonal eps/mu, sigma, sigma*)."""

    eps: tuple = (1.0, 1.0, 1.0)
    mu: tuple = (1.0, 1.0, 1.0)
    sigma: float = 0.0
    sigma_m: float = 0.0

    @classmethod
    def isotropic(cls, eps, mu=1.0, sigma=0.0):
        return cls((eps,) * 3, (mu,) * 3, sigma, 0.0)

    @classmethod
    def matched_absorber(cls, eps, mu, nu):
        return cls((eps,) * 3, (mu,) * 3, nu * eps, nu * mu)


def flux_alpha(flux: str) -> float:
    """session.rs:97-103."""
    if flux == "upwind":
        return 1.0
    if flux == "central":
        return 0.0
    raise ValueError(f"flux must be one of ['central', 'upwind'], got {flux!r}")


class DebyeMaterial:
    """dispersive.rs:28-54 DebyeMaterial: eps(w) = eps_inf + (eps_s - eps_inf)
    / (1 + i w tau); the polarisation obeys tau Pdot + P = (eps_s - eps_inf) E."""

    def __init__(self, eps_inf: float, eps_static: float, tau: float):
        self.eps_inf = float(eps_inf)
        self.eps_static = float(eps_static)
        self.tau = float(tau)

    def permittivity(self, omega: float) -> tuple[float, float]:
        """dispersive.rs:41-47: (Re, Im) of the relative permittivity at omega."""
        d = self.eps_static - self.eps_inf
        den = 1.0 + (omega * self.tau) ** 2
        return self.eps_inf + d / den, -d * omega * self.tau / den

    def relaxation_coeffs(self) -> tuple[float, float]:
        """dispersive.rs:50-53: (a, g) of Pdot = a P + g E."""
        return -1.0 / self.tau, (self.eps_static - self.eps_inf) / self.tau


class MaxwellOperator:
    """rhs.rs:502-527 MaxwellOperator, the data the GPU backend consumes.

    materials: None (vacuum), a list of ElemMaterial (one per tet), or a dict
    of arrays {'eps': (K,3), 'mu': (K,3), 'sigma': (K,), 'sigma_m': (K,)}.
    face_port: optional (K, 4) int array, port id per boundary face or -1
    (a port tag on an interior face is ignored, as rhs.rs:867-872 drops it).
    dispersive: optional list of (element, DebyeMaterial) (rhs.rs:782-788);
    each listed element gets a P-block slot in list order and its material
    eps should be eps_inf (rhs.rs:469-487, 963-985).
    """

    def __init__(self, mesh: Mesh, order: int, flux_alpha: float = 1.0, materials=None, face_port=None,
                 dispersive=None):
        self.mesh = mesh
        self.re = ReferenceElement.new(order)
        self.order = order
        self.flux_alpha = float(flux_alpha)
        K = mesh.n_tets
        self.n_elem = K
        np_, nfp = self.re.n_nodes, self.re.n_face_nodes
        eps = mu = sigma = sigma_m = None
        if materials is not None:
            if isinstance(materials, dict):
                eps = materials.get("eps")
                mu = materials.get("mu")
                sigma = materials.get("sigma")
                sigma_m = materials.get("sigma_m")
            else:
                if len(materials) != K:
                    raise ValueError("one material per element")
                eps = np.array([m.eps for m in materials], np.float64)
                mu = np.array([m.mu for m in materials], np.float64)
                sigma = np.array([m.sigma for m in materials], np.float64)
                sigma_m = np.array([m.sigma_m for m in materials], np.float64)
        keep = []

        def opt(a, shape, dt=np.float64):
            if a is None:
                return None
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt), shape))
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        self.jinv = np.zeros((K, 3, 3))
        self.det = np.zeros(K)
        self.normal = np.zeros((K, 4, 3))
        self.fscale = np.zeros((K, 4))
        self.nbr = np.zeros((K, 4), np.int32)
        self.nbr_lf = np.zeros((K, 4), np.int32)
        self.perm = np.zeros((K, 4, nfp), np.int32)
        self.port = np.zeros((K, 4), np.int32)
        self.inv_eps = np.zeros((K, 3))
        self.inv_mu = np.zeros((K, 3))
        self.sigma_eps = np.zeros((K, 3))
        self.sigma_mu = np.zeros((K, 3))
        nodes = np.ascontiguousarray(mesh.nodes, np.float64)
        tets = np.ascontiguousarray(mesh.tets, np.int64)
        rc = L.lib.rfhost_operator_arrays(
            nodes.shape[0], nodes, K, tets, order, opt(eps, (K, 3)), opt(mu, (K, 3)), opt(sigma, (K,)),
            opt(sigma_m, (K,)), opt(face_port, (K, 4), np.int32), self.jinv, self.det, self.normal,
            self.fscale, self.nbr, self.nbr_lf, self.perm, self.port, self.inv_eps, self.inv_mu,
            self.sigma_eps, self.sigma_mu)
        if rc == 2:
            raise RuntimeError("unmatched face node: the mesh is not conforming")
        if rc == 3:
            raise RuntimeError("non-manifold mesh: a face is shared by more than two tets")
        if rc != 0:
            raise RuntimeError(f"operator build failed ({rc})")
        # rhs.rs:963-985 Debye slots: element, a, g, 1/eps_inf per slot
        disp = list(dispersive or [])
        self.disp_elem = np.array([int(e) for e, _ in disp], np.int64)
        if len(set(self.disp_elem.tolist())) != len(disp) or np.any(self.disp_elem < 0) or np.any(
                self.disp_elem >= K):
            raise ValueError("dispersive element out of range or listed twice")
        self.disp_a = np.array([m.relaxation_coeffs()[0] for _, m in disp], np.float64)
        self.disp_g = np.array([m.relaxation_coeffs()[1] for _, m in disp], np.float64)
        self.disp_inv_eps_inf = np.array([1.0 / m.eps_inf for _, m in disp], np.float64)
        self.n_dispersive = len(disp)
        self.n_dof = 6 * np_ * K + 3 * np_ * self.n_dispersive

    def node_coords(self) -> np.ndarray:
        """(K, Np, 3) physical coordinates of every DG node."""
        out = np.zeros((self.n_elem, self.re.n_nodes, 3))
        L.lib.rfhost_node_coords(self.n_elem, np.ascontiguousarray(self.mesh.nodes, np.float64),
                                 np.ascontiguousarray(self.mesh.tets, np.int64), self.order, out)
        return out

    def desc(self) -> tuple[L.OpDesc, list]:
        """The rfhip_op_desc of this operator, plus the arrays it points to."""
        re = self.re
        arrs = dict(
            diff_r=np.ascontiguousarray(re.diff_r), diff_s=np.ascontiguousarray(re.diff_s),
            diff_t=np.ascontiguousarray(re.diff_t), lift=np.ascontiguousarray(re.lift),
            face_nodes=np.ascontiguousarray(re.face_nodes, np.int32), jinv=self.jinv,
            inv_eps=self.inv_eps, inv_mu=self.inv_mu, sigma_eps=self.sigma_eps, sigma_mu=self.sigma_mu,
            face_normal=self.normal, face_fscale=self.fscale, face_neighbor=self.nbr,
            face_nbr_local=self.nbr_lf, face_port=self.port, face_perm=self.perm)
        d = L.OpDesc()
        d.np = re.n_nodes
        d.nfp = re.n_face_nodes
        d.n_elem = self.n_elem
        for k, v in arrs.items():
            setattr(d, k, v.ctypes.data_as(C.c_void_p))
        d.flux_alpha = self.flux_alpha
        return d, list(arrs.values())


def port_face_data(op: MaxwellOperator, faces, profile):
    """Arrays of a port for GpuOperator.set_port, as MaxwellOperator's
    PortData holds them (rhs.rs:319-329): `faces` is a list of (element,
    local face) boundary faces, `profile(xyz) -> (e_t, h_t)` the mode's
    transverse profiles at points xyz (n, 3). Weights are 2*area times the
    face-node weight (rhs.rs:1224-1228), area = fscale*|det J|/2."""
    faces = np.asarray(faces, np.int64).reshape(-1, 2)
    re = op.re
    nfp = re.n_face_nodes
    x = op.node_coords()
    elem = faces[:, 0].astype(np.int32)
    lf = faces[:, 1].astype(np.int32)
    pts = x[faces[:, 0][:, None], re.face_nodes[faces[:, 1]]]  # (n_faces, Nfp, 3)
    et, ht = profile(pts.reshape(-1, 3))
    prof = np.concatenate([np.asarray(et, np.float64).reshape(-1, 3), np.asarray(ht, np.float64).reshape(-1, 3)],
                          axis=1).reshape(len(faces), nfp, 6)
    area = op.fscale[faces[:, 0], faces[:, 1]] * np.abs(op.det[faces[:, 0]]) / 2.0
    w = 2.0 * area[:, None] * re.face_node_weights[faces[:, 1]]
    return elem, lf, prof, w


def boundary_faces_on_plane(op: MaxwellOperator, axis: int, value: float, tol: float = 1e-9):
    """(element, local face) of every boundary face lying in the plane
    x[axis] = value."""
    x = op.node_coords()
    re = op.re
    out = []
    for e in np.nonzero((op.nbr < 0).any(axis=1))[0]:
        for f in range(4):
            if op.nbr[e, f] >= 0:
                continue
            if np.all(np.abs(x[e, re.face_nodes[f], axis] - value) < tol):
                out.append((int(e), f))
    return out
