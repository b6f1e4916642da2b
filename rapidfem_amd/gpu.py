"""GpuContext / GpuOperator over the rfhip C ABI.

Mirrors rapidfem_td::gpu::{GpuContext, GpuOperator}
(crates/rapidfem-td/src/gpu/mod.rs:33-135, gpu/operator.rs:205-1486): same
method names and argument meaning; errors raise RfhipError where the
reference returns Err(String), and length mismatches (asserts in the
reference, gpu/operator.rs:518,552) raise as well.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .maxwell import MaxwellOperator


@dataclass
class Controller:
    """session.rs:130-155 Controller with its defaults (constants.rs:268-299)."""

    atol: float = 1e-8
    rtol: float = 1e-4
    safety: float = 0.9
    growth_limit: float = 5.0
    shrink_limit: float = 0.2
    pi_alpha: float = 0.7 / 4.0
    pi_beta: float = 0.4 / 4.0
    min_step_factor: float = 1e-10

    def c(self) -> L.ControllerC:
        return L.ControllerC(self.atol, self.rtol, self.safety, self.growth_limit, self.shrink_limit,
                             self.pi_alpha, self.pi_beta, self.min_step_factor)


@dataclass
class KclStats:
    """session.rs:218-223 KclStats (h in operator time here)."""

    accepted: int
    rejected: int
    h_min: float
    h_max: float


class Drive:
    """session.rs:107-114: Free, Point(dof) or Vector(pattern)."""

    def __init__(self, kind: int, dof: int = 0, vec: np.ndarray | None = None):
        self.kind, self.dof, self.vec = kind, dof, vec

    @classmethod
    def free(cls):
        return cls(L.DRIVE_FREE)

    @classmethod
    def point(cls, dof: int):
        return cls(L.DRIVE_POINT, int(dof))

    @classmethod
    def vector(cls, b):
        return cls(L.DRIVE_VECTOR, 0, np.ascontiguousarray(b, np.float64))

    @classmethod
    def port_source(cls):
        """Vector drive with the port source last assembled on the device
        (GpuOperator.port_source), without a host copy of the pattern."""
        return cls(L.DRIVE_VECTOR, 0, None)


class GpuContext:
    """gpu/mod.rs:33-65: the first HIP device."""

    def __init__(self):
        h = C.c_void_p()
        L.check(L.lib.rfhip_ctx_create(C.byref(h)))
        self._h = h
        buf = C.create_string_buffer(256)
        L.check(L.lib.rfhip_ctx_device_name(self._h, buf, 256))
        self.device_name = buf.value.decode()
        self.fp64 = bool(L.lib.rfhip_ctx_has_fp64(self._h))

    @classmethod
    def new(cls) -> "GpuContext":
        return cls()

    def __del__(self):
        if getattr(self, "_h", None):
            L.lib.rfhip_ctx_destroy(self._h)
            self._h = None


class GpuOperator:
    """gpu/operator.rs:78-153 GpuOperator; form selects the kernel family
    ('plain_f32' = the reference's structure and precision, 'plain_f64',
    'fused_f64', 'fused_f64c', and the precision forms 'fused_bf16x4' /
    'fused_bf16x3' / 'fused_bf16x2': fused_f64c with the volume contraction
    on the matrix cores in 4-, 3- or 2-piece split BF16 with FP32
    accumulation)."""

    def __init__(self, ctx: GpuContext, op: MaxwellOperator | None, form: str = "fused_f64", *, desc=None,
                 mass=None, det=None, dispersive=None):
        self.ctx = ctx  # keep the context alive
        self.form = form
        if op is not None:
            d, keep = op.desc()
            mass, det = op.re.mass, op.det
        else:
            d, keep = desc, None
        h = C.c_void_p()
        if op is not None and getattr(op, "n_dispersive", 0):
            dispersive = (op.disp_elem, op.disp_a, op.disp_g, op.disp_inv_eps_inf)
        if dispersive is not None:
            de, da, dg, di = (np.ascontiguousarray(dispersive[0], np.int64),
                              *(np.ascontiguousarray(np.broadcast_to(v, (len(dispersive[0]),)), np.float64)
                                for v in dispersive[1:]))
            L.check(L.lib.rfhip_op_create_dispersive(ctx._h, C.byref(d), L.FORMS[form], de.shape[0], de, da, dg,
                                                     di, C.byref(h)))
        else:
            L.check(L.lib.rfhip_op_create(ctx._h, C.byref(d), L.FORMS[form], C.byref(h)))
        del keep
        self._h = h
        self._n = int(L.lib.rfhip_op_n_dof(self._h))
        self.n_elem = int(d.n_elem)
        self._mass = None if mass is None else np.ascontiguousarray(mass, np.float64)
        self._det = None if det is None else np.ascontiguousarray(det, np.float64)

    @classmethod
    def from_desc(cls, ctx: GpuContext, desc: L.OpDesc, form: str = "fused_f64", mass=None, det=None,
                  dispersive=None):
        """From raw rfhip_op_desc data (e.g. another host's MaxwellOperator);
        dispersive: optional (elements, a, g, 1/eps_inf) of the Debye slots."""
        return cls(ctx, None, form, desc=desc, mass=mass, det=det, dispersive=dispersive)

    @classmethod
    def new(cls, ctx: GpuContext, op: MaxwellOperator, form: str = "fused_f64") -> "GpuOperator":
        return cls(ctx, op, form)

    def __del__(self):
        if getattr(self, "_h", None):
            L.lib.rfhip_op_destroy(self._h)
            self._h = None

    def n_dof(self) -> int:
        return self._n

    @property
    def n_dispersive(self) -> int:
        """rhs.rs:1021-1024: number of Debye elements (P-block slots)."""
        return int(L.lib.rfhip_op_n_dispersive(self._h))

    def _vec(self, a, dt):
        a = np.ascontiguousarray(a, dt)
        if a.shape != (self._n,):
            raise ValueError(f"state length mismatch: {a.shape} vs ({self._n},)")
        return a

    # gpu/operator.rs:513-523
    def apply(self, ctx: GpuContext, y) -> np.ndarray:
        y = self._vec(y, np.float32)
        dy = np.empty_like(y)
        L.check(L.lib.rfhip_op_apply(self._h, y, dy))
        return dy

    def apply_f64(self, y) -> np.ndarray:
        y = self._vec(y, np.float64)
        dy = np.empty_like(y)
        L.check(L.lib.rfhip_op_apply_f64(self._h, y, dy))
        return dy

    # gpu/operator.rs:906-953
    def transient(self, ctx, y0, dt, steps, substeps=1, drive: Drive | None = None, g=None, record=True,
                  precision="f32") -> np.ndarray:
        drive = drive or Drive.free()
        f64 = precision == "f64"
        dtp = np.float64 if f64 else np.float32
        y0 = self._vec(y0, dtp)
        substeps = max(1, int(substeps))
        gg = None if g is None else np.ascontiguousarray(g, dtp)
        n_g = 0 if gg is None else gg.shape[0]
        out = np.empty(((steps + 1) if record else 1) * self._n, dtp)
        fn = L.lib.rfhip_op_transient_f64 if f64 else L.lib.rfhip_op_transient
        L.check(fn(self._h, y0, dt, int(steps), substeps, drive.kind, drive.dof, L.ptr(drive.vec), L.ptr(gg),
                   n_g, 1 if record else 0, out))
        return out.reshape(steps + 1, self._n) if record else out

    def transient_probes(self, ctx, y0, dt, steps, probes, substeps=1, drive: Drive | None = None, g=None):
        """Explicit transient recording only the DOFs in `probes`
        (session.rs:118-123 Record::Dofs) on the device: returns
        ((steps+1, len(probes)) rows, final state), both f64."""
        drive = drive or Drive.free()
        y0 = self._vec(y0, np.float64)
        substeps = max(1, int(substeps))
        gg = None if g is None else np.ascontiguousarray(g, np.float64)
        n_g = 0 if gg is None else gg.shape[0]
        pr = np.ascontiguousarray(probes, np.int64)
        rows = np.empty((steps + 1) * max(1, pr.shape[0]))
        yf = np.empty(self._n)
        L.check(L.lib.rfhip_op_transient_probes_f64(self._h, y0, float(dt), int(steps), substeps, drive.kind,
                                                    drive.dof, L.ptr(drive.vec), L.ptr(gg), n_g, pr, pr.shape[0],
                                                    rows, L.ptr(yf)))
        return rows[:(steps + 1) * pr.shape[0]].reshape(steps + 1, pr.shape[0]), yf

    # ---- port modal extraction on the device (rhs.rs:1211-1242) ----
    def set_port(self, port: int, elem, local_face, profiles, weights) -> None:
        """Port `port`: (element, local face) pairs, (e_t, h_t) profiles per
        face node (n_faces, Nfp, 6) and weights 2*area*face-node weight
        (n_faces, Nfp)."""
        elem = np.ascontiguousarray(elem, np.int32)
        lf = np.ascontiguousarray(local_face, np.int32)
        prof = np.ascontiguousarray(profiles, np.float64).reshape(-1)
        w = np.ascontiguousarray(weights, np.float64).reshape(-1)
        L.check(L.lib.rfhip_op_set_port(self._h, int(port), elem.shape[0], elem, lf, prof, w))

    @property
    def n_ports(self) -> int:
        return int(L.lib.rfhip_op_n_ports(self._h))

    def port_source(self, port: int, download: bool = True) -> np.ndarray | None:
        """rhs.rs:1157-1199 MaxwellOperator::port_source, assembled on the
        device from set_port's faces and profiles; it stays there as the
        vector source for Drive.port_source(). Returns the n_dof pattern when
        download is true."""
        b = np.empty(self._n) if download else None
        L.check(L.lib.rfhip_op_port_source_f64(self._h, int(port), L.ptr(b)))
        return b

    def transient_ports(self, ctx, y0, dt, steps, substeps=1, drive: Drive | None = None, g=None):
        """Explicit transient recording (P_e, P_h) of every port per frame on
        the device: ((steps+1, n_ports, 2) signals, final state)."""
        drive = drive or Drive.free()
        y0 = self._vec(y0, np.float64)
        substeps = max(1, int(substeps))
        gg = None if g is None else np.ascontiguousarray(g, np.float64)
        n_g = 0 if gg is None else gg.shape[0]
        npo = self.n_ports
        sig = np.zeros(max(1, (steps + 1) * npo * 2))
        yf = np.empty(self._n)
        L.check(L.lib.rfhip_op_transient_ports_f64(self._h, y0, float(dt), int(steps), substeps, drive.kind,
                                                   drive.dof, L.ptr(drive.vec), L.ptr(gg), n_g, sig, L.ptr(yf)))
        return sig[:(steps + 1) * npo * 2].reshape(steps + 1, npo, 2), yf

    # gpu/operator.rs:823-893
    def transient_kcl(self, ctx, y0, dt, steps, drive: Drive | None = None, g=None, ctl: Controller | None = None,
                      precision="f32"):
        drive = drive or Drive.free()
        ctl = ctl or Controller()
        f64 = precision == "f64"
        dtp = np.float64 if f64 else np.float32
        y0 = self._vec(y0, dtp)
        gg = None if g is None else np.ascontiguousarray(g, dtp)
        n_g = 0 if gg is None else gg.shape[0]
        out = np.empty((steps + 1) * self._n, dtp)
        st = L.KclStatsC()
        cc = ctl.c()
        fn = L.lib.rfhip_op_transient_kcl_f64 if f64 else L.lib.rfhip_op_transient_kcl
        L.check(fn(self._h, y0, dt, int(steps), drive.kind, drive.dof, L.ptr(drive.vec), L.ptr(gg), n_g,
                   C.byref(cc), out, C.byref(st)))
        return out.reshape(steps + 1, self._n), KclStats(st.accepted, st.rejected, st.h_min, st.h_max)

    def set_kcl_registers(self, mode: str) -> None:
        """KCL registers for later transient_kcl calls (rfhip_op_set_kcl_registers;
        no reference counterpart): "f64" (default; the update in the stage
        kernel's epilogue where the operator allows it), "compressed" (y and
        its backup f32 + bf16, err f32; forms with compressed state registers;
        fused the same way), or "compressed_separate" / "f64_separate" (the
        same registers, the update always in a pass of its own)."""
        modes = {"f64": 0, "compressed": 1, "compressed_separate": 2, "f64_separate": 3}
        if mode not in modes:
            raise ValueError(f"unknown KCL register mode {mode!r}")
        L.check(L.lib.rfhip_op_set_kcl_registers(self._h, modes[mode]))

    def set_krylov_basis(self, mode: str) -> None:
        """Basis storage of later expmv_adaptive calls (rfhip_op_set_krylov_basis;
        no reference counterpart): "f64" (default) or "compressed" (finished
        columns f32 + bf16, 6 bytes per entry; for tolerances well above 1e-9)."""
        modes = {"f64": 0, "compressed": 1}
        if mode not in modes:
            raise ValueError(f"unknown Krylov basis mode {mode!r}")
        L.check(L.lib.rfhip_op_set_krylov_basis(self._h, modes[mode]))

    # gpu/operator.rs:1175-1197 (CPU algorithm propagator.rs:186-326)
    def expmv(self, ctx, v, t, m, tol=0.0):
        v = self._vec(v, np.float64)
        out = np.empty_like(v)
        dim = C.c_int32()
        L.check(L.lib.rfhip_op_expmv(self._h, v, float(t), int(m), float(tol), out, C.byref(dim)))
        self.last_dim = dim.value
        return out

    def expmv_adaptive(self, ctx, v, t_end, ctl=None, proposer=None, log_cap: int = 100_000, hess_dim: int = 0):
        """exp(t_end A) v with (h, m) chosen per step (rfhip_op_expmv_adaptive;
        no reference counterpart: the reference's exponential route steps at a
        fixed h and m, session.rs:700-726). ctl: an ExpController; proposer:
        a trained Proposer, or None for the plain controller. Returns
        (out, step records (numpy EXP_STEP_DTYPE), Hessenberg blocks
        (n_steps, hess_dim+1, hess_dim) or None)."""
        from .proposer import ExpController

        ctl = ctl or ExpController()
        v = self._vec(v, np.float64)
        out = np.empty_like(v)
        steps = np.zeros(max(int(log_cap), 1), L.EXP_STEP_DTYPE)
        hess = np.zeros((max(int(log_cap), 1), hess_dim + 1, hess_dim)) if hess_dim > 0 else None
        n = C.c_int64()
        cc = ctl.c()
        L.check(L.lib.rfhip_op_expmv_adaptive(self._h, v, float(t_end), C.byref(cc),
                                              None if proposer is None else proposer._h,
                                              steps.ctypes.data_as(C.c_void_p), int(log_cap),
                                              None if hess is None else hess.ctypes.data_as(C.c_void_p),
                                              int(hess_dim), C.byref(n), out))
        k = min(n.value, int(log_cap))
        self.last_steps = n.value
        return out, steps[:k].copy(), None if hess is None else hess[:k].copy()

    # gpu/operator.rs:1267-1298 (CPU algorithm propagator.rs:338-373)
    def etd_step(self, ctx, y, b, h, m, tol=0.0):
        y = self._vec(y, np.float64)
        b = self._vec(b, np.float64)
        out = np.empty_like(y)
        dim = C.c_int32()
        L.check(L.lib.rfhip_op_etd_step(self._h, y, b, float(h), int(m), float(tol), out, C.byref(dim)))
        self.last_dim = dim.value
        return out

    # ---- kernel variants (no reference counterpart) ----
    def n_variants(self) -> int:
        return int(L.lib.rfhip_op_n_variants(self._h))

    def autotune(self) -> tuple[int, float]:
        """Time every kernel variant on this operator and keep the fastest."""
        v, ms = C.c_int(), C.c_double()
        L.check(L.lib.rfhip_op_autotune(self._h, C.byref(v), C.byref(ms)))
        return v.value, ms.value

    def set_variant(self, v: int) -> None:
        L.check(L.lib.rfhip_op_set_variant(self._h, int(v)))

    @property
    def variant(self) -> int:
        return int(L.lib.rfhip_op_variant(self._h))

    def variant_available(self, v: int) -> bool:
        return bool(L.lib.rfhip_op_variant_available(self._h, int(v)))

    def poison_lds(self) -> None:
        """Fill every CU's LDS with NaN (a test aid: a kernel that reads LDS
        it did not write then fails its parity check)."""
        L.check(L.lib.rfhip_op_poison_lds(self._h))

    @property
    def last_call_ms(self) -> float:
        """Device-part wall time of the last expmv / etd_step / transient_kcl
        call, host transfers excluded (no reference counterpart)."""
        ms = C.c_double()
        L.check(L.lib.rfhip_op_last_call_ms(self._h, C.byref(ms)))
        return ms.value

    # ---- device-resident state (no reference counterpart) ----
    def state_load(self, y0):
        L.check(L.lib.rfhip_op_state_load(self._h, self._vec(y0, np.float64)))

    def state_store(self) -> np.ndarray:
        out = np.empty(self._n)
        L.check(L.lib.rfhip_op_state_store(self._h, out))
        return out

    def state_lserk(self, dt, nsteps):
        L.check(L.lib.rfhip_op_state_lserk(self._h, float(dt), int(nsteps)))

    def state_lserk_graph(self, dt, nsteps):
        """state_lserk replayed from a captured HIP graph of two steps."""
        L.check(L.lib.rfhip_op_state_lserk_graph(self._h, float(dt), int(nsteps)))

    def state_lserk_timed(self, dt, nsteps):
        tot, stage = C.c_double(), C.c_double()
        L.check(L.lib.rfhip_op_state_lserk_timed(self._h, float(dt), int(nsteps), C.byref(tot), C.byref(stage)))
        return tot.value, stage.value

    def state_energy(self) -> float:
        e = C.c_double()
        L.check(L.lib.rfhip_op_state_energy(self._h, self._mass, self._det, C.byref(e)))
        return e.value


def batch_lserk(ops: list[GpuOperator], dts, nsteps: int) -> None:
    """nsteps free LSERK4 steps of every operator's resident state, dts[i]
    for ops[i], replayed from one graph in which the operators' streams run
    concurrently (the batching form; no reference counterpart)."""
    n = len(ops)
    handles = (C.c_void_p * n)(*[g._h.value for g in ops])
    d = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, np.float64), (n,)))
    L.check(L.lib.rfhip_batch_lserk(handles, n, d, int(nsteps)))


def batch_lserk_ports(ops: list[GpuOperator], dts, nsteps: int, waves) -> None:
    """As batch_lserk, each operator with a waveform (waves[i], nsteps
    samples held per step; None = free) driven by its device-assembled port
    source (GpuOperator.port_source) (the batching form with port
    excitations; no reference counterpart)."""
    n = len(ops)
    if len(waves) != n:
        raise ValueError("one waveform (or None) per operator")
    handles = (C.c_void_p * n)(*[g._h.value for g in ops])
    d = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, np.float64), (n,)))
    keep = []
    ptrs = (C.c_void_p * n)()
    for i, w in enumerate(waves):
        if w is None:
            ptrs[i] = None
            continue
        w = np.ascontiguousarray(w, np.float64)
        if w.shape != (int(nsteps),):
            raise ValueError(f"waveform {i} needs {nsteps} samples, got {w.shape}")
        keep.append(w)
        ptrs[i] = w.ctypes.data_as(C.c_void_p).value
    L.check(L.lib.rfhip_batch_lserk_ports(handles, n, d, int(nsteps), ptrs))


@dataclass
class CsrMatrix:
    """rhs.rs:1575-1591 CsrMatrix: dimension n, row_ptr (n+1), col_idx and
    values (nnz), column indices int64 (the reference's usize). The arrays
    are None when the matrix was assembled without download."""
    n: int
    row_ptr: np.ndarray | None
    col_idx: np.ndarray | None
    values: np.ndarray | None
    n_stored: int
    count_ms: float = 0.0  # device time of the counting pass over the rows
    fill_ms: float = 0.0   # device time of the filling pass

    def nnz(self) -> int:
        return self.n_stored


def assemble_sparse(ctx: GpuContext, op: MaxwellOperator | None = None, *, desc: L.OpDesc | None = None,
                    dispersive=None, download: bool = True) -> CsrMatrix:
    """MaxwellOperator::assemble_sparse (rhs.rs:1621-1730) on the device: the
    state-space A as CSR, pattern and values those of the reference's
    unit-column probe. From a MaxwellOperator (its Debye elements included)
    or from raw rfhip_op_desc data with dispersive = (elements, a, g,
    1/eps_inf) as for GpuOperator.from_desc. download=False keeps only the
    shape and the pass timings (the device copy is released)."""
    keep = None
    if op is not None:
        desc, keep = op.desc()
        if getattr(op, "n_dispersive", 0):
            dispersive = (op.disp_elem, op.disp_a, op.disp_g, op.disp_inv_eps_inf)
    if desc is None:
        raise ValueError("assemble_sparse needs a MaxwellOperator or a desc")
    n_disp = 0
    arrs = (None, None, None, None)
    if dispersive is not None:
        de = np.ascontiguousarray(dispersive[0], np.int64)
        n_disp = de.shape[0]
        arrs = (de, *(np.ascontiguousarray(np.broadcast_to(v, (n_disp,)), np.float64) for v in dispersive[1:]))
    h = C.c_void_p()
    L.check(L.lib.rfhip_assemble_sparse(ctx._h, C.byref(desc), n_disp, *(L.ptr(a) for a in arrs), C.byref(h)))
    del keep
    try:
        n, nnz = C.c_int64(), C.c_int64()
        L.check(L.lib.rfhip_csr_shape(h, C.byref(n), C.byref(nnz)))
        rp = ci = va = None
        if download:
            rp = np.empty(n.value + 1, np.int64)
            ci = np.empty(nnz.value, np.int64)
            va = np.empty(nnz.value, np.float64)
            L.check(L.lib.rfhip_csr_download(h, L.ptr(rp), L.ptr(ci), L.ptr(va)))
        cm, fm = C.c_double(), C.c_double()
        L.check(L.lib.rfhip_csr_timing(h, C.byref(cm), C.byref(fm)))
    finally:
        L.lib.rfhip_csr_destroy(h)
    return CsrMatrix(int(n.value), rp, ci, va, int(nnz.value), cm.value, fm.value)
