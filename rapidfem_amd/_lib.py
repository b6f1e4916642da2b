"""ctypes binding of librfhip.so (include/rfhip.h and include/rfhost.h).

The library is required: importing the package on a machine without the
built librfhip.so raises immediately. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "librfhip.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_d = C.c_double

RFHIP_OK = 0
ERRORS = {1: "EINVAL", 2: "ENODEV", 3: "ENOMEM", 4: "EHIP", 5: "ECOLLAPSE", 6: "EUNSUPPORTED"}

FORM_PLAIN_F32 = 0
FORM_PLAIN_F64 = 1
FORM_FUSED_F64 = 2
FORM_FUSED_F64C = 3
FORM_FUSED_BF16X3 = 4
FORM_FUSED_BF16X2 = 5
FORM_FUSED_BF16X4 = 6
FORMS = {"plain_f32": 0, "plain_f64": 1, "fused_f64": 2, "fused_f64c": 3, "fused_bf16x3": 4, "fused_bf16x2": 5,
         "fused_bf16x4": 6}

DRIVE_FREE, DRIVE_POINT, DRIVE_VECTOR = 0, 1, 2


class OpDesc(C.Structure):
    """rfhip_op_desc: the arrays GpuOperator::new reads (gpu/operator.rs:216-292)."""

    _fields_ = [
        ("np", _i32), ("nfp", _i32), ("n_elem", _i64),
        ("diff_r", _vp), ("diff_s", _vp), ("diff_t", _vp), ("lift", _vp), ("face_nodes", _vp),
        ("jinv", _vp), ("inv_eps", _vp), ("inv_mu", _vp), ("sigma_eps", _vp), ("sigma_mu", _vp),
        ("face_normal", _vp), ("face_fscale", _vp), ("face_neighbor", _vp), ("face_nbr_local", _vp),
        ("face_port", _vp), ("face_perm", _vp), ("flux_alpha", _d),
    ]


class ControllerC(C.Structure):
    _fields_ = [(n, _d) for n in ("atol", "rtol", "safety", "growth_limit", "shrink_limit",
                                  "pi_alpha", "pi_beta", "min_step_factor")]


class KclStatsC(C.Structure):
    _fields_ = [("accepted", _i64), ("rejected", _i64), ("h_min", _d), ("h_max", _d)]


class ExpCtlC(C.Structure):
    """rfhip_expctl: the adaptive Krylov exponential's controller."""

    _fields_ = [("tol", _d), ("m", _i32), ("m_min", _i32), ("m_max", _i32), ("h0", _d), ("safety", _d),
                ("growth_limit", _d), ("shrink_limit", _d), ("max_steps", _i64), ("h_scale", _d)]


# rfhip_exp_step as a numpy record (same layout as the C struct)
EXP_STEP_DTYPE = np.dtype([("t", "f8"), ("h", "f8"), ("h_first", "f8"), ("est_first", "f8"), ("bound_first", "f8"),
                           ("est", "f8"), ("bound", "f8"), ("beta", "f8"), ("t_left", "f8"), ("ms", "f8"),
                           ("m", "i4"), ("proposer", "i4"), ("n_judged", "i4"), ("breakdown", "i4"),
                           ("feat", "f4", (12,))])


# Exported symbols and signatures, in header order. tests/ check every one
# of these is present in the built library.
SIGNATURES = {
    # rfhip.h
    "rfhip_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "rfhip_ctx_create": (C.c_int, [C.POINTER(_vp)]),
    "rfhip_ctx_destroy": (None, [_vp]),
    "rfhip_ctx_device_name": (C.c_int, [_vp, C.c_char_p, C.c_size_t]),
    "rfhip_ctx_has_fp64": (C.c_int, [_vp]),
    "rfhip_op_create": (C.c_int, [_vp, C.POINTER(OpDesc), C.c_int, C.POINTER(_vp)]),
    "rfhip_op_create_dispersive": (C.c_int, [_vp, C.POINTER(OpDesc), C.c_int, _i64, _lp, _dp, _dp, _dp,
                                             C.POINTER(_vp)]),
    "rfhip_op_n_dispersive": (_i64, [_vp]),
    "rfhip_op_destroy": (None, [_vp]),
    "rfhip_op_n_dof": (_i64, [_vp]),
    "rfhip_op_form": (C.c_int, [_vp]),
    "rfhip_op_apply": (C.c_int, [_vp, _fp, _fp]),
    "rfhip_op_apply_f64": (C.c_int, [_vp, _dp, _dp]),
    "rfhip_op_transient": (C.c_int, [_vp, _fp, C.c_float, _i64, _i64, C.c_int, _i64, _vp, _vp, _i64,
                                     C.c_int, _fp]),
    "rfhip_op_transient_f64": (C.c_int, [_vp, _dp, _d, _i64, _i64, C.c_int, _i64, _vp, _vp, _i64,
                                         C.c_int, _dp]),
    "rfhip_op_transient_probes_f64": (C.c_int, [_vp, _dp, _d, _i64, _i64, C.c_int, _i64, _vp, _vp, _i64, _lp, _i64,
                                                _dp, _vp]),
    "rfhip_op_set_port": (C.c_int, [_vp, _i32, _i64, _ip, _ip, _dp, _dp]),
    "rfhip_op_n_ports": (C.c_int, [_vp]),
    "rfhip_op_port_source_f64": (C.c_int, [_vp, C.c_int32, _vp]),
    "rfhip_op_transient_ports_f64": (C.c_int, [_vp, _dp, _d, _i64, _i64, C.c_int, _i64, _vp, _vp, _i64, _dp, _vp]),
    "rfhip_op_transient_kcl": (C.c_int, [_vp, _fp, C.c_float, _i64, C.c_int, _i64, _vp, _vp, _i64,
                                         C.POINTER(ControllerC), _fp, C.POINTER(KclStatsC)]),
    "rfhip_op_set_kcl_registers": (C.c_int, [_vp, C.c_int32]),
    "rfhip_op_set_krylov_basis": (C.c_int, [_vp, C.c_int32]),
    "rfhip_op_transient_kcl_f64": (C.c_int, [_vp, _dp, _d, _i64, C.c_int, _i64, _vp, _vp, _i64,
                                             C.POINTER(ControllerC), _dp, C.POINTER(KclStatsC)]),
    "rfhip_op_expmv": (C.c_int, [_vp, _dp, _d, _i32, _d, _dp, C.POINTER(_i32)]),
    "rfhip_op_etd_step": (C.c_int, [_vp, _dp, _dp, _d, _i32, _d, _dp, C.POINTER(_i32)]),
    "rfhip_op_expmv_adaptive": (C.c_int, [_vp, _dp, _d, C.POINTER(ExpCtlC), _vp, _vp, _i64, _vp, _i32,
                                          C.POINTER(_i64), _dp]),
    "rfhip_krylov_labels": (C.c_int, [_dp, _i32, _i64, _ip, _dp, _dp, _dp, _dp, _ip, _i32, _dp, _i32, _dp, _ip,
                                      _vp]),
    "rfhip_expm_pade": (C.c_int, [_dp, _i32, _dp]),
    "rfhip_mlp_create": (C.c_int, [_vp, C.c_uint64, C.POINTER(_vp)]),
    "rfhip_mlp_destroy": (None, [_vp]),
    "rfhip_mlp_n_params": (C.c_int, []),
    "rfhip_mlp_get_params": (C.c_int, [_vp, _vp, _vp]),
    "rfhip_mlp_set_params": (C.c_int, [_vp, _vp, _vp]),
    "rfhip_mlp_forward": (C.c_int, [_vp, _fp, _i64, _fp]),
    "rfhip_mlp_loss_grad": (C.c_int, [_vp, _fp, _fp, _vp, _i64, C.POINTER(_d), _vp]),
    "rfhip_mlp_train": (C.c_int, [_vp, _fp, _fp, _vp, _i64, _i32, _i32, C.c_float, C.c_uint64, _vp]),
    "rfhip_op_last_call_ms": (C.c_int, [_vp, C.POINTER(_d)]),
    "rfhip_op_n_variants": (C.c_int, [_vp]),
    "rfhip_op_autotune": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(_d)]),
    "rfhip_op_set_variant": (C.c_int, [_vp, C.c_int]),
    "rfhip_op_poison_lds": (C.c_int, [_vp]),
    "rfhip_op_variant": (C.c_int, [_vp]),
    "rfhip_op_variant_available": (C.c_int, [_vp, C.c_int]),
    "rfhip_op_state_load": (C.c_int, [_vp, _dp]),
    "rfhip_op_state_store": (C.c_int, [_vp, _dp]),
    "rfhip_op_state_lserk": (C.c_int, [_vp, _d, _i64]),
    "rfhip_op_state_lserk_graph": (C.c_int, [_vp, _d, _i64]),
    "rfhip_batch_lserk": (C.c_int, [C.POINTER(_vp), _i32, _dp, _i64]),
    "rfhip_batch_lserk_ports": (C.c_int, [C.POINTER(_vp), _i32, _dp, _i64, C.POINTER(_vp)]),
    "rfhip_op_state_lserk_timed": (C.c_int, [_vp, _d, _i64, C.POINTER(_d), C.POINTER(_d)]),
    "rfhip_op_state_energy": (C.c_int, [_vp, _dp, _dp, C.POINTER(_d)]),
    "rfhip_op_energy_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rfhip_assemble_sparse": (C.c_int, [_vp, C.POINTER(OpDesc), _i64, _vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "rfhip_csr_destroy": (None, [_vp]),
    "rfhip_csr_shape": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64)]),
    "rfhip_csr_download": (C.c_int, [_vp, _vp, _vp, _vp]),
    "rfhip_csr_device": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "rfhip_csr_timing": (C.c_int, [_vp, C.POINTER(_d), C.POINTER(_d)]),
    # rfhost.h
    "rfhost_np": (_i32, [_i32]),
    "rfhost_nfp": (_i32, [_i32]),
    "rfhost_refelem": (C.c_int, [_i32, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp]),
    "rfhost_box_mesh": (None, [_i32, _i32, _i32, _d, _d, _d, _d, C.c_uint64, _i32, _dp, _lp]),
    "rfhost_operator_arrays": (C.c_int, [_i64, _dp, _i64, _lp, _i32, _vp, _vp, _vp, _vp, _vp,
                                         _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp]),
    "rfhost_node_coords": (None, [_i64, _dp, _lp, _i32, _dp]),
}


class RfhipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rfhip {ERRORS.get(code, code)}: {msg}")
        self.code = code


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc for gfx950). There is no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    lib.rfhip_last_error(buf, 4096)
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    if rc != RFHIP_OK:
        raise RfhipError(rc, last_error())


def ptr(a) -> C.c_void_p | None:
    return None if a is None else a.ctypes.data_as(C.c_void_p)
