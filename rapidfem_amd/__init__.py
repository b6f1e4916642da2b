"""rapidfem_amd: MI355X (gfx950) backend for rapidfem-td's nodal DGTD
Maxwell time-stepper.

The compute path is librfhip.so (hand-written HIP kernels behind a C ABI,
include/rfhip.h); this package is the Python host mirror of the reference's
GpuContext / GpuOperator / TdSession GPU routes and the operator builder.
"""
from ._lib import RfhipError  # noqa: F401  (fails loudly if librfhip.so is missing)
from .gpu import (Controller, CsrMatrix, Drive, GpuContext, GpuOperator, KclStats, assemble_sparse,  # noqa: F401
                  batch_lserk, batch_lserk_ports)
from .maxwell import (DebyeMaterial, ElemMaterial, MaxwellOperator, Mesh, ReferenceElement, flux_alpha,  # noqa: F401
                      structured_box, structured_box_jittered)
from .session import GaussianPulse, TdSession  # noqa: F401

__all__ = [
    "RfhipError", "Controller", "CsrMatrix", "assemble_sparse", "batch_lserk", "batch_lserk_ports", "Drive", "GpuContext", "GpuOperator", "KclStats", "DebyeMaterial", "ElemMaterial",
    "MaxwellOperator", "Mesh", "ReferenceElement", "flux_alpha", "structured_box", "structured_box_jittered",
    "GaussianPulse", "TdSession",
]
