"""Shared test setup: markers, import paths and oracle/operator helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


def oracle_desc(oop, flux_alpha=None):
    """rfhip_op_desc built from the oracle's own arrays (reference f64 values),
    so GPU parity compares kernels on identical operator data."""
    import ctypes as C

    from oracle import oracle as O
    from rapidfem_amd import _lib as L

    re_ = O.refelem(oop.order)
    ex = oop.export()
    arrs = dict(diff_r=re_["dr"], diff_s=re_["ds"], diff_t=re_["dt"], lift=re_["lift"],
                face_nodes=np.ascontiguousarray(re_["face_nodes"], np.int32), jinv=ex["jinv"],
                inv_eps=ex["inv_eps"], inv_mu=ex["inv_mu"], sigma_eps=ex["sig_eps"], sigma_mu=ex["sig_mu"],
                face_normal=ex["normal"], face_fscale=ex["fscale"], face_neighbor=ex["nbr"],
                face_nbr_local=ex["nbr_lf"], face_port=ex["port"], face_perm=ex["perm"])
    arrs = {k: np.ascontiguousarray(v) for k, v in arrs.items()}
    d = L.OpDesc()
    d.np, d.nfp, d.n_elem = re_["np"], re_["nfp"], oop.K
    for k, v in arrs.items():
        setattr(d, k, v.ctypes.data_as(C.c_void_p))
    d.flux_alpha = oop.flux_alpha if flux_alpha is None else flux_alpha
    return d, arrs, re_, ex


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
