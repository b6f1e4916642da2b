"""GPU parity: every kernel form against the CPU oracle, through the C ABI.

Operator data comes from the oracle (the reference's f64 values), so these
compare kernels on identical inputs. Tolerances: FP64 forms 1e-12 relative
L2 per operator apply; the f32 form the reference's own GPU gate
GPU_REL_TOL = 1e-3 (constants.rs:50).
"""
import numpy as np
import pytest

from conftest import oracle_desc, rel_l2

pytestmark = pytest.mark.gpu

FORMS64 = ["plain_f64", "fused_f64", "fused_f64c"]
ALL_FORMS = ["plain_f32"] + FORMS64


@pytest.fixture(scope="module")
def ctx():
    from rapidfem_amd import GpuContext

    return GpuContext()


def _mesh(oracle, n=3, jitter=True, lx=1.0):
    return oracle.box_mesh(n, n, n, lx, 1.0, 1.0, 0.1 if jitter else 0.0, 7, jitter)


def _gop(ctx, oop, form):
    from rapidfem_amd import GpuOperator

    d, arrs, re_, ex = oracle_desc(oop)
    g = GpuOperator.from_desc(ctx, d, form, mass=re_["mass"], det=ex["det"])
    return g


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("form", ALL_FORMS)
def test_apply_matches_oracle(ctx, oracle, order, form):
    xyz, tets = _mesh(oracle, 3)
    K = tets.shape[0]
    eps = np.ones((K, 3))
    eps[::3] = 4.0
    sig = np.zeros(K)
    sig[::5] = 0.3
    oop = oracle.Operator(xyz, tets, order, 1.0, eps=eps, sigma=sig)
    n = oop.n_dof
    y = np.sin(0.3 + 0.017 * np.arange(n))
    want = oop.apply(y)
    g = _gop(ctx, oop, form)
    if form == "plain_f32":
        got = g.apply(ctx, y.astype(np.float32))
        tol = 1e-3
    else:
        got = g.apply_f64(y)
        tol = 1e-12
    err = rel_l2(got, want)
    print(f"apply order={order} form={form} rel_l2={err:.3e}")
    assert err < tol
