"""rfhip_op_energy_noalloc, the allocation-free variant of rfhip_op_state_energy
(no reference counterpart), through the C ABI: two calls in a row, the second
on a smaller mesh, each equal to the allocating entry point's energy of the
same state (FP64, different reduction order: 1e-13 relative), plus its
refusals (host pointers, meshes beyond the scratch's 1,048,576 elements).
Device memory for the inputs comes from the HIP runtime librfhip.so itself
links (hipMalloc through ctypes; plumbing of the test only)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from rapidfem_amd import GpuContext

    return GpuContext()


class _Dev:
    """A device copy of a numpy array, from the HIP runtime librfhip.so links."""

    hip = None

    def __init__(self, a):
        if _Dev.hip is None:
            hip = C.CDLL("libamdhip64.so.7")
            hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            hip.hipFree.argtypes = [C.c_void_p]
            for f in (hip.hipMalloc, hip.hipMemcpy, hip.hipFree):
                f.restype = C.c_int
            _Dev.hip = hip
        a = np.ascontiguousarray(a, np.float64)
        self.nbytes = max(a.nbytes, 8)
        self.ptr = C.c_void_p()
        assert _Dev.hip.hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        assert _Dev.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device

    def get(self, n):
        out = np.empty(n)
        assert _Dev.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, 8 * n, 2) == 0  # device to host
        return out

    def __del__(self):
        if _Dev.hip is not None and self.ptr:
            _Dev.hip.hipFree(self.ptr)


def _noalloc_energy(g, op, y):
    from rapidfem_amd import _lib as L

    args = [_Dev(y), _Dev(g._mass), _Dev(g._det), _Dev(op.inv_eps), _Dev(op.inv_mu)]
    out = _Dev(np.array([np.nan]))
    L.check(L.lib.rfhip_op_energy_noalloc(g._h, *[a.ptr for a in args], out.ptr))
    return float(out.get(1)[0])


def test_noalloc_energy_equals_state_energy_on_two_calls(ctx):
    import rapidfem_amd as R
    from rapidfem_amd import workloads as W

    cases = [W.pulse_box(20, 10, 10, 3), W.cavity_te101(4, 2)]  # 12,000 then 384 tets
    for k, (op, y0) in enumerate(cases):
        g = R.GpuOperator(ctx, op, "fused_f64")
        g.state_load(y0)
        g.state_lserk(1e-3, 3)
        y = g.state_store()
        want = g.state_energy()
        got = _noalloc_energy(g, op, y)
        print(f"call {k}: K={op.n_elem}, energy {got:.15e} vs state_energy {want:.15e}", flush=True)
        assert want > 0
        assert abs(got - want) <= 1e-13 * want


def test_noalloc_energy_refuses_host_pointers_and_oversized_meshes(ctx):
    import rapidfem_amd as R
    from rapidfem_amd import _lib as L
    from rapidfem_amd import workloads as W

    op, y0 = W.cavity_te101(3, 2)
    g = R.GpuOperator(ctx, op, "fused_f64")
    host = np.zeros(op.n_dof)
    dev = _Dev(np.zeros(8))
    p = dev.ptr
    rc = L.lib.rfhip_op_energy_noalloc(g._h, host.ctypes.data_as(C.c_void_p), p, p, p, p, p)
    assert rc == 1 and "device pointers" in L.last_error()
    from rapidfem_amd.maxwell import MaxwellOperator, structured_box

    big = MaxwellOperator(structured_box(58, 58, 58, 1.0, 1.0, 1.0), 1, 1.0)  # 1,170,672 tets > 1,048,576
    gb = R.GpuOperator(ctx, big, "fused_f64")
    rc = L.lib.rfhip_op_energy_noalloc(gb._h, p, p, p, p, p, p)
    assert rc == 1 and "exceed the scratch" in L.last_error()
