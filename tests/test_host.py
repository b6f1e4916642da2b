"""CPU tests of the host side: the C ABI library loads and exports every
declared symbol, the C++ operator builder reproduces the oracle's arrays,
and the Python mirror of the reference session helpers. No compute call
reaches a GPU here.
"""
import ctypes as C
import json
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

EXACT = json.load(open(os.path.join(GOLDEN, "exact_basis.json")))["orders"]


def header_functions(name):
    """Function names declared in include/<name>."""
    text = open(os.path.join(ROOT, "include", name)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^[A-Za-z_][\w \*]*?\b(rf\w+)\s*\(", text, flags=re.M)


def test_library_exports_every_declared_symbol():
    from rapidfem_amd import _lib as L

    declared = header_functions("rfhip.h") + header_functions("rfhost.h")
    assert len(declared) >= 30
    so = C.CDLL(L.LIB_PATH)
    for fn in declared:
        assert hasattr(so, fn), f"{fn} declared but not exported"
    # the ctypes binding covers exactly the declared surface
    assert sorted(declared) == sorted(L.SIGNATURES)


def test_context_without_device_reports_enodev():
    """GpuContext::new's "no device" error (gpu/mod.rs:48-50) becomes
    RFHIP_ENODEV; run in a child so a machine with a GPU is untouched."""
    code = (
        "import ctypes as C, sys; sys.path.insert(0, %r)\n"
        "from rapidfem_amd import _lib as L\n"
        "h = C.c_void_p()\n"
        "rc = L.lib.rfhip_ctx_create(C.byref(h))\n"
        "print(rc, L.last_error())\n"
        "if rc == 0: L.lib.rfhip_ctx_destroy(h)\n" % ROOT)
    env = dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    rc = int(r.stdout.split()[0])
    import torch

    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    assert rc == 2, r.stdout


def test_error_message_is_thread_local_and_truncates():
    from rapidfem_amd import _lib as L

    h = C.c_void_p()
    rc = L.lib.rfhip_op_create(None, None, 0, C.byref(h))
    assert rc == 1
    full = L.last_error()
    assert full
    buf = C.create_string_buffer(4)
    n = L.lib.rfhip_last_error(buf, 4)
    assert n == len(full) and buf.value == full[:3].encode()


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_host_reference_element_matches_exact(oracle, p):
    """The builder's reference element follows the reference's f64 operation
    sequence: its tables equal the oracle's (pinned by the reference goldens)
    bit for bit, round-off residues included, and the exact-rational tables
    to the reference's own accuracy (p = 4 LIFT leaves ~1e-10 relative)."""
    from rapidfem_amd.maxwell import ReferenceElement

    re_ = ReferenceElement.new(p)
    ex = EXACT[f"p{p}"]
    orc = oracle.refelem(p)
    assert (re_.nodes == orc["nodes"]).all()
    assert (re_.face_nodes == orc["face_nodes"]).all()
    for key, arr in (("mass", re_.mass), ("dr", re_.diff_r), ("ds", re_.diff_s), ("dt", re_.diff_t),
                     ("lift", re_.lift)):
        assert np.array_equal(arr.ravel(), np.asarray(orc[key]).ravel()), key
        want = np.array(ex[key])
        assert np.abs(arr.ravel() - want).max() < 1e-9 * np.abs(want).max(), key
    assert np.array_equal(re_.face_node_weights.ravel(), np.asarray(orc["fweights"]).ravel())


@pytest.mark.parametrize("jit", [False, True])
def test_host_box_mesh_matches_oracle(oracle, jit):
    from rapidfem_amd.maxwell import structured_box, structured_box_jittered

    if jit:
        m = structured_box_jittered(5, 4, 3, 1.0, 2.0, 0.5, 0.1, 11)
        xyz, tets = oracle.box_mesh(5, 4, 3, 1.0, 2.0, 0.5, amplitude=0.1, seed=11, jittered=True)
    else:
        m = structured_box(5, 4, 3, 1.0, 2.0, 0.5)
        xyz, tets = oracle.box_mesh(5, 4, 3, 1.0, 2.0, 0.5)
    assert (m.nodes == xyz).all() and (m.tets == tets).all()


def test_host_operator_arrays_match_oracle(oracle):
    """rfhost_operator_arrays against the oracle's MaxwellOperator::new:
    topology, perm, geometry and materials equal bit for bit (the sparse
    state-space export stores exactly the entries these values make
    nonzero)."""
    from rapidfem_amd.maxwell import MaxwellOperator, structured_box_jittered

    mesh = structured_box_jittered(4, 3, 3, 1.0, 1.0, 1.0, 0.15, 5)
    K = mesh.n_tets
    rng = np.random.default_rng(0)
    eps = rng.uniform(1, 4, (K, 3))
    mu = rng.uniform(1, 2, (K, 3))
    sig = rng.uniform(0, 1, K)
    port = np.full((K, 4), -1, np.int32)
    port[:7, 0] = 3
    op = MaxwellOperator(mesh, 3, 0.6, materials={"eps": eps, "mu": mu, "sigma": sig}, face_port=port)
    oop = oracle.Operator(mesh.nodes, mesh.tets, 3, 0.6, eps=eps, mu=mu, sigma=sig, face_port=port.ravel())
    ex = oop.export()
    assert (op.nbr == ex["nbr"]).all() and (op.nbr_lf == ex["nbr_lf"]).all()
    assert (op.perm == ex["perm"]).all() and (op.port == ex["port"]).all()
    for a, b in ((op.jinv, ex["jinv"]), (op.det, ex["det"]), (op.normal, ex["normal"]),
                 (op.fscale, ex["fscale"]), (op.inv_eps, ex["inv_eps"]), (op.inv_mu, ex["inv_mu"]),
                 (op.sigma_eps, ex["sig_eps"]), (op.sigma_mu, ex["sig_mu"])):
        assert np.array_equal(np.asarray(a).ravel(), np.asarray(b).ravel())
    # a port tag survives only on boundary faces (rhs.rs:862-880)
    assert ((op.port.ravel() >= 0) <= (op.nbr.ravel() < 0)).all()


def test_host_builder_rejects_nonconforming_mesh():
    from rapidfem_amd.maxwell import MaxwellOperator, Mesh

    # two tets sharing a face whose third vertex is displaced on one side
    nodes = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0]])
    tets = np.array([[0, 1, 2, 3], [0, 1, 2, 4], [0, 1, 2, 5]])
    with pytest.raises(RuntimeError, match="non-manifold"):
        MaxwellOperator(Mesh(nodes, tets), 1)


def test_node_coords_map_reference_nodes():
    from rapidfem_amd.maxwell import MaxwellOperator, structured_box_jittered

    mesh = structured_box_jittered(2, 2, 2, 1.0, 1.0, 1.0, 0.2, 1)
    op = MaxwellOperator(mesh, 2)
    x = op.node_coords()
    v = mesh.nodes[mesh.tets]
    for e in (0, 17, 47):
        want = v[e, 0] + op.re.nodes @ (v[e, 1:] - v[e, 0])
        assert np.abs(x[e] - want).max() < 1e-15


def test_gaussian_pulse_closed_forms():
    """session.rs:1366-1372."""
    from rapidfem_amd.session import GaussianPulse

    g = GaussianPulse(1.0, 0.5, 2.0)
    assert g.eval(1.0) == 1.0
    assert abs(GaussianPulse(1.0, 0.5).eval(1.5) - math.exp(-1.0)) < 1e-12
    assert abs(g.eval(1.125) - math.exp(-0.0625) * math.cos(2 * math.pi * 2.0 * 0.125)) < 1e-15


def test_normal_vector_matches_oracle(oracle):
    from rapidfem_amd.session import normal_vector

    # same splitmix64 uniforms; numpy's vectorised log/cos may differ from
    # libm's by an ulp
    for seed in (0, 1, 7):
        a, b = normal_vector(1001, seed), oracle.normal_vector(1001, seed)
        assert np.abs(a - b).max() <= 4e-16 * np.abs(b).max()


def test_workload_shapes_and_unit_costs():
    from rapidfem_amd import workloads as W

    op, y0 = W.cavity_te101()
    assert op.n_elem == 1296 and op.n_dof == 77760
    x = op.node_coords().reshape(-1, 3)
    y = y0.reshape(-1, 6)
    assert np.abs(y[:, 1] - np.sin(math.pi * x[:, 0]) * np.sin(math.pi * x[:, 2])).max() < 1e-15
    assert np.abs(y[:, [0, 2, 3, 4, 5]]).max() == 0.0
    op2, y2 = W.pulse_box(6, 3, 3)
    assert op2.n_elem == 6 * 3 * 3 * 6 and op2.order == 3
    inv = op2.inv_eps[:, 0]
    assert set(np.round(1.0 / inv, 12)) == {1.0, 4.0}
    # algorithmic flops per DOF-stage: 166 / 253 / 379 at N = 2 / 3 / 4
    assert [round(W.flops_per_dof_stage(p)) for p in (2, 3, 4)] == [166, 253, 379]


def test_slab_box_grading_and_materials():
    """The stiff Krylov workload: cell widths graded 20:1 toward a seeded slab
    of eps_r 10, sigma 50; the remap keeps the box's extent and conformity."""
    from rapidfem_amd import workloads as W

    op, y0 = W.slab_box(12, 9, 8, 3)
    assert op.n_elem == 12 * 9 * 8 * 6 and op.n_dof == y0.size
    xs = np.unique(np.round(op.mesh.nodes[:, 0], 14))
    assert xs.size == 13 and xs[0] == 0.0 and abs(xs[-1] - 1.0) < 1e-15
    w = np.diff(xs)
    assert 10.0 < w.max() / w.min() <= 20.0 + 1e-9
    eps = 1.0 / op.inv_eps[:, 0]
    slab = eps > 1.0
    assert set(np.round(eps, 12)) == {1.0, 10.0}
    sigma = op.sigma_eps[:, 0] / op.inv_eps[:, 0]
    assert np.all(sigma[slab] == 50.0) and np.all(sigma[~slab] == 0.0)
    # the slab sits where the cells are finest
    cx = op.mesh.nodes[op.mesh.tets, 0].mean(axis=1)
    xc = 0.5 * (xs[1:] + xs[:-1])
    assert abs(cx[slab].mean() - xc[np.argmin(w)]) < 2 * w.max()
    assert np.linalg.norm(y0) > 0.0


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_port_face_data_weights_and_projection(oracle):
    """Port extraction arrays (rhs.rs:319-329, 1211-1242): the weights of the
    x = 0 wall of the unit box sum to its area 1, and a state equal to the
    profile projects to P_e = P_h = 1 (closed forms)."""
    from rapidfem_amd.maxwell import MaxwellOperator, boundary_faces_on_plane, port_face_data, structured_box_jittered

    mesh = structured_box_jittered(3, 3, 3, 1.0, 1.0, 1.0, 0.1, 4)
    op = MaxwellOperator(mesh, 2)
    faces = boundary_faces_on_plane(op, 0, 0.0)
    assert len(faces) == 2 * 9

    def prof(x):
        et = np.stack([0 * x[:, 0], np.sin(np.pi * x[:, 2]), 0 * x[:, 0]], axis=1)
        ht = np.stack([0 * x[:, 0], 0 * x[:, 0], np.sin(np.pi * x[:, 1])], axis=1)
        return et, ht

    elem, lf, pr, w = port_face_data(op, faces, prof)
    assert abs(w.sum() - 1.0) < 1e-12
    y = np.zeros(op.n_dof)
    nfp = op.re.n_face_nodes
    dof = ((elem[:, None].astype(np.int64) * op.re.n_nodes + op.re.face_nodes[lf]) * 6).reshape(-1)
    y[dof[:, None] + np.arange(6)] = pr.reshape(-1, 6)
    pe, ph = oracle.port_projections(y, dof, pr, w)
    assert abs(pe - 1.0) < 1e-14 and abs(ph - 1.0) < 1e-14
    assert pr.shape == (len(faces), nfp, 6)


def test_debye_material_and_dispersive_operator_host_data():
    """dispersive.rs:41-53 closed forms (the values of its own test material
    eps_inf 2, eps_s 5, tau 0.3) and the host operator's Debye slots:
    n_dof gains 3 Np per slot (rhs.rs:1003-1010), slots follow list order with
    (a, g, 1/eps_inf), and an element out of range or listed twice is refused
    (rhs.rs:970-974)."""
    import rapidfem_amd as R

    m = R.DebyeMaterial(2.0, 5.0, 0.3)
    a, g = m.relaxation_coeffs()
    assert a == -1.0 / 0.3 and g == 3.0 / 0.3
    for w in (0.0, 0.5, 4.0, 1e6):
        re, im = m.permittivity(w)
        z = 2.0 + 3.0 / (1.0 + 1j * w * 0.3)
        assert abs(re - z.real) < 1e-13 and abs(im - z.imag) < 1e-13
    mesh = R.structured_box(1, 1, 1)
    K = mesh.n_tets
    mats = [R.ElemMaterial.isotropic(2.0) for _ in range(K)]
    op = R.MaxwellOperator(mesh, 2, 1.0, mats, dispersive=[(4, m), (1, m)])
    assert op.n_dispersive == 2 and op.n_dof == 6 * 10 * K + 3 * 10 * 2
    assert op.disp_elem.tolist() == [4, 1]
    assert np.allclose(op.disp_a, a) and np.allclose(op.disp_g, g) and np.allclose(op.disp_inv_eps_inf, 0.5)
    for bad in ([(K, m)], [(2, m), (2, m)], [(-1, m)]):
        with pytest.raises(ValueError):
            R.MaxwellOperator(mesh, 2, 1.0, mats, dispersive=bad)
