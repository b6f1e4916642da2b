"""The compiled reference-element tables of the component-lane stage kernel
(rapidfem_amd/csrc/rfhip_elem_tables.hpp) against the host builder they are
generated from: face node sets and the zero pattern of D_r, D_s, D_t at
orders 1-4."""
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GEN = os.path.join(HERE, "..", "rapidfem_amd", "csrc", "gen_elem_tables.py")


def _gen():
    spec = importlib.util.spec_from_file_location("gen_elem_tables", GEN)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_committed_header_matches_the_host_builder():
    g = _gen()
    with open(g.HEADER) as f:
        assert f.read() == g.render()


def test_pattern_separates_exact_zeros_from_entries():
    # every entry outside the pattern is round-off (below 1e-12 of the row
    # scale) and every entry inside is at least 1/3: no threshold ambiguity
    from rapidfem_amd.maxwell import ReferenceElement

    g = _gen()
    for p in g.ORDERS:
        np_, _, fn, masks = g.tables(p)
        r = ReferenceElement.new(p)
        assert np.array_equal(np.asarray(fn), r.face_nodes)
        for d, mk in zip((r.diff_r, r.diff_s, r.diff_t), masks):
            inside = np.array([[(mk[i] >> j) & 1 for j in range(np_)] for i in range(np_)], bool)
            assert np.abs(d[~inside]).max(initial=0.0) < 1e-12 * np.abs(d).max()
            assert np.abs(d[inside]).min() >= 1.0 / 3.0 - 1e-12
