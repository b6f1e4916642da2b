"""Extract the golden vectors the reference's own tests hold, as data.

Run in the build container only (it reads the reference checkout); the GPU
box and the test suite read the committed JSON fixtures it writes.

Sources (relative to the reference root; all tables were generated there by
derivations/td_dg/*.py in exact sympy arithmetic):
  crates/rapidfem-td/tests/dg_basis_golden_test.rs   mass, Dr, Ds, Dt, p=1..3
  crates/rapidfem-td/tests/rhs_curl_golden_test.rs   element_curl, p=2
  crates/rapidfem-td/tests/propagator_expm_test.rs   exp(hA)v0, h=0.1/0.3/1.0

Only the numeric tables are copied (as JSON arrays); no code is.
"""
from __future__ import annotations

import json
import os
import re
import sys

REF = os.environ.get("RAPIDFEM_REFERENCE", "/root/reference")
TESTS = os.path.join(REF, "crates", "rapidfem-td", "tests")
OUT = os.path.dirname(os.path.abspath(__file__))

_CONST = re.compile(r"^const\s+(\w+)\s*:\s*[^=]+=\s*(.+);\s*$", re.M)


def _numbers(text: str):
    return [float(x) for x in re.findall(r"[-+]?\d+\.\d*(?:[eE][-+]?\d+)?|[-+]?\d+[eE][-+]?\d+", text)]


def _consts(path: str) -> dict:
    src = open(path).read()
    out = {}
    for name, body in _CONST.findall(src):
        if body.strip().startswith("["):
            out[name] = _numbers(body)
        else:
            vals = _numbers(body)
            out[name] = vals[0] if len(vals) == 1 else body.strip()
    return out


def main() -> int:
    basis = _consts(os.path.join(TESTS, "dg_basis_golden_test.rs"))
    fixture = {}
    for p in (1, 2, 3):
        fixture[f"p{p}"] = {k.lower(): basis[f"{k}_P{p}"] for k in ("MASS", "DR", "DS", "DT")}
    json.dump({"source": "crates/rapidfem-td/tests/dg_basis_golden_test.rs", "tol": 1e-9,
               "orders": fixture}, open(os.path.join(OUT, "dg_basis_golden.json"), "w"))

    curl = _consts(os.path.join(TESTS, "rhs_curl_golden_test.rs"))
    json.dump({"source": "crates/rapidfem-td/tests/rhs_curl_golden_test.rs", "p": 2, "tol": 1e-9,
               "tet": curl["TET"], "field": curl["FIELD"], "curl": curl["CURL"]},
              open(os.path.join(OUT, "rhs_curl_golden.json"), "w"))

    ex = _consts(os.path.join(TESTS, "propagator_expm_test.rs"))
    json.dump({"source": "crates/rapidfem-td/tests/propagator_expm_test.rs", "m": 12, "tol": 1e-10,
               "n": 6, "a": ex["A"], "v0": ex["V0"],
               "cases": [{"h": ex["H0"], "ref": ex["REF0"]}, {"h": ex["H1"], "ref": ex["REF1"]},
                         {"h": ex["H2"], "ref": ex["REF2"]}]},
              open(os.path.join(OUT, "propagator_expm_golden.json"), "w"))
    print("wrote fixtures to", OUT)
    return 0


if __name__ == "__main__":
    sys.exit(main())
