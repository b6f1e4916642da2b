"""Exact-rational nodal basis of the equispaced tetrahedron, p = 1..4.

Writes tests/golden/exact_basis.json: mass, Dr, Ds, Dt, LIFT and face_nodes
for each order, computed in exact rational arithmetic (fractions.Fraction)
and rounded once to the nearest double. It pins the parts of the basis the
reference's own sympy goldens (crates/rapidfem-td/tests/dg_basis_golden_test.rs,
p = 1..3, no LIFT) leave open: p = 4 and LIFT at every order.

Conventions it has to share with the reference so the tables compare entry
by entry (crates/rapidfem-td/src/dg_basis.rs):
  * node order: (i, j, k) / p for i, then j, then k ascending with
    i + j + k <= p (equispaced_ijk, :270-280);
  * face node sets in node order: face 0 k = 0, face 1 i = 0, face 2 j = 0,
    face 3 i + j + k = p (:295-301);
  * face parameters (u, v): face 0 (i, j), 1 (j, k), 2 (i, k), 3 (i, j),
    with the face mass integrated over the unit right triangle in (u, v)
    (:303-330); the physical face area enters later through fscale.
The construction itself is independent of the reference's code: Lagrange
polynomials from an exact Vandermonde inverse, exact monomial moments
a! b! c! / (a + b + c + 3)! on the tetrahedron and a! b! / (a + b + 2)! on
the triangle.

Run: python tests/golden/make_exact_basis.py   (a few seconds)
"""
from __future__ import annotations

import json
import math
import os
from fractions import Fraction as F

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "exact_basis.json")


def ijk_nodes(p):
    return [(i, j, k) for i in range(p + 1) for j in range(p + 1 - i) for k in range(p + 1 - i - j)]


def monos3(p):
    return [(a, b, c) for a in range(p + 1) for b in range(p + 1 - a) for c in range(p + 1 - a - b)]


def monos2(p):
    return [(a, b) for a in range(p + 1) for b in range(p + 1 - a)]


def inverse(m):
    """Exact Gauss-Jordan inverse of a square matrix of Fractions."""
    n = len(m)
    a = [row[:] + [F(int(i == j)) for j in range(n)] for i, row in enumerate(m)]
    for col in range(n):
        piv = next(r for r in range(col, n) if a[r][col] != 0)
        a[col], a[piv] = a[piv], a[col]
        inv = 1 / a[col][col]
        a[col] = [x * inv for x in a[col]]
        for r in range(n):
            if r != col and a[r][col] != 0:
                f = a[r][col]
                a[r] = [x - f * y for x, y in zip(a[r], a[col])]
    return [row[n:] for row in a]


def matmul(a, b):
    bt = list(zip(*b))
    return [[sum((x * y for x, y in zip(row, col)), F(0)) for col in bt] for row in a]


def tet_moment(a, b, c):
    return F(math.factorial(a) * math.factorial(b) * math.factorial(c), math.factorial(a + b + c + 3))


def tri_moment(a, b):
    return F(math.factorial(a) * math.factorial(b), math.factorial(a + b + 2))


def basis(p):
    nodes = ijk_nodes(p)
    x = [tuple(F(q, p) for q in nd) for nd in nodes]
    mon = monos3(p)
    n = len(nodes)
    # coef[m][j]: coefficient of monomial m in the Lagrange polynomial of node j
    vand = [[xi[0] ** a * xi[1] ** b * xi[2] ** c for (a, b, c) in mon] for xi in x]
    coef = inverse(vand)
    gram = [[tet_moment(a1 + a2, b1 + b2, c1 + c2) for (a2, b2, c2) in mon] for (a1, b1, c1) in mon]
    coef_t = [list(r) for r in zip(*coef)]
    mass = matmul(matmul(coef_t, gram), coef)

    def dmono(m, xi, axis):
        e = m[axis]
        if e == 0:
            return F(0)
        d = list(m)
        d[axis] -= 1
        return e * xi[0] ** d[0] * xi[1] ** d[1] * xi[2] ** d[2]

    diffs = []
    for axis in range(3):
        vd = [[dmono(m, xi, axis) for m in mon] for xi in x]
        diffs.append(matmul(vd, coef))

    preds = [lambda q: q[2] == 0, lambda q: q[0] == 0, lambda q: q[1] == 0, lambda q: sum(q) == p]
    params = [lambda q: (q[0], q[1]), lambda q: (q[1], q[2]), lambda q: (q[0], q[2]), lambda q: (q[0], q[1])]
    face_nodes = [[i for i, q in enumerate(nodes) if preds[f](q)] for f in range(4)]
    nfp = len(face_nodes[0])
    m2 = monos2(p)
    emat = [[F(0)] * (4 * nfp) for _ in range(n)]
    for f in range(4):
        uv = [tuple(F(t, p) for t in params[f](nodes[i])) for i in face_nodes[f]]
        v2 = [[u ** a * v ** b for (a, b) in m2] for (u, v) in uv]
        c2 = inverse(v2)
        g2 = [[tri_moment(a1 + a2, b1 + b2) for (a2, b2) in m2] for (a1, b1) in m2]
        mf = matmul(matmul([list(r) for r in zip(*c2)], g2), c2)
        for a, vi in enumerate(face_nodes[f]):
            for b in range(nfp):
                emat[vi][f * nfp + b] = mf[a][b]
    lift = matmul(inverse(mass), emat)
    flt = lambda m: [float(v) for row in m for v in row]  # noqa: E731
    return dict(np=n, nfp=nfp, mass=flt(mass), dr=flt(diffs[0]), ds=flt(diffs[1]), dt=flt(diffs[2]),
                lift=flt(lift), face_nodes=face_nodes)


def main():
    out = {"source": "tests/golden/make_exact_basis.py (exact rationals, rounded once)",
           "orders": {f"p{p}": basis(p) for p in range(1, 5)}}
    with open(OUT, "w") as fh:
        json.dump(out, fh)
    print(OUT)


if __name__ == "__main__":
    main()
