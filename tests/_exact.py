"""Exact rational reference computations for pinning the oracle and the kernels.

Everything here uses fractions.Fraction on the exact binary values of the
doubles, so the answers are the mathematical ones.
"""
from __future__ import annotations

from fractions import Fraction as F

import numpy as np

FACES = [(1, 3, 2), (0, 2, 3), (0, 3, 1), (0, 1, 2)]


def fr(v):
    return [F(float(x)) for x in v]


def det3(u, v, w):
    return (u[0] * (v[1] * w[2] - v[2] * w[1]) + u[1] * (v[2] * w[0] - v[0] * w[2])
            + u[2] * (v[0] * w[1] - v[1] * w[0]))


def orient3d(a, b, c, d) -> int:
    a, b, c, d = fr(a), fr(b), fr(c), fr(d)
    s = det3([b[i] - a[i] for i in range(3)], [c[i] - a[i] for i in range(3)],
             [d[i] - a[i] for i in range(3)])
    return (s > 0) - (s < 0)


def in_tet(nodes, tet, p) -> int:
    """0 outside, 1 on the boundary, 2 strictly inside (any orientation)."""
    P = [nodes[v] for v in tet]
    D = orient3d(*P)
    strict = True
    for f in range(4):
        a, b, c = (P[k] for k in FACES[f])
        s = orient3d(a, b, c, p) * D
        if s < 0:
            return 0
        if s == 0:
            strict = False
    return 2 if strict else 1


def brute_locate(nodes, tets, p) -> int:
    """Minimum-id tet whose closed volume holds p, -1 if none."""
    for t, tet in enumerate(tets):
        if in_tet(nodes, tet, p):
            return t
    return -1


def seg_tri_param(p, q, a, b, c):
    """Exact intersection of segment p->q with triangle abc for the non-coplanar
    case by Cramer's rule: p + t(q-p) = a + u(b-a) + v(c-a).  Returns t as a
    Fraction, None for no intersection, 'coplanar' when the plane holds the line."""
    p, q, a, b, c = fr(p), fr(q), fr(a), fr(b), fr(c)
    d = [q[i] - p[i] for i in range(3)]
    e1 = [b[i] - a[i] for i in range(3)]
    e2 = [c[i] - a[i] for i in range(3)]
    r = [p[i] - a[i] for i in range(3)]
    # t*d - u*e1 - v*e2 = -r
    M = [[d[i], -e1[i], -e2[i]] for i in range(3)]
    rhs = [-r[i] for i in range(3)]
    cols = [[M[i][k] for i in range(3)] for k in range(3)]
    D = det3(*cols)
    if D == 0:
        return "coplanar"
    def rep(k):
        cc = [list(col) for col in cols]
        cc[k] = rhs
        return det3(*cc) / D
    t, u, v = rep(0), rep(1), rep(2)
    if 0 <= t <= 1 and u >= 0 and v >= 0 and u + v <= 1:
        return t
    return None


def brute_first_hit(nodes, tris, p, q):
    best, best_t = -1, None
    for s, tri in enumerate(tris):
        t = seg_tri_param(p, q, *(nodes[v] for v in tri))
        if t is None or t == "coplanar":
            continue
        if best_t is None or t < best_t:
            best, best_t = s, t
    return best


def barycentric(nodes, tet, p):
    P = np.array([nodes[v] for v in tet], dtype=np.float64)
    A = np.vstack([P.T, np.ones(4)])
    return np.linalg.solve(A, np.append(np.asarray(p, dtype=np.float64), 1.0))
