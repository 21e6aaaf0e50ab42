"""A conforming tet mesh of the unit box full of slivers: the Delaunay
tetrahedralization (scipy's Qhull) of a cubic lattice whose interior points are
jittered by a tiny amount.  The eight corners of every lattice cell are nearly
cospherical, so each cell's Delaunay tets include a flat one; with a jitter of
1e-5 of the spacing their condition numbers reach ~1e6.  Boundary points are
jittered only within their face planes, so the boundary is the box and its
faces are the surface triangles."""
from __future__ import annotations

import numpy as np
from scipy.spatial import Delaunay

from start_amd.meshgen import TetMesh


def sliver_box(n: int = 6, jitter: float = 1e-5, side: float = 0.06, seed: int = 0) -> TetMesh:
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, n + 1)
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    J = rng.uniform(-1.0, 1.0, P.shape) * (jitter / n)
    J[(P == 0.0) | (P == 1.0)] = 0.0  # a coordinate on a box face stays on it
    P += J
    T = Delaunay(P).simplices.astype(np.int32)
    nodes = P * side
    p = nodes[T]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    neg = vol < 0
    T[neg] = T[neg][:, [0, 2, 1, 3]]  # positive orientation
    # boundary faces: the faces one tet has
    faces = np.concatenate([T[:, [1, 2, 3]], T[:, [0, 2, 3]], T[:, [0, 1, 3]], T[:, [0, 1, 2]]])
    key = np.sort(faces, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    tris = np.ascontiguousarray(faces[cnt[inv.reshape(-1)] == 1], dtype=np.int32)
    return TetMesh(np.ascontiguousarray(nodes), np.ascontiguousarray(T), tris)


def condition_numbers(m: TetMesh) -> np.ndarray:
    p = m.nodes[m.tets]
    A = np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], 1)
    return np.linalg.cond(A)
