"""Structured tetrahedral meshes for the workloads the kernel is measured on.

The reference reads Gmsh meshes (VolumetricMesh.cpp:35-121, GmshUtils.cpp:96-152);
there is no Gmsh here, so the box workloads are generated directly:
a structured nx*ny*nz hex grid, every hex split into the 6 Kuhn tetrahedra
that share its main diagonal (all positively oriented, conforming across hexes),
and the boundary faces of the tet mesh as the surface triangles.

Node id:   (k*(ny+1) + j)*(nx+1) + i
Tet id:    6*((k*ny + j)*nx + i) + p,  p indexes the axis permutation
Triangle:  boundary faces in tet-id order, face f opposite local vertex f.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass

import numpy as np

# face f of a tet is opposite local vertex f; vertex order makes
# orient3d(face, opposite vertex) > 0 for a positively oriented tet
FACES = np.array([[1, 3, 2], [0, 2, 3], [0, 3, 1], [0, 1, 2]], dtype=np.int64)


@dataclass
class TetMesh:
    nodes: np.ndarray  # (M, 3) float64
    tets: np.ndarray   # (T, 4) int32
    tris: np.ndarray   # (S, 3) int32 surface triangles, id = row

    @property
    def n_nodes(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def n_tets(self) -> int:
        return int(self.tets.shape[0])

    @property
    def n_tris(self) -> int:
        return int(self.tris.shape[0])


def box_kuhn(nx: int, ny: int, nz: int, lx: float, ly: float, lz: float,
             origin=(0.0, 0.0, 0.0)) -> TetMesh:
    xs = origin[0] + np.arange(nx + 1, dtype=np.float64) * lx / nx
    ys = origin[1] + np.arange(ny + 1, dtype=np.float64) * ly / ny
    zs = origin[2] + np.arange(nz + 1, dtype=np.float64) * lz / nz
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    unit = np.eye(3, dtype=np.int64)
    tets = np.empty((i.size, 6, 4), dtype=np.int64)
    for p, perm in enumerate(itertools.permutations(range(3))):
        c = np.zeros(3, dtype=np.int64)
        corners = [c.copy()]
        for ax in perm:
            c = c + unit[ax]
            corners.append(c.copy())
        for v, off in enumerate(corners):
            tets[:, p, v] = nid(i + off[0], j + off[1], k + off[2])
    tets = tets.reshape(-1, 4)
    tets = _orient_positive(nodes, tets)
    tris = boundary_faces(tets)
    return TetMesh(nodes=nodes, tets=tets.astype(np.int32), tris=tris.astype(np.int32))


def _orient_positive(nodes: np.ndarray, tets: np.ndarray) -> np.ndarray:
    a, b, c, d = (nodes[tets[:, q]] for q in range(4))
    det = np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a))
    tets = tets.copy()
    neg = det < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    return tets


def boundary_faces(tets: np.ndarray) -> np.ndarray:
    """Faces used by exactly one tet, in (tet id, local face) order."""
    faces = tets[:, FACES].reshape(-1, 3)
    s = np.sort(faces, axis=1).astype(np.int64)
    m = int(tets.max()) + 1
    if m ** 3 < 2 ** 62:
        key = (s[:, 0] * m + s[:, 1]) * m + s[:, 2]
        order = np.argsort(key)
        ks = key[order]
        dup = np.zeros(len(ks), dtype=bool)
        same = ks[1:] == ks[:-1]
        dup[1:] |= same
        dup[:-1] |= same
        once = np.zeros(len(ks), dtype=bool)
        once[order] = ~dup
    else:
        _, inv, cnt = np.unique(s, axis=0, return_inverse=True, return_counts=True)
        once = cnt[inv.ravel()] == 1
    return faces[once]


def concentric_disk(a: np.ndarray, b: np.ndarray):
    """Shirley-Chiu concentric map of the square [-1, 1]^2 onto the unit disk:
    area preserving (Jacobian pi/4 everywhere), orientation preserving, and
    square rings go to circles."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    use_a = np.abs(a) > np.abs(b)
    r = np.where(use_a, a, b)
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = np.where(use_a, (np.pi / 4.0) * (b / a), np.pi / 2.0 - (np.pi / 4.0) * (a / b))
    phi = np.where(r == 0.0, 0.0, phi)
    return r * np.cos(phi), r * np.sin(phi)


@dataclass
class Chamber:
    mesh: TetMesh        # tris: the boundary faces, then the substrate triangles
    substrate: np.ndarray  # bool per triangle
    target: np.ndarray     # bool per triangle: boundary faces on the target disk
    radius: float
    height: float


def cylinder_chamber(n: int = 80, nz: int = 78, radius: float = 0.15, height: float = 0.3,
                     r_target: float = 0.05, z_substrate: float = 0.1, r_substrate: float = 0.1325) -> Chamber:
    """Sputtering chamber: a cylinder (radius, height) of n x n x nz hexes x 6
    tets (80 x 80 x 78 gives 2,995,200 tets and 518,319 nodes) mapped from a box
    by the concentric map.  Surface triangles: every boundary face (walls, the
    top with the target disk r <= r_target, the floor), then the internal faces
    of the layer z = z_substrate within r_substrate (the substrate plate; about
    10,000 triangles at the default sizes).  z_substrate must be a layer of the
    grid (height * k / nz)."""
    box = box_kuhn(n, n, nz, 2.0, 2.0, height, origin=(-1.0, -1.0, 0.0))
    x, y = concentric_disk(box.nodes[:, 0], box.nodes[:, 1])
    nodes = np.stack([radius * x, radius * y, box.nodes[:, 2]], axis=1)
    tets = _orient_positive(nodes, box.tets.astype(np.int64))
    a, b, c, d = (nodes[tets[:, q]] for q in range(4))
    vol = np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a))
    if not (vol > 0).all():
        raise ValueError("concentric map produced a degenerate tet")
    walls = boundary_faces(tets)
    k = int(round(z_substrate / height * nz))
    z_layer = box.nodes[:, 2][((k * (n + 1)) * (n + 1))]  # the grid's own value of that layer
    faces = tets[:, FACES].reshape(-1, 3)
    on_layer = np.all(nodes[faces][:, :, 2] == z_layer, axis=1)
    cen = nodes[faces[on_layer]].mean(axis=1)
    inside = np.hypot(cen[:, 0], cen[:, 1]) <= r_substrate
    sub = faces[on_layer][inside]
    key = np.sort(sub, axis=1)
    _, first = np.unique(key, axis=0, return_index=True)
    sub = sub[np.sort(first)]
    tris = np.concatenate([walls, sub]).astype(np.int32)
    wc = nodes[walls].mean(axis=1)
    target = np.zeros(len(tris), dtype=bool)
    target[: len(walls)] = (np.all(nodes[walls][:, :, 2] == height, axis=1)
                            & (np.hypot(wc[:, 0], wc[:, 1]) <= r_target))
    substrate = np.zeros(len(tris), dtype=bool)
    substrate[len(walls):] = True
    return Chamber(TetMesh(nodes, tets.astype(np.int32), tris), substrate, target, radius, height)


def linear_potential(mesh: TetMesh, v_bottom: float, v_top: float, axis: int = 2) -> np.ndarray:
    """Node potential of two Dirichlet plates at the min/max of `axis`."""
    c = mesh.nodes[:, axis]
    lo, hi = c.min(), c.max()
    return v_bottom + (v_top - v_bottom) * (c - lo) / (hi - lo)
