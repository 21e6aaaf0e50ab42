"""The Gmsh 4.1 reader (start_amd/host/GmshReader.cpp) on meshes from the
reference's meshes/ directory (tests/golden/gmsh)."""
import os
from collections import Counter

import numpy as np
import pytest

from start_amd import gmsh

G = os.path.join(os.path.dirname(__file__), "golden", "gmsh")


def volumes(m):
    p = m.nodes[m.tets]
    return np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])) / 6.0


def face_counts(tets):
    f = np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 2, 3]], tets[:, [0, 1, 3]], tets[:, [0, 1, 2]]])
    return Counter(map(tuple, np.sort(f, axis=1)))


def parse_ascii(path):
    """Independent minimal reading of a 4.1 ASCII file: node tag -> xyz and
    the (tag, nodes) of every element of types 2 and 4."""
    lines = [l.strip() for l in open(path)]
    i = lines.index("$Nodes")
    nb = int(lines[i + 1].split()[0])
    j = i + 2
    xyz = {}
    for _ in range(nb):
        dim, _, par, n = map(int, lines[j].split())
        tags = [int(lines[j + 1 + k]) for k in range(n)]
        for k in range(n):
            xyz[tags[k]] = [float(v) for v in lines[j + 1 + n + k].split()[:3]]
        j += 1 + 2 * n
    i = lines.index("$Elements")
    nb = int(lines[i + 1].split()[0])
    j = i + 2
    el = {2: [], 4: []}
    for _ in range(nb):
        _, _, ty, n = map(int, lines[j].split())
        for k in range(n):
            v = list(map(int, lines[j + 1 + k].split()))
            if ty in el:
                el[ty].append(v)
        j += 1 + n
    return xyz, el


@pytest.mark.parametrize("name,volume", [("cube", 125.0), ("two_cubes", 250.0), ("plates", 1250.0)])
def test_volumes_closed_form(name, volume):
    m = gmsh.read(os.path.join(G, name + ".msh")).mesh
    v = volumes(m)
    assert np.all(np.abs(v) > 0)
    assert abs(np.abs(v).sum() - volume) <= 1e-12 * volume


@pytest.mark.parametrize("name", ["cube", "two_cubes", "sphere", "plates"])
def test_triangles_are_faces_and_cover_the_boundary(name):
    """Every triangle is a face of the tet mesh; every boundary face (one
    tet) is a triangle; no face has more than two tets.  (two_cubes: two
    5^3 cubes touching at one corner, which carries a node of each.)"""
    m = gmsh.read(os.path.join(G, name + ".msh")).mesh
    fc = face_counts(m.tets)
    tri = set(map(tuple, np.sort(m.tris, axis=1)))
    assert all(t in fc for t in tri)
    boundary = {f for f, c in fc.items() if c == 1}
    assert boundary <= tri
    assert all(c in (1, 2) for c in fc.values())


@pytest.mark.parametrize("name", ["cube", "sphere"])
def test_against_an_independent_parse(name):
    path = os.path.join(G, name + ".msh")
    g = gmsh.read(path)
    xyz, el = parse_ascii(path)
    assert list(g.node_tags) == sorted(xyz)
    assert np.array_equal(g.mesh.nodes, np.array([xyz[t] for t in sorted(xyz)]))
    row = {t: r for r, t in enumerate(sorted(xyz))}
    assert list(g.tet_tags) == [e[0] for e in el[4]]
    assert np.array_equal(g.mesh.tets, np.array([[row[t] for t in e[1:]] for e in el[4]]))
    assert list(g.tri_tags) == [e[0] for e in el[2]]
    assert np.array_equal(g.mesh.tris, np.array([[row[t] for t in e[1:]] for e in el[2]]))


def test_binary_file_is_a_closed_surface():
    """The binary 4.1 file (surface only): every triangle edge is shared by an
    even number of triangles (2, or 4 where two parts meet along an edge), so
    its 1,702 triangles close a surface (empty boundary mod 2) -- a misread
    node or element breaks that."""
    g = gmsh.read(os.path.join(G, "binary_surface.msh"))
    m = g.mesh
    assert m.n_nodes == 856 and len(m.tris) == 1702 and m.n_tets == 0
    e = np.concatenate([m.tris[:, [0, 1]], m.tris[:, [1, 2]], m.tris[:, [2, 0]]])
    counts = Counter(map(tuple, np.sort(e, axis=1)))
    assert all(c % 2 == 0 for c in counts.values())
    assert sum(c == 2 for c in counts.values()) > 2500
    assert len(np.unique(g.tri_tags)) == 1702
    assert np.all(np.isfinite(m.nodes)) and np.abs(m.nodes).max() < 100.0


def test_writer_round_trips_generated_meshes(tmp_path):
    """Generated box and chamber meshes written as Gmsh 4.1 read back bit for
    bit, and the independent parse of the written file agrees."""
    from start_amd import meshgen
    for name, m in [("box", meshgen.box_kuhn(4, 3, 2, 0.4, 0.3, 0.2)),
                    ("chamber", meshgen.cylinder_chamber(n=8, nz=6).mesh)]:
        p = str(tmp_path / f"{name}.msh")
        gmsh.write(p, m)
        g = gmsh.read(p)
        assert np.array_equal(g.mesh.nodes, m.nodes) and np.array_equal(g.mesh.tets, m.tets)
        assert np.array_equal(g.mesh.tris, m.tris)
        S = len(m.tris)
        assert list(g.tri_tags) == list(range(1, S + 1)) and list(g.tet_tags) == list(range(S + 1, S + 1 + m.n_tets))
        xyz, el = parse_ascii(p)
        assert np.array_equal(np.array([xyz[t] for t in sorted(xyz)]), m.nodes)
        assert len(el[4]) == m.n_tets and len(el[2]) == S


def test_writer_keeps_tags_of_a_reference_mesh(tmp_path):
    g = gmsh.read(os.path.join(G, "sphere.msh"))
    p = str(tmp_path / "sphere_again.msh")
    gmsh.write(p, g.mesh, g.node_tags, g.tet_tags, g.tri_tags)
    h = gmsh.read(p)
    for a, b in [(g.mesh.nodes, h.mesh.nodes), (g.mesh.tets, h.mesh.tets), (g.mesh.tris, h.mesh.tris),
                 (g.node_tags, h.node_tags), (g.tet_tags, h.tet_tags), (g.tri_tags, h.tri_tags)]:
        assert np.array_equal(a, b)


def test_writer_errors(tmp_path):
    from start_amd.meshgen import TetMesh
    bad = TetMesh(np.zeros((3, 3)), np.array([[0, 1, 2, 3]], dtype=np.int32), np.zeros((0, 3), np.int32))
    with pytest.raises(RuntimeError, match="out of range"):
        gmsh.write(str(tmp_path / "bad.msh"), bad)
    with pytest.raises(RuntimeError, match="cannot open"):
        gmsh.write(str(tmp_path / "no_such_dir" / "x.msh"), gmsh.read(os.path.join(G, "cube.msh")).mesh)


REF_MESHES = "/root/reference/meshes"


@pytest.mark.skipif(not os.path.isdir(REF_MESHES), reason="no reference checkout")
def test_every_reference_mesh():
    """Every .msh of the reference checkout reads; tets have non-zero volume,
    no face has more than two tets, every boundary face is a triangle, and
    surface-only files close (every edge shared an even number of times)."""
    names = sorted(f for f in os.listdir(REF_MESHES) if f.endswith(".msh"))
    assert len(names) > 20
    for f in names:
        g = gmsh.read(os.path.join(REF_MESHES, f))
        m = g.mesh
        assert np.array_equal(g.node_tags, np.arange(1, m.n_nodes + 1)), f
        assert len(np.unique(g.tri_tags)) == len(m.tris), f
        if m.n_tets:
            assert np.all(np.abs(volumes(m)) > 0), f
            fc = face_counts(m.tets)
            assert all(c in (1, 2) for c in fc.values()), f
            assert {t for t, c in fc.items() if c == 1} <= set(map(tuple, np.sort(m.tris, axis=1))), f
        else:
            e = np.concatenate([m.tris[:, [0, 1]], m.tris[:, [1, 2]], m.tris[:, [2, 0]]])
            assert all(c % 2 == 0 for c in Counter(map(tuple, np.sort(e, axis=1))).values()), f


def test_errors(tmp_path):
    with pytest.raises(RuntimeError, match="cannot open"):
        gmsh.read(str(tmp_path / "missing.msh"))
    p = tmp_path / "old.msh"
    p.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
    with pytest.raises(RuntimeError, match="4.1"):
        gmsh.read(str(p))
    p = tmp_path / "bad.msh"
    p.write_text("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 1 1 1\n3 1 0 1\n7\n0 0 0\n$EndNodes\n"
                 "$Elements\n1 1 1 1\n2 1 2 1\n1 7 8 9\n$EndElements\n")
    with pytest.raises(RuntimeError, match="unknown node tag"):
        gmsh.read(str(p))


def test_elements_come_back_in_tag_order(tmp_path):
    """A file whose element tags do not ascend in file order (a renumbered or
    merged mesh): tets and triangles come back sorted by tag, with their node
    rows, so the device's smallest-row tie-break is the reference's
    smallest-tag one (NodeChargeDensityProcessor.cpp:51-56)."""
    from start_amd import meshgen
    m = meshgen.box_kuhn(2, 2, 2, 0.2, 0.2, 0.2)
    rng = np.random.default_rng(3)
    tet_tags = rng.permutation(np.arange(1000, 1000 + m.n_tets)).astype(np.uint64)
    tri_tags = rng.permutation(np.arange(1, 1 + m.n_tris)).astype(np.uint64)
    p = str(tmp_path / "permuted.msh")
    gmsh.write(p, m, None, tet_tags, tri_tags)
    g = gmsh.read(p)
    ot, os_ = np.argsort(tet_tags), np.argsort(tri_tags)
    assert np.array_equal(g.tet_tags, tet_tags[ot]) and np.array_equal(g.mesh.tets, m.tets[ot])
    assert np.array_equal(g.tri_tags, tri_tags[os_]) and np.array_equal(g.mesh.tris, m.tris[os_])
    # a repeated tet tag is an error
    dup = tet_tags.copy()
    dup[1] = dup[0]
    gmsh.write(p, m, None, dup, tri_tags)
    with pytest.raises(RuntimeError, match="duplicate tetrahedron tag"):
        gmsh.read(p)
