"""The HDF5 deposition file (libstart_h5.so, include/start_hdf5.h) against the
layout TriangleMeshHdf5Manager writes (TriangleMeshHdf5Manager.cpp:97-139):
round trips through the reader (readMeshFromHDF5, :141-178), the file's
structure and types as the HDF5 tools print them (h5dump, an independent
reader of the format), areas against closed forms, and the error paths.
No file written by the reference exists to compare against, so the byte layout
beyond what h5dump shows is not pinned."""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from start_amd import gmsh, meshgen, outputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

pytestmark = pytest.mark.skipif(not os.path.exists(outputs.H5_LIB_PATH), reason="libstart_h5.so not built")


def test_round_trip_on_a_reference_mesh(tmp_path):
    """The plates mesh of the reference's meshes/: groups named by the
    triangles' element tags, read back in tag order."""
    g = gmsh.read(os.path.join(ROOT, "tests", "golden", "gmsh", "plates.msh"))
    m, tags = g.mesh, g.tri_tags.astype(np.int64)
    rng = np.random.default_rng(7)
    counts = rng.integers(0, 1 << 40, m.n_tris).astype(np.uint64)
    p = str(tmp_path / "plates.hdf5")
    outputs.write_triangle_hdf5(p, m, counts, ids=tags)
    order = np.argsort(tags)
    first = int(tags.min())
    assert np.array_equal(np.sort(tags), np.arange(first, first + m.n_tris)), "tags are consecutive"
    r = outputs.read_triangle_hdf5(p, first)
    assert len(r["count"]) == m.n_tris
    assert np.array_equal(r["count"], counts[order])
    assert np.array_equal(r["coords"], m.nodes[m.tris[order]].reshape(-1, 9))
    assert np.array_equal(r["area"], [outputs.triangle_area(t) for t in r["coords"]])


def test_structure_as_h5dump_prints_it(tmp_path):
    if H5DUMP is None:
        pytest.skip("h5dump not available")
    m = meshgen.box_kuhn(1, 1, 1, 0.1, 0.2, 0.3)
    counts = np.arange(m.n_tris, dtype=np.uint64) * 1000003
    p = str(tmp_path / "box.hdf5")
    outputs.write_triangle_hdf5(p, m, counts, ids=np.arange(5, 5 + m.n_tris))
    txt = subprocess.run([H5DUMP, p], capture_output=True, text=True, check=True).stdout
    groups = re.findall(r'GROUP "(Triangle_\d+)"', txt)
    assert sorted(groups) == sorted(f"Triangle_{5 + k}" for k in range(m.n_tris))
    # per group: three datasets, their types and one-dimensional extents
    blocks = re.split(r'GROUP "Triangle_', txt)[1:]
    for b in blocks:
        k = int(b.split('"', 1)[0]) - 5
        ds = dict(re.findall(r'DATASET "(\w+)" \{\s*DATATYPE\s+(\S+)', b))
        assert ds == {"Area": "H5T_IEEE_F64LE", "Coordinates": "H5T_IEEE_F64LE", "Counter": "H5T_STD_U64LE"}
        dims = dict(re.findall(r'DATASET "(\w+)" \{\s*DATATYPE\s+\S+\s+DATASPACE\s+SIMPLE \{ \( (\d+) \)', b))
        assert dims == {"Area": "1", "Coordinates": "9", "Counter": "1"}
        counter = re.search(r'DATASET "Counter".*?DATA \{\s*\(0\): (\d+)', b, re.S).group(1)
        assert int(counter) == int(counts[k])
        coords = re.search(r'DATASET "Coordinates".*?DATA \{\s*\(0\): ([^}]*)\}', b, re.S).group(1)
        vals = [float(v) for v in re.sub(r"\(\d+\):", "", coords).replace(",", " ").split()]
        assert np.allclose(vals, m.nodes[m.tris[k]].reshape(9), rtol=0, atol=1e-15)


def test_areas_closed_forms(tmp_path):
    assert outputs.triangle_area([0, 0, 0, 3, 0, 0, 0, 4, 0]) == 6.0
    s = 2.0
    eq = [0, 0, 0, s, 0, 0, s / 2, s * np.sqrt(3) / 2, 0]
    assert abs(outputs.triangle_area(eq) - np.sqrt(3) / 4 * s * s) <= 1e-15
    assert outputs.triangle_area([1, 1, 1, 2, 2, 2, 3, 3, 3]) == 0.0  # collinear
    # the closed surface of a box: the areas add up to its surface area
    a, b, c = 0.1, 0.2, 0.3
    m = meshgen.box_kuhn(3, 4, 5, a, b, c)
    p = str(tmp_path / "box.hdf5")
    outputs.write_triangle_hdf5(p, m, np.zeros(m.n_tris, dtype=np.uint64))
    r = outputs.read_triangle_hdf5(p, 1)
    assert abs(r["area"].sum() - 2 * (a * b + b * c + a * c)) <= 1e-15


def test_existing_file_is_replaced_and_empty_maps(tmp_path):
    m = meshgen.box_kuhn(2, 2, 2, 0.1, 0.1, 0.1)
    p = str(tmp_path / "mesh.hdf5")
    outputs.write_triangle_hdf5(p, m, np.ones(m.n_tris, dtype=np.uint64))
    small = meshgen.TetMesh(m.nodes, m.tets, m.tris[:5])
    outputs.write_triangle_hdf5(p, small, np.full(5, 9, dtype=np.uint64))
    r = outputs.read_triangle_hdf5(p, 1)
    assert len(r["count"]) == 5 and (r["count"] == 9).all()
    empty = meshgen.TetMesh(m.nodes, m.tets, m.tris[:0])
    outputs.write_triangle_hdf5(p, empty, np.zeros(0, dtype=np.uint64))
    assert len(outputs.read_triangle_hdf5(p, 1)["count"]) == 0


def test_errors(tmp_path):
    m = meshgen.box_kuhn(1, 1, 1, 0.1, 0.1, 0.1)
    p = str(tmp_path / "e.hdf5")
    with pytest.raises(RuntimeError, match="duplicate id"):
        outputs.write_triangle_hdf5(p, m, np.zeros(m.n_tris, dtype=np.uint64), ids=np.ones(m.n_tris))
    with pytest.raises(RuntimeError, match="failed to open"):
        outputs.read_triangle_hdf5(str(tmp_path / "missing.hdf5"))
    with pytest.raises(RuntimeError, match="failed to create"):
        outputs.write_triangle_hdf5(str(tmp_path / "no" / "dir.hdf5"), m, np.zeros(m.n_tris, dtype=np.uint64))
    outputs.write_triangle_hdf5(p, m, np.zeros(m.n_tris, dtype=np.uint64), ids=np.arange(m.n_tris) * 2)
    with pytest.raises(RuntimeError, match="no group Triangle_1"):
        outputs.read_triangle_hdf5(p, 0)  # the ids are not consecutive


def test_settled_positions_file(tmp_path):
    """SettledParticleHDF5Writer's layout: one n x 3 FP64 dataset."""
    if H5DUMP is None:
        pytest.skip("h5dump not available")
    xyz = np.array([[0.5, -1.25, 3.0], [1e-3, 2.0, -0.0625]])
    p = str(tmp_path / "settled_particles.hdf5")
    assert outputs.write_settled_hdf5(p, xyz)
    txt = subprocess.run([H5DUMP, p], capture_output=True, text=True, check=True).stdout
    assert re.search(r'DATASET "settled_particles" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s+'
                     r'DATASPACE\s+SIMPLE \{ \( 2, 3 \) / \( 2, 3 \) \}', txt), txt
    data = re.search(r"DATA \{(.*?)\}", txt, re.S).group(1)
    vals = [float(v) for v in re.sub(r"\(\d+,\d+\):", "", data).replace(",", " ").split()]
    assert vals == list(xyz.reshape(-1))
    assert outputs.write_settled_hdf5(str(tmp_path / "none.hdf5"), np.zeros((0, 3))) is False
    assert not os.path.exists(tmp_path / "none.hdf5")
