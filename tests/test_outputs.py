"""Output writers (start_amd/host/Outputs.cpp): movement JSON, deposition
histogram and KDE file, against the reference's own output file
(tests/golden/outputs/kde.dat), JSON's published layout and closed forms."""
import json
import os

import numpy as np
import pytest

from start_amd import gmsh, outputs
from start_amd.meshgen import TetMesh

G = os.path.join(os.path.dirname(__file__), "golden", "outputs")


# ---- movements JSON ----------------------------------------------------------
def read_text(p):
    with open(p) as f:
        return f.read()


def test_movements_json_layout_matches_an_indented_json_dump(tmp_path):
    """Numbers of moderate size print as Python's repr does (both the shortest
    round-trip digits), so the whole file equals json.dumps(indent=4) of the
    same object with keys in string order ("10" < "2")."""
    rng = np.random.default_rng(3)
    mv = {}
    for pid in [2, 10, 7, 123456789012, 1, 0]:
        k = int(rng.integers(0, 6))
        mv[pid] = rng.normal(0.0, 0.05, size=(k, 3)) + 0.01
    p = tmp_path / "m.json"
    outputs.write_movements_json(str(p), mv)
    expect = {str(k): [{"x": float(a), "y": float(b), "z": float(c)} for a, b, c in v]
              for k, v in mv.items() if len(v) > 1}
    text = read_text(p)
    assert text == json.dumps(expect, indent=4, sort_keys=True)
    back = json.loads(text)
    for k, v in back.items():  # exact round trip
        assert np.array_equal(np.array([[q["x"], q["y"], q["z"]] for q in v]), mv[int(k)])


@pytest.mark.parametrize("v,text", [
    (1.0, "1.0"), (-2.5, "-2.5"), (0.0, "0.0"), (0.1, "0.1"), (100.0, "100.0"),
    (0.0001, "0.0001"), (1e-05, "1e-05"), (1.5e-10, "1.5e-10"), (-3.25e-7, "-3.25e-07"),
    (1e14, "100000000000000.0"), (1e15, "1e+15"), (1.25e16, "1.25e+16"), (1e100, "1e+100"),
    (0.30000000000000004, "0.30000000000000004"), (123.456, "123.456"),
])
def test_json_number_text(tmp_path, v, text):
    """nlohmann::json dump(): plain notation while the decimal point is within
    15 digits and no more than 3 zeros follow it (-4 < n <= 15), a two-digit
    minimum exponent otherwise; ".0" on whole numbers."""
    p = tmp_path / "n.json"
    outputs.write_movements_json(str(p), {5: np.array([[v, 0.0, 1.0], [0.5, 0.5, 0.5]])})
    first = read_text(p).splitlines()[3]
    assert first == f'            "x": {text},'
    assert json.loads(read_text(p))["5"][0]["x"] == v


def test_movements_json_edge_cases(tmp_path):
    p = tmp_path / "e.json"
    outputs.write_movements_json(str(p), {})
    assert not p.exists()  # an empty map writes nothing
    outputs.write_movements_json(str(p), {3: np.zeros((1, 3))})
    assert read_text(p) == "null"  # no particle with more than one point


# ---- band overlap and histogram ----------------------------------------------
def tri(*v):
    return np.array([[x, y, 0.0] for x, y in v], dtype=np.float64)


def width_integral(t, a, b, n=20001):
    """Area of the triangle's xy projection between x = a and x = b by
    Simpson's rule on its vertical width (independent of polygon clipping)."""
    P = t[:, :2]
    xs = np.linspace(a, b, n)
    w = np.zeros(n)
    for i, x in enumerate(xs):
        ys = []
        for j in range(3):
            p, q = P[j], P[(j + 1) % 3]
            if (p[0] - x) * (q[0] - x) <= 0 and p[0] != q[0]:
                ys.append(p[1] + (q[1] - p[1]) * (x - p[0]) / (q[0] - p[0]))
        w[i] = max(ys) - min(ys) if len(ys) >= 2 else 0.0
    h = (b - a) / (n - 1)
    return h / 3 * (w[0] + w[-1] + 4 * w[1:-1:2].sum() + 2 * w[2:-1:2].sum())


def test_band_overlap_closed_forms():
    assert outputs.band_overlap(tri((48.5, 0), (51, 0), (49, 3)), 48, 52) == 1.0
    assert outputs.band_overlap(tri((10, 0), (11, 0), (10, 1)), 48, 52) == 0.0
    # right triangle (46,0),(54,0),(46,8): 16 of its 32 units lie in 48..52
    assert abs(outputs.band_overlap(tri((46, 0), (54, 0), (46, 8)), 48, 52) - 0.5) < 1e-15
    # degenerate projection (a vertical triangle) counts 0
    assert outputs.band_overlap(np.array([[50, 0, 0], [50, 1, 0], [50, 0, 1.0]]), 48, 52) == 0.0


def test_band_overlap_against_integration():
    rng = np.random.default_rng(11)
    for _ in range(40):
        t = np.c_[rng.uniform(45, 55, (3, 1)), rng.uniform(0, 5, (3, 1)), np.zeros((3, 1))]
        P = t[:, :2]
        area = abs((P[1, 0] - P[0, 0]) * (P[2, 1] - P[0, 1]) - (P[2, 0] - P[0, 0]) * (P[1, 1] - P[0, 1])) / 2
        if area < 0.5:
            continue
        ref = width_integral(t, 48.0, 52.0) / area
        assert abs(outputs.band_overlap(t, 48.0, 52.0) - ref) < 1e-6


def test_histogram_bins_and_file(tmp_path):
    """Triangles fully in the band add their counts to floor(y) bins; one
    75% inside counts 0.75 of its deposits; 50% inside and centroid-outside
    triangles are left out; the reference centroid (v0 + v1 + v0) / 3 decides
    band and bin."""
    tris = [
        tri((49, 0.2), (50, 0.2), (49, 0.7)),      # bin 0, overlap 1, count 5
        tri((49, 2.1), (50, 2.1), (49.5, 2.9)),    # bin 2, overlap 1, count 7
        tri((46, 3.1), (54, 3.1), (46, 3.5)),      # 50% inside: excluded
        tri((60, 1.0), (61, 1.0), (60, 1.5)),      # outside the band
        # (50,4),(54,4),(50,4.5): 1/2 * 4 * 0.5 = 1; part with x <= 52 is 0.75 of it
        tri((50, 4.0), (54, 4.0), (50, 4.5)),
        # centroid in the band, but 1.5 of its 3.92 units lie beyond x = 52: excluded
        tri((48.2, 1.1), (58, 1.1), (48.4, 1.9)),
    ]
    nodes = np.concatenate(tris)
    m = TetMesh(nodes, np.zeros((0, 4), np.int32), np.arange(len(nodes), dtype=np.int32).reshape(-1, 3))
    counts = np.array([5, 7, 100, 1000, 8, 50], dtype=np.uint64)
    assert abs(outputs.band_overlap(tris[4], 48, 52) - 0.75) < 1e-15
    y, tot = outputs.deposition_histogram(m, counts)
    assert list(y) == [0.0, 2.0, 4.0]
    assert tot[0] == 5.0 and tot[1] == 7.0 and abs(tot[2] - 6.0) < 1e-12
    p = tmp_path / "h.dat"
    outputs.write_histogram(str(p), m, counts)
    assert read_text(p) == "0 5\n2 7\n4 %g\n" % tot[2]


# ---- KDE file against the reference's own output ------------------------------
def test_kde_rows_are_the_reference_centroids(tmp_path):
    """Every (x, y) of the reference's results/kde.dat is the text of a
    reference centroid of a triangle of its plates mesh; with the true centroid
    none is."""
    g = gmsh.read(os.path.join(G, "two_plates.msh"))
    m = g.mesh
    ones = np.ones(len(m.tris), dtype=np.uint64)
    ref_rows = read_text(os.path.join(G, "kde.dat")).splitlines()
    assert ref_rows[0] == "# X-coordinate(cm) Y-coordinate(cm) Count"
    ref_xy = {" ".join(r.split()[:2]) for r in ref_rows[1:]}
    assert len(ref_rows) == 4903
    for flag, expect_all in [(True, True), (False, False)]:
        p = tmp_path / f"k{int(flag)}.dat"
        outputs.write_kde(str(p), m, ones, reference_centroid=flag)
        rows = read_text(p).splitlines()
        assert rows[0] == ref_rows[0] and len(rows) == len(m.tris) + 1
        mine = {" ".join(r.split()[:2]) for r in rows[1:]}
        hits = len(ref_xy & mine)
        if expect_all:
            assert hits == len(ref_xy)
        else:
            assert hits == 0
    # rows carry the count, and triangles with none are left out
    counts = np.zeros(len(m.tris), dtype=np.uint64)
    counts[[0, 5, 9]] = [3, 1, 12]
    p = tmp_path / "k.dat"
    outputs.write_kde(str(p), m, counts)
    rows = read_text(p).splitlines()[1:]
    assert [r.split()[2] for r in rows] == ["3", "1", "12"]
