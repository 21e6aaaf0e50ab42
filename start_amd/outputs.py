"""Output files of a run through the C++ writers (libstart_io.so,
include/start_output.h): particle movements as JSON
(ParticleMovementTracker.cpp:18-69), the deposition histogram along a band and
the per-triangle deposition file (SputteringModel.cpp:183-252).

`reference_centroid=True` bins and prints triangles at (v0 + v1 + v0) / 3, the
centroid the reference computes (TriangleCell.cpp:11-14); False uses the true
centroid.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .gmsh import LIB_PATH

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run start_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        d, i32, i64, u64 = C.c_double, C.c_int32, C.c_int64, C.c_uint64
        P = C.POINTER
        L.start_write_movements_json.argtypes = [C.c_char_p, i64, P(u64), P(i32), i32, P(d)]
        L.start_deposition_histogram.argtypes = [i64, P(d), P(u64), d, d, d, i32, P(i64), P(d), P(d)]
        L.start_write_histogram.argtypes = [C.c_char_p, i64, P(d), P(u64), d, d, d, i32]
        L.start_band_overlap.argtypes = [P(d), d, d]
        L.start_band_overlap.restype = d
        L.start_write_kde.argtypes = [C.c_char_p, i64, P(d), P(u64), i32]
        L.start_output_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise RuntimeError(_lib().start_output_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _tris(mesh) -> np.ndarray:
    return np.ascontiguousarray(mesh.nodes[mesh.tris].reshape(-1, 9), dtype=np.float64)


def write_movements_json(path: str, movements: dict):
    """movements: {particle id: (k, 3) points}, as ParticleInCellStep.movements() gives."""
    ids = np.array(list(movements.keys()), dtype=np.uint64)
    counts = np.array([len(v) for v in movements.values()], dtype=np.int32)
    kmax = max(int(counts.max()) if len(counts) else 1, 1)
    xyz = np.zeros((max(len(ids), 1), kmax, 3))
    for s, v in enumerate(movements.values()):
        xyz[s, : len(v)] = v
    _check(_lib().start_write_movements_json(os.fsencode(path), len(ids), _p(ids, C.c_uint64),
                                             _p(counts, C.c_int32), kmax, _p(xyz, C.c_double)))


def band_overlap(tri_xyz, x_min: float, x_max: float) -> float:
    t = np.ascontiguousarray(tri_xyz, dtype=np.float64).reshape(9)
    return float(_lib().start_band_overlap(_p(t, C.c_double), x_min, x_max))


def deposition_histogram(mesh, counts, x_min=48.0, x_max=52.0, bin=1.0, reference_centroid=True):
    """(bin y, total) in ascending y; defaults are the reference's band."""
    t = _tris(mesh)
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    L = _lib()
    nb = C.c_int64(0)
    _check(L.start_deposition_histogram(len(c), _p(t, C.c_double), _p(c, C.c_uint64), x_min, x_max, bin,
                                        int(reference_centroid), C.byref(nb), None, None))
    y = np.empty(max(nb.value, 1))
    tot = np.empty(max(nb.value, 1))
    _check(L.start_deposition_histogram(len(c), _p(t, C.c_double), _p(c, C.c_uint64), x_min, x_max, bin,
                                        int(reference_centroid), C.byref(nb), _p(y, C.c_double),
                                        _p(tot, C.c_double)))
    return y[: nb.value], tot[: nb.value]


def write_histogram(path, mesh, counts, x_min=48.0, x_max=52.0, bin=1.0, reference_centroid=True):
    t = _tris(mesh)
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    _check(_lib().start_write_histogram(os.fsencode(path), len(c), _p(t, C.c_double), _p(c, C.c_uint64), x_min,
                                        x_max, bin, int(reference_centroid)))


def write_kde(path, mesh, counts, reference_centroid=True):
    t = _tris(mesh)
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    _check(_lib().start_write_kde(os.fsencode(path), len(c), _p(t, C.c_double), _p(c, C.c_uint64),
                                  int(reference_centroid)))


# ---------------------------------------------------------------------------
# HDF5 deposition file (libstart_h5.so, include/start_hdf5.h)
# ---------------------------------------------------------------------------
H5_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libstart_h5.so")
_H5 = None


def _h5():
    global _H5
    if _H5 is None:
        if not os.path.exists(H5_LIB_PATH):
            raise RuntimeError(f"{H5_LIB_PATH} is not built (the HDF5 C library was not found at build time)")
        L = C.CDLL(H5_LIB_PATH)
        d, i32, i64, u64 = C.c_double, C.c_int32, C.c_int64, C.c_uint64
        P = C.POINTER
        L.start_write_triangle_hdf5.argtypes = [C.c_char_p, i64, P(i64), P(d), P(i32), P(u64)]
        L.start_read_triangle_hdf5.argtypes = [C.c_char_p, i64, P(i64), P(d), P(d), P(u64)]
        L.start_write_settled_hdf5.argtypes = [C.c_char_p, i64, P(d), C.c_char_p]
        L.start_triangle_area.argtypes = [P(d)]
        L.start_triangle_area.restype = d
        L.start_hdf5_last_error.restype = C.c_char_p
        _H5 = L
    return _H5


def _h5check(rc):
    if rc != 0:
        raise RuntimeError(_h5().start_hdf5_last_error().decode())


def write_triangle_hdf5(path: str, mesh, counts: np.ndarray, ids=None):
    """Triangle_<id>/{Coordinates, Area, Counter} for every surface triangle,
    as TriangleMeshHdf5Manager::saveMeshToHDF5 writes them
    (TriangleMeshHdf5Manager.cpp:97-139).  ids: the triangles' element tags
    (default: row + 1); counts: settled particles per triangle."""
    n = mesh.n_tris
    ids = np.arange(1, n + 1, dtype=np.int64) if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
    counts = np.ascontiguousarray(counts, dtype=np.uint64)
    if len(ids) != n or len(counts) != n:
        raise ValueError("one id and one count per triangle")
    nodes = np.ascontiguousarray(mesh.nodes, dtype=np.float64)
    tris = np.ascontiguousarray(mesh.tris, dtype=np.int32)
    _h5check(_h5().start_write_triangle_hdf5(path.encode(), n, _p(ids, C.c_int64), _p(nodes, C.c_double),
                                             _p(tris, C.c_int32), _p(counts, C.c_uint64)))


def read_triangle_hdf5(path: str, first_id: int = 1) -> dict:
    """The groups Triangle_<first_id>.. of a file, as readMeshFromHDF5 reads
    them (TriangleMeshHdf5Manager.cpp:141-178): {"coords": (n, 9), "area": (n,),
    "count": (n,)}."""
    L = _h5()
    n = C.c_int64(0)
    _h5check(L.start_read_triangle_hdf5(path.encode(), int(first_id), C.byref(n), None, None, None))
    m = n.value
    coords = np.zeros((max(m, 1), 9))
    area = np.zeros(max(m, 1))
    count = np.zeros(max(m, 1), dtype=np.uint64)
    n = C.c_int64(max(m, 1))
    _h5check(L.start_read_triangle_hdf5(path.encode(), int(first_id), C.byref(n), _p(coords, C.c_double),
                                        _p(area, C.c_double), _p(count, C.c_uint64)))
    return {"coords": coords[:m], "area": area[:m], "count": count[:m]}


def write_settled_hdf5(path: str, xyz: np.ndarray, dataset: str = "settled_particles") -> bool:
    """Settled positions as one n x 3 dataset, as
    SettledParticleHDF5Writer::saveParticlesToHDF5 writes them; False (and no
    file) for no particles."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    rc = _h5().start_write_settled_hdf5(path.encode(), len(xyz), _p(xyz, C.c_double), dataset.encode())
    if rc == 1:
        return False
    _h5check(rc)
    return True


def triangle_area(tri_xyz: np.ndarray) -> float:
    t = np.ascontiguousarray(tri_xyz, dtype=np.float64).reshape(9)
    return float(_h5().start_triangle_area(_p(t, C.c_double)))
