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
