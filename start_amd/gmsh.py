"""Gmsh 4.1 meshes through the C++ reader (libstart_io.so, include/start_mesh.h).

`read(path)` gives a `GmshMesh`: the `TetMesh` the engine takes (node rows in
ascending tag order, tets and triangles as node rows) and the element tags
the reference uses as tet and triangle ids (VolumetricMesh.cpp:71-80,
GmshUtils.cpp:118-136), so device triangle rows map back to them.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .meshgen import TetMesh

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libstart_io.so")


class _Mesh(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("nodes", C.POINTER(C.c_double)), ("node_tags", C.POINTER(C.c_uint64)),
                ("n_tets", C.c_int64), ("tets", C.POINTER(C.c_int32)), ("tet_tags", C.POINTER(C.c_uint64)),
                ("n_tris", C.c_int64), ("tris", C.POINTER(C.c_int32)), ("tri_tags", C.POINTER(C.c_uint64))]


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run start_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        L.start_gmsh_read.argtypes = [C.c_char_p, C.POINTER(_Mesh)]
        L.start_gmsh_read.restype = C.c_int
        L.start_gmsh_free.argtypes = [C.POINTER(_Mesh)]
        L.start_gmsh_free.restype = None
        L.start_gmsh_last_error.restype = C.c_char_p
        L.start_gmsh_write.argtypes = [C.c_char_p, C.POINTER(_Mesh)]
        L.start_gmsh_write.restype = C.c_int
        L.start_gmsh_write_error.restype = C.c_char_p
        _LIB = L
    return _LIB


@dataclass
class GmshMesh:
    mesh: TetMesh
    node_tags: np.ndarray
    tet_tags: np.ndarray
    tri_tags: np.ndarray


def _arr(p, n, dtype):
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True) if n else np.zeros(0, dtype=dtype)


def write(path: str, mesh: TetMesh, node_tags=None, tet_tags=None, tri_tags=None):
    """The mesh as a Gmsh 4.1 ASCII file (tags default to 1..N, triangles
    1..S, tets S+1..S+T)."""
    nodes = np.ascontiguousarray(mesh.nodes, dtype=np.float64)
    tets = np.ascontiguousarray(mesh.tets, dtype=np.int32)
    tris = np.ascontiguousarray(mesh.tris, dtype=np.int32)
    tags = [None if t is None else np.ascontiguousarray(t, dtype=np.uint64) for t in (node_tags, tet_tags, tri_tags)]
    up = C.POINTER(C.c_uint64)
    tp = [C.cast(None, up) if t is None else t.ctypes.data_as(up) for t in tags]
    m = _Mesh(len(nodes), nodes.ctypes.data_as(C.POINTER(C.c_double)), tp[0],
              len(tets), tets.ctypes.data_as(C.POINTER(C.c_int32)), tp[1],
              len(tris), tris.ctypes.data_as(C.POINTER(C.c_int32)), tp[2])
    L = _lib()
    if L.start_gmsh_write(os.fsencode(path), C.byref(m)) != 0:
        raise RuntimeError(L.start_gmsh_write_error().decode())


def read(path: str) -> GmshMesh:
    L = _lib()
    m = _Mesh()
    if L.start_gmsh_read(os.fsencode(path), C.byref(m)) != 0:
        raise RuntimeError(L.start_gmsh_last_error().decode())
    try:
        nodes = _arr(m.nodes, 3 * m.n_nodes, np.float64).reshape(-1, 3)
        tets = _arr(m.tets, 4 * m.n_tets, np.int32).reshape(-1, 4)
        tris = _arr(m.tris, 3 * m.n_tris, np.int32).reshape(-1, 3)
        out = GmshMesh(TetMesh(nodes, tets, tris), _arr(m.node_tags, m.n_nodes, np.uint64),
                       _arr(m.tet_tags, m.n_tets, np.uint64), _arr(m.tri_tags, m.n_tris, np.uint64))
    finally:
        L.start_gmsh_free(C.byref(m))
    return out
