"""ctypes binding of the CPU oracle (oracle/liboracle.so) — TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker / the timed CPU baseline.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MODEL = {"NONE": 0, "HS": 1, "VHS": 2, "VSS": 3}
ACTIVE, SETTLED = 0, 1


class SpeciesDesc(C.Structure):
    _fields_ = [("mass", C.c_double), ("charge", C.c_double), ("diameter", C.c_double)]


class GasDesc(C.Structure):
    _fields_ = [("mass", C.c_double), ("omega", C.c_double), ("alpha", C.c_double),
                ("concentration", C.c_double), ("model", C.c_int32), ("pad", C.c_int32)]


class SpeciesK(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("mass", "charge", "sigma_hs", "vhs_c", "two_kt", "m_r",
                                          "vhs_exp", "mass_cp", "mass_ct", "inv_alpha")]


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("concentration", C.c_double),
                ("bbox_lo", C.c_double * 3), ("bbox_hi", C.c_double * 3),
                ("seed", C.c_uint64), ("run", C.c_uint32), ("model", C.c_int32),
                ("em_push", C.c_int32), ("nspecies", C.c_int32)]


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: run `make -C {_HERE}`")
        L = C.CDLL(path)
        P = C.POINTER
        d, i32, i64, u32, u64, vp = C.c_double, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_void_p
        L.or_philox4x32_10.argtypes = [P(u32), P(u32), P(u32)]
        L.or_variates.argtypes = [u64, u32, u64, u32, u32, P(d)]
        for f in ("or_orient3d", "or_orient3d_exact"):
            getattr(L, f).argtypes = [P(d)] * 4
            getattr(L, f).restype = C.c_int
        L.or_orient2d_exact.argtypes = [P(d)] * 3
        L.or_orient2d_exact.restype = C.c_int
        L.or_seg_tri_intersect.argtypes = [P(d)] * 5 + [P(d)]
        L.or_seg_tri_intersect.restype = C.c_int
        L.or_gas_concentration.argtypes = [d, d]
        L.or_gas_concentration.restype = d
        L.or_set_libm.argtypes = [C.c_int]
        L.or_get_libm.restype = C.c_int
        L.or_math_eval.argtypes = [C.c_int, C.c_int, i64, P(d), P(d), P(d), P(d)]
        L.or_species_constants.argtypes = [P(SpeciesDesc), P(GasDesc), P(SpeciesK)]
        L.or_surface_bbox.argtypes = [i64, P(d), i64, P(i32), d, P(d), P(d)]
        L.or_collide.argtypes = [C.c_int, P(SpeciesK), d, d, P(d), d, d, d]
        L.or_collide.restype = C.c_int
        L.or_push.argtypes = [P(d), d, d, d, P(d)]
        L.or_mesh_build.argtypes = [i64, P(d), i64, P(i32), i64, P(i32)]
        L.or_mesh_build.restype = vp
        L.or_mesh_free.argtypes = [vp]
        L.or_mesh_tets.argtypes = [vp, P(i32)]
        L.or_locate.argtypes = [vp, P(d)]
        L.or_locate.restype = i32
        L.or_locate_many.argtypes = [vp, i64, P(d), P(d), P(d), P(i32)]
        L.or_first_hit.argtypes = [vp, P(d), P(d)]
        L.or_first_hit.restype = i32
        L.or_tet_field.argtypes = [vp, P(d), P(d)]
        L.or_barycentric.argtypes = [vp, i32, P(d), P(d)]
        L.or_step.argtypes = [vp, P(Params), P(SpeciesK), P(d), i64] + [P(d)] * 6 + \
            [P(i32), P(i32), P(i32), P(i32), P(u64), u32, P(u64), P(d), C.c_int, P(u64)]
        L.or_step.restype = C.c_int
        L.or_step_moved.argtypes = L.or_step.argtypes + [P(C.c_uint8)]
        L.or_step_moved.restype = C.c_int
        L.or_scatter_charge.argtypes = [vp, P(SpeciesK), i64, P(d), P(d), P(d), P(i32), P(i32),
                                        P(i32), P(d)]
        L.or_tet_volumes.argtypes = [vp, P(d)]
        L.or_node_volumes.argtypes = [vp, P(d)]
        L.or_node_density.argtypes = [vp, P(SpeciesK), i64, P(i32), P(i32), P(i32), P(d), P(i32)]
        _LIB = L
    return _LIB


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _vec(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64))


def philox(ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def variates(seed, run, pid, step, draw):
    u = (C.c_double * 2)()
    lib().or_variates(seed, run, pid, step, draw, u)
    return u[0], u[1]


def orient3d(a, b, c, d, exact_only=False) -> int:
    f = lib().or_orient3d_exact if exact_only else lib().or_orient3d
    A, B, Cc, D = (_vec(v) for v in (a, b, c, d))
    return f(_p(A, C.c_double), _p(B, C.c_double), _p(Cc, C.c_double), _p(D, C.c_double))


def orient2d(a, b, c) -> int:
    A, B, Cc = (_vec(v) for v in (a, b, c))
    return lib().or_orient2d_exact(_p(A, C.c_double), _p(B, C.c_double), _p(Cc, C.c_double))


def seg_tri(p, q, a, b, c):
    arr = [_vec(v) for v in (p, q, a, b, c)]
    t = C.c_double(-1.0)
    hit = lib().or_seg_tri_intersect(*[_p(x, C.c_double) for x in arr], C.byref(t))
    return bool(hit), t.value


def set_libm(on: bool) -> None:
    """Collision-path elementary functions: the C library's (True) or the
    shared deterministic ones of start_amd/csrc/detmath.h (False, the default)."""
    lib().or_set_libm(int(bool(on)))


def get_libm() -> bool:
    return bool(lib().or_get_libm())


MATH_FN = {"pow": 0, "sincos": 1, "acos": 2, "atan2": 3, "log": 4, "exp": 5}


def math_eval(fn: str, x, y=None, libm: bool = False):
    """fn over arrays: pow(x, y), sincos(x) -> (sin, cos), acos(x), atan2(y, x), log(x), exp(x)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, dtype=np.float64)
    assert x.shape == y.shape and x.ndim == 1
    out, out2 = np.empty_like(x), np.empty_like(x)
    lib().or_math_eval(MATH_FN[fn], int(libm), len(x), _p(x, C.c_double), _p(y, C.c_double),
                       _p(out, C.c_double), _p(out2, C.c_double))
    return (out, out2) if fn == "sincos" else out


def gas_concentration(P: float, T: float) -> float:
    return lib().or_gas_concentration(P, T)


def species_constants(mass, charge, diameter, gas_mass, omega, alpha, model) -> SpeciesK:
    s = SpeciesDesc(mass, charge, diameter)
    g = GasDesc(gas_mass, omega, alpha, 0.0, model, 0)
    k = SpeciesK()
    lib().or_species_constants(C.byref(s), C.byref(g), C.byref(k))
    return k


def collide(model, k: SpeciesK, n, dt, v, u1, u2, u3):
    vv = (C.c_double * 3)(*v)
    r = lib().or_collide(model, C.byref(k), n, dt, vv, u1, u2, u3)
    return bool(r), np.array(list(vv))


def push(E, q, m, dt, v):
    vv = (C.c_double * 3)(*v)
    EE = (C.c_double * 3)(*E)
    lib().or_push(EE, q, m, dt, vv)
    return np.array(list(vv))


class Mesh:
    def __init__(self, nodes: np.ndarray, tets: np.ndarray, tris: np.ndarray):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        t = np.ascontiguousarray(tets, dtype=np.int32)
        s = np.ascontiguousarray(tris, dtype=np.int32)
        self.n_nodes, self.n_tets, self.n_tris = len(self.nodes), len(t), len(s)
        self.tris = s
        self.h = lib().or_mesh_build(self.n_nodes, _p(self.nodes, C.c_double), self.n_tets,
                                     _p(t, C.c_int32), self.n_tris, _p(s, C.c_int32))
        self.tets = np.empty_like(t)
        lib().or_mesh_tets(self.h, _p(self.tets, C.c_int32))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.or_mesh_free(self.h)
            self.h = None

    def locate(self, p) -> int:
        P = _vec(p)
        return lib().or_locate(self.h, _p(P, C.c_double))

    def locate_many(self, x, y, z) -> np.ndarray:
        x, y, z = _vec(x), _vec(y), _vec(z)
        out = np.empty(len(x), dtype=np.int32)
        lib().or_locate_many(self.h, len(x), _p(x, C.c_double), _p(y, C.c_double),
                             _p(z, C.c_double), _p(out, C.c_int32))
        return out

    def first_hit(self, p, q) -> int:
        P, Q = _vec(p), _vec(q)
        return lib().or_first_hit(self.h, _p(P, C.c_double), _p(Q, C.c_double))

    def tet_field(self, phi) -> np.ndarray:
        phi = _vec(phi)
        E = np.empty((self.n_tets, 3), dtype=np.float64)
        lib().or_tet_field(self.h, _p(phi, C.c_double), _p(E, C.c_double))
        return E

    def barycentric(self, t, p) -> np.ndarray:
        P = _vec(p)
        lam = np.empty(4, dtype=np.float64)
        lib().or_barycentric(self.h, int(t), _p(P, C.c_double), _p(lam, C.c_double))
        return lam

    def tet_volumes(self) -> np.ndarray:
        v = np.empty(self.n_tets)
        lib().or_tet_volumes(self.h, _p(v, C.c_double))
        return v

    def node_volumes(self) -> np.ndarray:
        v = np.empty(self.n_nodes)
        lib().or_node_volumes(self.h, _p(v, C.c_double))
        return v

    def surface_bbox(self, expansion=1.01):
        lo = np.empty(3)
        hi = np.empty(3)
        lib().or_surface_bbox(self.n_nodes, _p(self.nodes, C.c_double), self.n_tris,
                              _p(self.tris, C.c_int32), expansion, _p(lo, C.c_double),
                              _p(hi, C.c_double))
        return lo, hi


@dataclass
class State:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    tet: np.ndarray
    status: np.ndarray
    hit: np.ndarray
    species: np.ndarray
    pid: np.ndarray

    @classmethod
    def new(cls, pos, vel, species, pid, tet):
        n = len(pos)
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64).copy()
        return cls(f(pos[:, 0]), f(pos[:, 1]), f(pos[:, 2]), f(vel[:, 0]), f(vel[:, 1]), f(vel[:, 2]),
                   np.ascontiguousarray(tet, dtype=np.int32).copy(),
                   np.zeros(n, dtype=np.int32), np.full(n, -1, dtype=np.int32),
                   np.ascontiguousarray(species, dtype=np.int32),
                   np.ascontiguousarray(pid, dtype=np.uint64))

    def copy(self) -> "State":
        return State(*[getattr(self, f).copy() for f in self.__dataclass_fields__])


class Sim:
    """The oracle's step loop over one particle population."""

    def __init__(self, mesh: Mesh, species_k: list, model: int, concentration: float, dt: float,
                 seed: int, run: int = 0, em_push: bool = True, E=None, bbox=None):
        self.mesh = mesh
        self.sp = (SpeciesK * len(species_k))(*species_k)
        self.E = _vec(E if E is not None else np.zeros((mesh.n_tets, 3))).reshape(-1)
        lo, hi = bbox if bbox is not None else mesh.surface_bbox()
        self.prm = Params(dt, concentration, (C.c_double * 3)(*lo), (C.c_double * 3)(*hi), seed, run,
                          model, int(em_push), len(species_k))
        self.tri_count = np.zeros(max(mesh.n_tris, 1), dtype=np.uint64)

    def step(self, st: State, step: int, node_charge: np.ndarray | None = None, nthreads: int = 1,
             moved: np.ndarray | None = None) -> int:
        """One step; with `moved` (uint8, one per particle) also the particles
        whose start position the step records as a movement."""
        processed = C.c_uint64(0)
        nc = _p(node_charge, C.c_double) if node_charge is not None else None
        mv = _p(moved, C.c_uint8) if moved is not None else None
        lib().or_step_moved(self.mesh.h, C.byref(self.prm), self.sp, _p(self.E, C.c_double), len(st.x),
                            _p(st.x, C.c_double), _p(st.y, C.c_double), _p(st.z, C.c_double),
                            _p(st.vx, C.c_double), _p(st.vy, C.c_double), _p(st.vz, C.c_double),
                            _p(st.tet, C.c_int32), _p(st.status, C.c_int32), _p(st.hit, C.c_int32),
                            _p(st.species, C.c_int32), _p(st.pid, C.c_uint64), step,
                            _p(self.tri_count, C.c_uint64), nc, nthreads, C.byref(processed), mv)
        return processed.value

    def scatter_charge(self, st: State, node_charge: np.ndarray):
        lib().or_scatter_charge(self.mesh.h, self.sp, len(st.x), _p(st.x, C.c_double),
                                _p(st.y, C.c_double), _p(st.z, C.c_double), _p(st.tet, C.c_int32),
                                _p(st.status, C.c_int32), _p(st.species, C.c_int32),
                                _p(node_charge, C.c_double))

    def node_density(self, st: State):
        """Tet-average node charge density of the current owners; returns (rho, written)."""
        rho = np.empty(self.mesh.n_nodes)
        wr = np.empty(self.mesh.n_nodes, dtype=np.int32)
        lib().or_node_density(self.mesh.h, self.sp, len(st.x), _p(st.tet, C.c_int32),
                              _p(st.status, C.c_int32), _p(st.species, C.c_int32),
                              _p(rho, C.c_double), _p(wr, C.c_int32))
        return rho, wr.astype(bool)
