"""ctypes binding of libstart_gpu.so (C ABI in include/start_gpu.h).

The library is built in-tree by start_amd/build.py.  Loading fails loudly if it
is missing: there is no CPU fallback for the particle step.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstart_gpu.so")

START_OK = 0
START_ERR_NOMEM = -6  # device allocation failed (include/start_gpu.h)
MODEL = {"NONE": 0, "HS": 1, "VHS": 2, "VSS": 3}
PREC_FP64, PREC_MIXED = 0, 1
ACTIVE, SETTLED = 0, 1

# every entry point declared in include/start_gpu.h
EXPORTS = [
    "start_gas_concentration", "start_create", "start_destroy", "start_last_error",
    "start_load_particles", "start_set_tet_field", "start_set_node_potential", "start_step",
    "start_gather_charge", "start_read_node_charge", "start_read_particles",
    "start_read_tri_counts", "start_read_settled", "start_read_stats", "start_read_stats_settled", "start_synchronize",
    "start_locate", "start_orient3d_batch", "start_math_batch", "start_timing_enable", "start_timing_read",
    "start_node_charge_density", "start_node_charge_density_view", "start_set_gas", "start_set_time_step",
    "start_load_batch", "start_set_tet_field_basis", "start_set_node_potential_basis",
    "start_read_node_charge_runs", "start_read_tri_counts_runs", "start_set_launch", "start_get_launch",
    "start_autotune", "start_autotune_cached", "start_record_settle", "start_read_hits", "start_adjoint_push",
    "start_read_tet_field", "start_generate", "start_track_movements", "start_read_movements",
    "start_clear_movements", "start_locate_noalloc", "start_append_particles", "start_mesh_maps", "start_timing_read_each", "start_read_settled_since",
]
NOALLOC_LOCATE_MAX = 131072
SOURCE = {"point": 0, "surface": 1}
ENERGY = {"mono": 0, "maxwellian": 1, "thompson": 2, "lognormal": 3}
ANGLE = {"fan": 0, "cosine": 1, "isotropic": 2, "cone": 3}
DENSITY = {"barycentric": 0, "tet_average": 1}


class MeshDesc(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("nodes", C.POINTER(C.c_double)),
                ("n_tets", C.c_int64), ("tets", C.POINTER(C.c_int32)),
                ("n_tris", C.c_int64), ("tris", C.POINTER(C.c_int32))]


class SpeciesDesc(C.Structure):
    _fields_ = [("mass", C.c_double), ("charge", C.c_double), ("diameter", C.c_double)]


class GasDesc(C.Structure):
    _fields_ = [("mass", C.c_double), ("omega", C.c_double), ("alpha", C.c_double),
                ("concentration", C.c_double), ("model", C.c_int32), ("pad", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("bbox_expansion", C.c_double), ("seed", C.c_uint64),
                ("run", C.c_uint32), ("em_push", C.c_int32), ("precision", C.c_int32),
                ("charge_every_step", C.c_int32), ("n_species", C.c_int32),
                ("steps_per_launch", C.c_int32), ("species", C.POINTER(SpeciesDesc)),
                ("gas", GasDesc), ("sort_interval", C.c_int32), ("reserved", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("step", C.c_uint64), ("processed", C.c_uint64), ("settled", C.c_uint64),
                ("n_particles", C.c_uint64), ("deferred", C.c_uint64),
                ("exhaustive", C.c_uint64), ("stop", C.c_int32), ("pad", C.c_int32)]


class RunDesc(C.Structure):
    _fields_ = [("n_particles", C.c_int64), ("field_scale", C.c_double * 2), ("concentration", C.c_double),
                ("run", C.c_uint32), ("pad", C.c_int32)]


class SourceDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("species", C.c_int32), ("count", C.c_int64),
                ("energy", C.c_int32), ("angle", C.c_int32), ("energy_eV", C.c_double),
                ("p0", C.c_double), ("p1", C.c_double), ("position", C.c_double * 3),
                ("theta", C.c_double), ("phi", C.c_double), ("expansion", C.c_double),
                ("displace", C.c_double), ("centres", C.POINTER(C.c_double)),
                ("normals", C.POINTER(C.c_double)), ("n_cells", C.c_int64)]


class LaunchCfg(C.Structure):
    _fields_ = [("steps_per_launch", C.c_int32), ("block", C.c_int32), ("persistent", C.c_int32),
                ("sort_interval", C.c_int32), ("graph", C.c_int32), ("walk", C.c_int32)]


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python start_amd/build.py` "
                           "(there is no CPU fallback for the particle step)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    d, i32, i64, u64, vp = C.c_double, C.c_int32, C.c_int64, C.c_uint64, C.c_void_p
    L.start_gas_concentration.argtypes = [d, d]
    L.start_gas_concentration.restype = d
    L.start_create.argtypes = [P(vp), P(MeshDesc), P(Params)]
    L.start_destroy.argtypes = [vp]
    L.start_destroy.restype = None
    L.start_last_error.argtypes = [vp]
    L.start_last_error.restype = C.c_char_p
    L.start_load_particles.argtypes = [vp, i64] + [P(d)] * 6 + [P(i32), P(u64)]
    L.start_append_particles.argtypes = [vp, i64] + [P(d)] * 6 + [P(i32), P(u64)]
    L.start_set_tet_field.argtypes = [vp, P(d)]
    L.start_set_node_potential.argtypes = [vp, P(d)]
    L.start_step.argtypes = [vp, i32]
    L.start_gather_charge.argtypes = [vp]
    L.start_read_node_charge.argtypes = [vp, P(d)]
    L.start_read_particles.argtypes = [vp] + [P(d)] * 6 + [P(i32)] * 3
    L.start_read_tri_counts.argtypes = [vp, P(u64)]
    L.start_read_settled.argtypes = [vp, P(u64), P(i64)]
    L.start_read_settled_since.argtypes = [vp, i64, P(u64), P(i64)]
    L.start_read_stats.argtypes = [vp, P(Stats)]
    L.start_read_stats_settled.argtypes = [vp, P(Stats), i64, P(u64), P(i64)]
    L.start_synchronize.argtypes = [vp]
    L.start_locate.argtypes = [vp, i64, P(d), P(d), P(d), P(i32)]
    L.start_locate_noalloc.argtypes = [vp, i64, P(d), P(d), P(d), P(i32)]
    L.start_orient3d_batch.argtypes = [i64, P(d), P(i32)]
    L.start_math_batch.argtypes = [i32, i64, P(d), P(d), P(d), P(d)]
    L.start_mesh_maps.argtypes = [P(MeshDesc), P(i32), P(d), P(d), P(d), P(C.c_float), P(d)]
    L.start_timing_enable.argtypes = [vp, i32]
    L.start_timing_read.argtypes = [vp, P(d), P(u64)]
    L.start_timing_read_each.argtypes = [vp, P(d), P(i64)]
    L.start_node_charge_density.argtypes = [vp, i32, P(d), P(i32)]
    L.start_node_charge_density_view.argtypes = [vp, i32, P(P(d)), P(P(i32))]
    L.start_set_gas.argtypes = [vp, P(GasDesc), P(d)]
    L.start_set_time_step.argtypes = [vp, d]
    L.start_load_batch.argtypes = [vp, i32, P(RunDesc)] + [P(d)] * 6 + [P(i32), P(u64)]
    L.start_set_tet_field_basis.argtypes = [vp, i32, P(d)]
    L.start_read_tet_field.argtypes = [vp, i32, P(d)]
    L.start_generate.argtypes = [vp, i32, P(SourceDesc), u64]
    L.start_set_node_potential_basis.argtypes = [vp, i32, P(d)]
    L.start_read_node_charge_runs.argtypes = [vp, P(d)]
    L.start_read_tri_counts_runs.argtypes = [vp, P(u64)]
    L.start_set_launch.argtypes = [vp, P(LaunchCfg)]
    L.start_get_launch.argtypes = [vp, P(LaunchCfg)]
    L.start_autotune.argtypes = [vp, i32, P(LaunchCfg), P(d)]
    L.start_autotune_cached.argtypes = [vp, C.c_char_p, i32, P(LaunchCfg), P(d), P(i32)]
    L.start_record_settle.argtypes = [vp, i32]
    L.start_read_hits.argtypes = [vp, P(d), P(i32), P(i32)]
    L.start_adjoint_push.argtypes = [vp, P(d), i32, P(d)]
    L.start_track_movements.argtypes = [vp, i64, i32]
    L.start_read_movements.argtypes = [vp, P(i64), P(u64), P(i32), P(d)]
    L.start_clear_movements.argtypes = [vp]
    for name in EXPORTS:
        f = getattr(L, name)
        if f.restype is C.c_int:  # default
            f.restype = C.c_int
    _LIB = L
    return L


class StartError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc: int, ctx=None):
    if rc != START_OK:
        msg = lib().start_last_error(ctx)
        raise StartError(rc, msg.decode() if msg else "")
