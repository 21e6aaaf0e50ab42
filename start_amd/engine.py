"""Host-side mirror of the reference step-loop interface over the C ABI.

Start's driver calls, once per step (ModelingMainDriver.cpp:146-184):
    NodeChargeDensityProcessor::gather(...)   relocation + node charge
    ChargeDensityEquationSolver::solve(...)   FEM (stays on Trilinos)
    ParticleDynamicsProcessor::process(...)   push/move/collide/hit/settle
`ParticleInCellStep` keeps one device context with the particles resident in
HBM; `gather()` and `process()` keep the reference's names and meaning.
Errors raise `StartError` (the reference throws from gather,
NodeChargeDensityProcessor.cpp:103-112, and only logs from process,
ParticleDynamicsProcessor.cpp:202-209; the C ABI reports both as status codes).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .meshgen import TetMesh
from .species import BBOX_EXPANSION, Gas, Species


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


class ParticleInCellStep:
    """One simulation context on the GPU."""

    def __init__(self, mesh: TetMesh, species: list[Species], gas: Gas, dt: float, seed: int,
                 run: int = 0, em_push: bool = True, precision: str = "fp64",
                 charge_every_step: bool = True, bbox_expansion: float = BBOX_EXPANSION,
                 steps_per_launch: int = 1, sort_interval: int = 20, block: int = 256,
                 persistent: bool = False, graph: bool = False):
        L = _lib.lib()
        self.mesh = mesh
        self._nodes = np.ascontiguousarray(mesh.nodes, dtype=np.float64)
        self._tets = np.ascontiguousarray(mesh.tets, dtype=np.int32)
        self._tris = np.ascontiguousarray(mesh.tris, dtype=np.int32)
        md = _lib.MeshDesc(len(self._nodes), _dp(self._nodes), len(self._tets), _ip(self._tets),
                           len(self._tris), _ip(self._tris))
        sp = (_lib.SpeciesDesc * len(species))(*[_lib.SpeciesDesc(s.mass, s.charge, s.diameter)
                                                 for s in species])
        self._sp = sp
        g = _lib.GasDesc(gas.mass, gas.omega, gas.alpha, gas.concentration, _lib.MODEL[gas.model], 0)
        prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
        prm = _lib.Params(dt, bbox_expansion, seed, run, int(em_push), prec, int(charge_every_step),
                          len(species), steps_per_launch, sp, g, int(sort_interval), 0)
        self.ctx = C.c_void_p()
        _lib.check(L.start_create(C.byref(self.ctx), C.byref(md), C.byref(prm)), None)
        self.n = 0
        self.n_runs = 1
        if block != 256 or persistent or graph:
            self.set_launch(steps_per_launch=steps_per_launch, block=block, persistent=persistent,
                            sort_interval=sort_interval, graph=graph)

    def close(self):
        if self.ctx:
            _lib.lib().start_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        _lib.check(rc, self.ctx)

    # --- inputs -----------------------------------------------------------------
    def load_particles(self, pos: np.ndarray, vel: np.ndarray, species: np.ndarray, ids=None):
        pos = np.asarray(pos, dtype=np.float64)
        vel = np.asarray(vel, dtype=np.float64)
        cols = [np.ascontiguousarray(a) for a in (pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0],
                                                  vel[:, 1], vel[:, 2])]
        sp = np.ascontiguousarray(species, dtype=np.int32)
        idp = None
        if ids is not None:
            self._ids = np.ascontiguousarray(ids, dtype=np.uint64)
            idp = _up(self._ids)
        self._check(_lib.lib().start_load_particles(self.ctx, len(pos), *[_dp(c) for c in cols],
                                                    _ip(sp), idp))
        self.n = len(pos)
        self.n_runs = 1

    def append_particles(self, pos: np.ndarray, vel: np.ndarray, species: np.ndarray, ids=None):
        """Appends particles after the loaded ones, keeping everything loaded
        (the continuous-sputtering spawn, ModelingMainDriver.cpp:141-142)."""
        pos = np.asarray(pos, dtype=np.float64)
        vel = np.asarray(vel, dtype=np.float64)
        cols = [np.ascontiguousarray(a) for a in (pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0],
                                                  vel[:, 1], vel[:, 2])]
        sp = np.ascontiguousarray(species, dtype=np.int32)
        idp = None
        if ids is not None:
            ids = np.ascontiguousarray(ids, dtype=np.uint64)
            idp = _up(ids)
        self._check(_lib.lib().start_append_particles(self.ctx, len(pos), *[_dp(c) for c in cols],
                                                      _ip(sp), idp))
        self.n += len(pos)

    def generate(self, sources: list[dict], seed: int = 0):
        """Particles drawn on the device from point / surface sources (replaces
        ParticleGeneratorHost / ParticleGeneratorDevice and the displacer).
        Each source: kind ("point" | "surface"), species, count, energy ("mono" |
        "maxwellian" | "thompson" | "lognormal"), angle ("fan" | "cosine" |
        "isotropic" | "cone"), energy_eV, p0, p1, position, theta, phi,
        expansion, displace, and for surface sources centres and normals
        (n_cells x 3)."""
        descs = (_lib.SourceDesc * len(sources))()
        keep = []
        n = 0
        for d, s in zip(descs, sources):
            d.kind = _lib.SOURCE[s.get("kind", "point")]
            d.species = int(s.get("species", 0))
            d.count = int(s["count"])
            d.energy = _lib.ENERGY[s.get("energy", "mono")]
            d.angle = _lib.ANGLE[s.get("angle", "fan")]
            d.energy_eV = float(s.get("energy_eV", 0.0))
            d.p0 = float(s.get("p0", 0.0))
            d.p1 = float(s.get("p1", 0.0))
            for q, v in enumerate(s.get("position", (0.0, 0.0, 0.0))):
                d.position[q] = float(v)
            d.theta = float(s.get("theta", 0.0))
            d.phi = float(s.get("phi", 0.0))
            d.expansion = float(s.get("expansion", 0.0))
            d.displace = float(s.get("displace", 0.0))
            if d.kind == _lib.SOURCE["surface"]:
                cen = np.ascontiguousarray(s["centres"], dtype=np.float64).reshape(-1, 3)
                nor = np.ascontiguousarray(s["normals"], dtype=np.float64).reshape(-1, 3)
                keep += [cen, nor]
                d.centres, d.normals, d.n_cells = _dp(cen), _dp(nor), len(cen)
            n += d.count
        self._check(_lib.lib().start_generate(self.ctx, len(sources), descs, int(seed)))
        self.n = n
        self.n_runs = 1

    def load_batch(self, runs: list[dict], pos: np.ndarray, vel: np.ndarray, species: np.ndarray, ids=None):
        """Independent runs on the shared mesh, their rows consecutive in run
        order.  Each run: {"n": rows, "scale": (s0, s1), "concentration": n,
        "run": variate key}; its field is s0 * basis 0 + s1 * basis 1."""
        pos = np.asarray(pos, dtype=np.float64)
        vel = np.asarray(vel, dtype=np.float64)
        rd = (_lib.RunDesc * len(runs))()
        for k, r in enumerate(runs):
            s = r.get("scale", (1.0, 0.0))
            rd[k] = _lib.RunDesc(int(r["n"]), (C.c_double * 2)(float(s[0]), float(s[1])),
                                 float(r["concentration"]), int(r.get("run", k)), 0)
        assert sum(int(r["n"]) for r in runs) == len(pos)
        cols = [np.ascontiguousarray(a) for a in (pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0],
                                                  vel[:, 1], vel[:, 2])]
        sp = np.ascontiguousarray(species, dtype=np.int32)
        idp = None
        if ids is not None:
            self._ids = np.ascontiguousarray(ids, dtype=np.uint64)
            idp = _up(self._ids)
        self._check(_lib.lib().start_load_batch(self.ctx, len(runs), rd, *[_dp(c) for c in cols], _ip(sp), idp))
        self.n = len(pos)
        self.n_runs = len(runs)

    def set_field_basis(self, which: int, phi: np.ndarray | None = None, E: np.ndarray | None = None):
        if phi is not None:
            phi = np.ascontiguousarray(phi, dtype=np.float64)
            self._check(_lib.lib().start_set_node_potential_basis(self.ctx, int(which), _dp(phi)))
        else:
            E = np.ascontiguousarray(E, dtype=np.float64).reshape(-1)
            self._check(_lib.lib().start_set_tet_field_basis(self.ctx, int(which), _dp(E)))

    def node_charge_runs(self) -> np.ndarray:
        out = np.empty((self.n_runs, self.mesh.n_nodes), dtype=np.float64)
        self._check(_lib.lib().start_read_node_charge_runs(self.ctx, _dp(out)))
        return out

    def tri_counts_runs(self) -> np.ndarray:
        out = np.zeros((self.n_runs, max(self.mesh.n_tris, 1)), dtype=np.uint64)
        self._check(_lib.lib().start_read_tri_counts_runs(self.ctx, _up(out)))
        return out[:, : self.mesh.n_tris]

    def set_launch(self, steps_per_launch=1, block=256, persistent=False, sort_interval=20, graph=False,
                   walk=None):
        """Launch shape; walk=None keeps the multi-tet walk setting (on unless
        switched off)."""
        if walk is None:
            walk = self.get_launch()["walk"]
        cfg = _lib.LaunchCfg(int(steps_per_launch), int(block), int(persistent), int(sort_interval), int(graph),
                             int(walk))
        self._check(_lib.lib().start_set_launch(self.ctx, C.byref(cfg)))

    def get_launch(self) -> dict:
        cfg = _lib.LaunchCfg()
        self._check(_lib.lib().start_get_launch(self.ctx, C.byref(cfg)))
        return {k: getattr(cfg, k) for k, _ in _lib.LaunchCfg._fields_}

    def autotune(self, trial_steps: int = 200, cache: str | None = None):
        """Project autotuner: times the launch configurations from the current
        state (restored afterwards) and keeps the fastest; returns it and its
        ms per step.  With a cache file, a context of the same shape on the
        same device takes the cached result without trials (self.autotune_hit)."""
        cfg = _lib.LaunchCfg()
        ms = C.c_double(0)
        if cache is None:
            self._check(_lib.lib().start_autotune(self.ctx, int(trial_steps), C.byref(cfg), C.byref(ms)))
            self.autotune_hit = False
        else:
            hit = C.c_int32(0)
            self._check(_lib.lib().start_autotune_cached(self.ctx, os.fsencode(cache), int(trial_steps),
                                                         C.byref(cfg), C.byref(ms), C.byref(hit)))
            self.autotune_hit = bool(hit.value)
        return {k: getattr(cfg, k) for k, _ in _lib.LaunchCfg._fields_}, ms.value

    def record_settle(self, on: bool = True):
        self._check(_lib.lib().start_record_settle(self.ctx, int(on)))

    def hits(self):
        """Hit point, triangle and settle step of every particle (load order; -1 if not settled)."""
        xh = np.empty((self.n, 3), dtype=np.float64)
        tri = np.empty(self.n, dtype=np.int32)
        step = np.empty(self.n, dtype=np.int32)
        self._check(_lib.lib().start_read_hits(self.ctx, _dp(xh), _ip(tri), _ip(step)))
        return xh, tri, step

    def track_movements(self, max_particles: int = 200_000, max_points: int = 1000):
        """Record the movements of the first `max_particles` moving particles
        (load order): the start of every step and the surface hit point, up to
        `max_points` each.  0 stops recording."""
        self._max_points = int(max_points)
        self._check(_lib.lib().start_track_movements(self.ctx, int(max_particles), int(max_points)))

    def movements(self) -> dict:
        """{particle id: (k, 3) array of recorded points} for the tracked
        particles (the first max_points of each; see movement_counts)."""
        L = _lib.lib()
        n = C.c_int64(0)
        self._check(L.start_read_movements(self.ctx, C.byref(n), None, None, None))
        m, kmax = n.value, getattr(self, "_max_points", 0)
        ids = np.empty(max(m, 1), dtype=np.uint64)
        counts = np.empty(max(m, 1), dtype=np.int32)
        xyz = np.empty(max(m * kmax * 3, 1), dtype=np.float64)
        self._check(L.start_read_movements(self.ctx, C.byref(n), _up(ids), _ip(counts), _dp(xyz)))
        xyz = xyz[: m * kmax * 3].reshape(m, kmax, 3)
        return {int(ids[s]): xyz[s, : min(int(counts[s]), kmax)].copy() for s in range(m)}

    def clear_movements(self):
        """Empties the recorded points; the tracked particles stay."""
        self._check(_lib.lib().start_clear_movements(self.ctx))

    def movement_counts(self) -> tuple[np.ndarray, np.ndarray]:
        """(ids, points recorded) per tracked particle, counts possibly above max_points."""
        L = _lib.lib()
        n = C.c_int64(0)
        self._check(L.start_read_movements(self.ctx, C.byref(n), None, None, None))
        ids = np.empty(max(n.value, 1), dtype=np.uint64)
        counts = np.empty(max(n.value, 1), dtype=np.int32)
        self._check(L.start_read_movements(self.ctx, C.byref(n), _up(ids), _ip(counts), None))
        return ids[: n.value], counts[: n.value]

    def adjoint_push(self, lam: np.ndarray, first_step: int = 0) -> np.ndarray:
        """dL/d(field scales) for lam = dL/d(hit point) per particle (n x 3, load order)."""
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
        g = np.zeros(2, dtype=np.float64)
        self._check(_lib.lib().start_adjoint_push(self.ctx, _dp(lam), int(first_step), _dp(g)))
        return g

    def set_tet_field(self, E: np.ndarray):
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(-1)
        self._check(_lib.lib().start_set_tet_field(self.ctx, _dp(E)))

    def set_node_potential(self, phi: np.ndarray):
        """Node potentials -> per-tet E on the device (asynchronous, stream-ordered)."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        self._check(_lib.lib().start_set_node_potential(self.ctx, _dp(phi)))

    def tet_field(self, which: int = 0) -> np.ndarray:
        """The per-tet field basis `which` as the step kernels see it, (n_tets, 3)."""
        out = np.empty((self.mesh.n_tets, 3), dtype=np.float64)
        self._check(_lib.lib().start_read_tet_field(self.ctx, int(which), _dp(out)))
        return out

    # --- the two step-loop calls ------------------------------------------------
    def process(self, nsteps: int = 1):
        """ParticleDynamicsProcessor::process, nsteps times (asynchronous)."""
        self._check(_lib.lib().start_step(self.ctx, int(nsteps)))

    def gather(self):
        """NodeChargeDensityProcessor::gather for the current positions."""
        self._check(_lib.lib().start_gather_charge(self.ctx))

    # --- outputs ----------------------------------------------------------------
    def node_charge(self) -> np.ndarray:
        out = np.empty(self.mesh.n_nodes, dtype=np.float64)
        self._check(_lib.lib().start_read_node_charge(self.ctx, _dp(out)))
        return out

    def node_charge_density(self, mode: str = "barycentric"):
        """Node charge density (C/m^3) and the written-node mask, as gather leaves
        NodeChargeDensitiesMap: "barycentric" (Q_n over the lumped node volume) or
        "tet_average" (the reference formula, NodeChargeDensityProcessor.cpp:60-93)."""
        rho = np.empty(self.mesh.n_nodes, dtype=np.float64)
        wr = np.empty(self.mesh.n_nodes, dtype=np.int32)
        self._check(_lib.lib().start_node_charge_density(self.ctx, _lib.DENSITY[mode], _dp(rho), _ip(wr)))
        return rho, wr.astype(bool)

    def node_charge_density_view(self, mode: str = "barycentric"):
        """The same through start_node_charge_density_view: copies of the
        context's page-locked buffers (which the next density call reuses)."""
        rp, wp = C.POINTER(C.c_double)(), C.POINTER(C.c_int32)()
        self._check(_lib.lib().start_node_charge_density_view(self.ctx, _lib.DENSITY[mode], C.byref(rp),
                                                               C.byref(wp)))
        M = self.mesh.n_nodes
        return (np.ctypeslib.as_array(rp, shape=(M,)).copy(),
                np.ctypeslib.as_array(wp, shape=(M,)).astype(bool))

    def particles(self) -> dict:
        n = self.n
        f = {k: np.empty(n, dtype=np.float64) for k in ("x", "y", "z", "vx", "vy", "vz")}
        i = {k: np.empty(n, dtype=np.int32) for k in ("tet", "status", "hit")}
        self._check(_lib.lib().start_read_particles(
            self.ctx, *[_dp(f[k]) for k in ("x", "y", "z", "vx", "vy", "vz")],
            *[_ip(i[k]) for k in ("tet", "status", "hit")]))
        f.update(i)
        return f

    def owners(self) -> tuple[np.ndarray, np.ndarray]:
        """Owning tet and status by load index: the integer part of particles()."""
        n = self.n
        tet = np.empty(n, dtype=np.int32)
        status = np.empty(n, dtype=np.int32)
        self._check(_lib.lib().start_read_particles(self.ctx, *([None] * 6), _ip(tet), _ip(status), None))
        return tet, status

    def tri_counts(self) -> np.ndarray:
        out = np.zeros(max(self.mesh.n_tris, 1), dtype=np.uint64)
        self._check(_lib.lib().start_read_tri_counts(self.ctx, _up(out)))
        return out[: self.mesh.n_tris]

    def settled_ids(self) -> np.ndarray:
        n = C.c_int64(self.n)
        out = np.empty(max(self.n, 1), dtype=np.uint64)
        self._check(_lib.lib().start_read_settled(self.ctx, _up(out), C.byref(n)))
        return out[: n.value]

    def settled_ids_since(self, first: int) -> np.ndarray:
        """Ids of the particles that settled after the first `first` to settle, in settle order."""
        L = _lib.lib()
        n = C.c_int64(0)
        self._check(L.start_read_settled_since(self.ctx, int(first), None, C.byref(n)))
        out = np.empty(max(n.value, 1), dtype=np.uint64)
        self._check(L.start_read_settled_since(self.ctx, int(first), _up(out), C.byref(n)))
        return out[: n.value]

    def stats(self) -> dict:
        s = _lib.Stats()
        self._check(_lib.lib().start_read_stats(self.ctx, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.Stats._fields_ if k != "pad"}

    def synchronize(self):
        self._check(_lib.lib().start_synchronize(self.ctx))

    def locate(self, pts: np.ndarray) -> np.ndarray:
        pts = np.asarray(pts, dtype=np.float64)
        x, y, z = (np.ascontiguousarray(pts[:, k]) for k in range(3))
        out = np.empty(len(pts), dtype=np.int32)
        self._check(_lib.lib().start_locate(self.ctx, len(pts), _dp(x), _dp(y), _dp(z), _ip(out)))
        return out

    def locate_noalloc(self, pts: np.ndarray) -> np.ndarray:
        """start_locate without device allocation (at most NOALLOC_LOCATE_MAX points)."""
        pts = np.asarray(pts, dtype=np.float64)
        x, y, z = (np.ascontiguousarray(pts[:, k]) for k in range(3))
        out = np.empty(len(pts), dtype=np.int32)
        self._check(_lib.lib().start_locate_noalloc(self.ctx, len(pts), _dp(x), _dp(y), _dp(z), _ip(out)))
        return out

    def timing(self, on: bool):
        self._check(_lib.lib().start_timing_enable(self.ctx, int(on)))

    def timing_read_each(self) -> np.ndarray:
        """Milliseconds of the streaming kernel in each timed launch (resets)."""
        L = _lib.lib()
        n = C.c_int64(0)
        self._check(L.start_timing_read_each(self.ctx, None, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.float64)
        self._check(L.start_timing_read_each(self.ctx, _dp(out), C.byref(n)))
        return out[: n.value]

    def timing_read(self):
        ms = C.c_double(0)
        n = C.c_uint64(0)
        self._check(_lib.lib().start_timing_read(self.ctx, C.byref(ms), C.byref(n)))
        return ms.value, n.value


def mesh_maps(mesh: TetMesh) -> dict:
    """The barycentric maps and certification thresholds the device uses
    (computed on the host; no GPU needed): positively oriented tets, geo
    (n_tets x 12: v0, A row-major), rho (per tet: bound on the row sums of
    |A E - I|), tau, geo32 (n_tets x 12: A, v0) and tau32."""
    nodes = np.ascontiguousarray(mesh.nodes, dtype=np.float64)
    tets = np.ascontiguousarray(mesh.tets, dtype=np.int32)
    tris = np.ascontiguousarray(mesh.tris, dtype=np.int32)
    md = _lib.MeshDesc(len(nodes), _dp(nodes), len(tets), _ip(tets), len(tris), _ip(tris))
    T = len(tets)
    out_tets = np.empty((T, 4), dtype=np.int32)
    geo = np.empty((T, 12), dtype=np.float64)
    rho = np.empty(T, dtype=np.float64)
    geo32 = np.empty((T, 12), dtype=np.float32)
    tau, tau32 = C.c_double(0), C.c_double(0)
    _lib.check(_lib.lib().start_mesh_maps(C.byref(md), _ip(out_tets), _dp(geo), _dp(rho), C.byref(tau),
                                          geo32.ctypes.data_as(C.POINTER(C.c_float)), C.byref(tau32)), None)
    return {"tets": out_tets, "geo": geo, "rho": rho, "tau": tau.value, "geo32": geo32, "tau32": tau32.value}


def orient3d_batch(pts: np.ndarray) -> np.ndarray:
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 12)
    out = np.empty(len(pts), dtype=np.int32)
    _lib.check(_lib.lib().start_orient3d_batch(len(pts), _dp(pts), _ip(out)), None)
    return out


MATH_FN = {"pow": 0, "sincos": 1, "acos": 2, "atan2": 3, "log": 4, "exp": 5}


def math_batch(fn: str, x, y=None):
    """The device's elementary functions of the collision path and the
    particle sources over arrays: pow(x, y), sincos(x) -> (sin, cos),
    acos(x), atan2(y, x), log(x), exp(x)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, dtype=np.float64).ravel()
    if x.shape != y.shape:
        raise ValueError("x and y must have the same length")
    out, out2 = np.empty_like(x), np.empty_like(x)
    _lib.check(_lib.lib().start_math_batch(MATH_FN[fn], len(x), _dp(x), _dp(y), _dp(out), _dp(out2)), None)
    return (out, out2) if fn == "sincos" else out
