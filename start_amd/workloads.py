"""Synthetic inputs of the measured workloads (seeded, deterministic).

small_box  unit cube 0.1 m, 10x10x10 hexes x 6 tets (6,000 tets, 1,331 nodes,
           1,200 boundary triangles); 10,000 Ti (charge 0) from a point source at
           (0.05, 0.05, 0.095), 5 eV along -z within a 5 degree cone; HS collisions
           with Ar at 0.5 Pa / 300 K, diameter 2.0e-10 m; dt 1e-7 s, 1,000 steps;
           potential 0 V at z=0 and 100 V at z=0.1 m.
large_box  0.5 m cube, 50^3 hexes x 6 tets (750,000 tets, 132,651 nodes, 30,000
           triangles); 1,000,000 Ar+ (+1e), 1 eV Maxwellian, uniform positions;
           VHS with Ar at 1 Pa / 300 K, d_ref 4.17e-10 m, omega 0.81f; charge
           scattered every step; dt 1e-9 s, 2,000 steps; potential 0 / 500 V.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import meshgen
from .species import EV_J, MASS, Gas, Species, concentration, gas, species


@dataclass
class Workload:
    name: str
    mesh: meshgen.TetMesh
    species: list
    gas: Gas
    dt: float
    steps: int
    seed: int
    pos: np.ndarray
    vel: np.ndarray
    sp_idx: np.ndarray
    phi: np.ndarray
    charge_every_step: bool = True
    extra: dict = field(default_factory=dict)


def cone_jet(n, rng, origin, energy_ev, mass, half_angle_deg):
    speed = np.sqrt(2.0 * (energy_ev * EV_J) / mass)
    cos_max = np.cos(np.deg2rad(half_angle_deg))
    ct = rng.uniform(cos_max, 1.0, n)
    st = np.sqrt(1.0 - ct * ct)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    vel = speed * np.stack([st * np.cos(ph), st * np.sin(ph), -ct], axis=1)
    pos = np.tile(np.asarray(origin, dtype=np.float64), (n, 1))
    return pos, vel


def small_box(n_particles: int = 10_000, steps: int = 1000, charge_units: float = 0.0,
              seed: int = 1001) -> Workload:
    mesh = meshgen.box_kuhn(10, 10, 10, 0.1, 0.1, 0.1)
    rng = np.random.default_rng(seed)
    ti = species("Ti", "Ar", charge_units=charge_units, diameter=2.0e-10)
    g = gas("Ar", 0.5, 300.0, "HS")
    pos, vel = cone_jet(n_particles, rng, (0.05, 0.05, 0.095), 5.0, MASS["Ti"], 5.0)
    phi = meshgen.linear_potential(mesh, 0.0, 100.0)
    return Workload("small_box", mesh, [ti], g, 1e-7, steps, seed, pos, vel,
                    np.zeros(n_particles, dtype=np.int32), phi)


def maxwellian_box(mesh, n, rng, mass, kT_ev, margin=0.0):
    lo = mesh.nodes.min(axis=0) + margin
    hi = mesh.nodes.max(axis=0) - margin
    pos = rng.uniform(lo, hi, size=(n, 3))
    sig = np.sqrt(kT_ev * EV_J / mass)
    vel = rng.normal(0.0, sig, size=(n, 3))
    return pos, vel


def large_box(n_particles: int = 1_000_000, steps: int = 2000, cells: int = 50,
              seed: int = 1002, model: str = "VHS") -> Workload:
    mesh = meshgen.box_kuhn(cells, cells, cells, 0.5, 0.5, 0.5)
    rng = np.random.default_rng(seed)
    ar = species("Ar", "Ar", charge_units=1.0, diameter=4.17e-10)
    g = gas("Ar", 1.0, 300.0, model)
    pos, vel = maxwellian_box(mesh, n_particles, rng, MASS["Ar"], 1.0)
    phi = meshgen.linear_potential(mesh, 0.0, 500.0)
    return Workload("large_box", mesh, [ar], g, 1e-9, steps, seed, pos, vel,
                    np.zeros(n_particles, dtype=np.int32), phi)


@dataclass
class Plates:
    """Two parallel electrode plates with particles emitted from the top one;
    the side faces of the gap are open (particles leave there)."""
    name: str
    mesh: meshgen.TetMesh
    species: list
    gas: Gas
    dt: float
    steps: int
    seed: int
    pos: np.ndarray
    vel: np.ndarray
    sp_idx: np.ndarray
    phi_top: np.ndarray     # top plate at 1 V, bottom at 0 V (field basis 0)
    phi_bottom: np.ndarray  # bottom plate at 1 V, top at 0 V (field basis 1)
    bottom: np.ndarray      # bool per surface triangle: on the bottom plate


def plates(n_particles: int = 262_144, nx: int = 51, ny: int = 51, nz: int = 13, steps: int = 1000,
           dt: float = 1e-8, seed: int = 1003, side: float = 0.2, gap: float = 0.05) -> Plates:
    """Plates side x side at a gap (51 x 51 x 13 hexes x 6 = 202,878 tets); Cu+
    (+1e) with lognormal energies (mean 10 eV, sigma_ln 0.5) and cosine-law
    directions into the gap from points uniform on the top plate, 0.1 mm below
    it; no gas."""
    m = meshgen.box_kuhn(nx, ny, nz, side, side, gap)
    z = m.nodes[m.tris][:, :, 2]
    bottom_all = np.all(z == 0.0, axis=1)
    on_plate = bottom_all | np.all(z == gap, axis=1)
    mesh = meshgen.TetMesh(m.nodes, m.tets, np.ascontiguousarray(m.tris[on_plate]))
    rng = np.random.default_rng(seed)
    cu = species("Cu", "Ar", charge_units=1.0)
    g = gas("Ar", 0.0, 300.0, "NONE")
    mu = np.log(10.0) - 0.5 * 0.5 ** 2
    energy = np.exp(rng.normal(mu, 0.5, n_particles))
    speed = np.sqrt(2.0 * energy * EV_J / MASS["Cu"])
    ct = np.sqrt(rng.uniform(0.0, 1.0, n_particles))
    st = np.sqrt(1.0 - ct * ct)
    ph = rng.uniform(0.0, 2.0 * np.pi, n_particles)
    vel = speed[:, None] * np.stack([st * np.cos(ph), st * np.sin(ph), -ct], axis=1)
    pos = np.stack([rng.uniform(0.0, side, n_particles), rng.uniform(0.0, side, n_particles),
                    np.full(n_particles, gap - 1e-4)], axis=1)
    return Plates("plates", mesh, [cu], g, dt, steps, seed, pos, vel, np.zeros(n_particles, dtype=np.int32),
                  meshgen.linear_potential(mesh, 0.0, 1.0), meshgen.linear_potential(mesh, 1.0, 0.0),
                  bottom_all[on_plate])


@dataclass
class Sweep:
    """Independent runs on one mesh: particle counts, electrode voltages and
    pressures per run, rows of the runs consecutive in run order."""
    name: str
    mesh: meshgen.TetMesh
    species: list
    gas: Gas
    dt: float
    steps: int
    seed: int
    runs: list          # {"n", "scale": (volts, 0), "concentration", "run", "pressure"}
    pos: np.ndarray
    vel: np.ndarray
    sp_idx: np.ndarray
    phi_unit: np.ndarray  # potential with the electrode at 1 V (field basis 0)


def sweep_box(n_runs: int = 64, steps: int = 500, cells: int = 50, seed: int = 1004,
              n_min: float = 1e4, n_max: float = 2e6, count_scale: float = 1.0) -> Sweep:
    """Voltage/pressure sweep on the large box mesh: per-run particle counts
    log-uniform in [n_min, n_max] (expected total 2.4e7 for 64 runs), electrode
    voltage uniform in [0, 1000] V (0 V at z = 0), Ar pressure uniform in
    [0.1, 5] Pa at 300 K with VSS collisions (alpha 1.4f, omega 0.81f); Ar+
    (+1e) 1 eV Maxwellian with uniform positions; dt 1e-9 s."""
    mesh = meshgen.box_kuhn(cells, cells, cells, 0.5, 0.5, 0.5)
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_runs))
    counts = np.maximum((counts * count_scale).astype(np.int64), 1)
    volts = rng.uniform(0.0, 1000.0, n_runs)
    press = rng.uniform(0.1, 5.0, n_runs)
    ar = species("Ar", "Ar", charge_units=1.0, diameter=4.17e-10)
    g = gas("Ar", 1.0, 300.0, "VSS")
    n = int(counts.sum())
    pos, vel = maxwellian_box(mesh, n, rng, MASS["Ar"], 1.0)
    runs = [{"n": int(counts[r]), "scale": (float(volts[r]), 0.0),
             "concentration": concentration(float(press[r]), 300.0), "run": r, "pressure": float(press[r])}
            for r in range(n_runs)]
    return Sweep("sweep_box", mesh, [ar], g, 1e-9, steps, seed, runs, pos, vel,
                 np.zeros(n, dtype=np.int32), meshgen.linear_potential(mesh, 0.0, 1.0))
