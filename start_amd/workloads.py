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
chamber    cylinder r 0.15 m, h 0.3 m (80 x 80 x 78 hexes x 6 tets: 2,995,200
           tets, 518,319 nodes), target disk r <= 0.05 m on top, substrate plate
           (10,080 triangles) at z = 0.1 m; 16,000,000 particles: Ti neutrals
           (Thompson, U_b 4.89 eV, cosine law) and Ar+ (1 eV Maxwellian); VSS Ar
           at 0.3 Pa; node potential refreshed every 10 steps; dt 5e-8 s,
           1,000 steps; mixed precision.
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


def thompson_energy(n, rng, u_b, e_max):
    """Thompson distribution f(E) ~ E / (E + U_b)^3 on [0, e_max] (eV) by
    inverse CDF: F(E) = E^2 / (E + U_b)^2, E = U_b sqrt(F) / (1 - sqrt(F))."""
    f_max = (e_max / (e_max + u_b)) ** 2
    r = np.sqrt(rng.uniform(0.0, f_max, n))
    return u_b * r / (1.0 - r)


@dataclass
class ChamberWorkload:
    """Sputtering chamber: target disk on top, substrate plate inside, mixed
    Ti neutrals from the target and Ar+ in the volume, the field refreshed
    from a node potential every `refresh` steps."""
    name: str
    chamber: meshgen.Chamber
    species: list
    gas: Gas
    dt: float
    steps: int
    refresh: int
    seed: int
    pos: np.ndarray
    vel: np.ndarray
    sp_idx: np.ndarray
    shape: np.ndarray | None = None  # spatial profile of the potential (built on first use)

    @property
    def mesh(self):
        return self.chamber.mesh

    def potential(self, step: int) -> np.ndarray:
        """Node potential in force from `step` on: the target biased around
        -300 V (modulated by 20% over 200 steps), decaying over 2 cm below
        it and to 0 V at the wall.  Refreshed every `refresh` steps."""
        if self.shape is None:
            nd = self.chamber.mesh.nodes
            r2 = (nd[:, 0] ** 2 + nd[:, 1] ** 2) / self.chamber.radius ** 2
            # wall nodes (r = radius up to the rounding of the map) at exactly 0 V
            self.shape = np.exp(-(self.chamber.height - nd[:, 2]) / 0.02) * np.where(r2 >= 1.0 - 1e-9, 0.0, 1.0 - r2)
        k = (step // self.refresh) * self.refresh
        return (-300.0 * (1.0 + 0.2 * np.sin(2.0 * np.pi * k / 200.0))) * self.shape


def chamber(n_particles: int = 16_000_000, n: int = 80, nz: int = 78, ti_fraction: float = 0.5,
            steps: int = 1000, dt: float = 5e-8, refresh: int = 10, seed: int = 1005) -> ChamberWorkload:
    """Production-like chamber (80 x 80 x 78 grid: 2,995,200 tets, 518,319
    nodes, 10,080 substrate triangles).  Ti neutrals (Thompson energies with
    U_b = 4.89 eV up to 100 eV, cosine law) leave the target disk (r <= 0.05 m)
    0.1 mm below it; Ar+ (+1e, 1 eV Maxwellian) fill the volume uniformly.  VSS
    collisions with Ar at 0.3 Pa, 300 K."""
    ch = meshgen.cylinder_chamber(n=n, nz=nz)
    rng = np.random.default_rng(seed)
    n_ti = int(round(n_particles * ti_fraction))
    n_ar = n_particles - n_ti
    ti = species("Ti", "Ar", charge_units=0.0)
    ar = species("Ar", "Ar", charge_units=1.0, diameter=4.17e-10)
    g = gas("Ar", 0.3, 300.0, "VSS")
    # Ti from the target disk, cosine law into the chamber
    rt = 0.05 * np.sqrt(rng.uniform(0.0, 1.0, n_ti))
    at = rng.uniform(0.0, 2.0 * np.pi, n_ti)
    pos_ti = np.stack([rt * np.cos(at), rt * np.sin(at), np.full(n_ti, ch.height - 1e-4)], axis=1)
    e = thompson_energy(n_ti, rng, 4.89, 100.0)
    speed = np.sqrt(2.0 * e * EV_J / MASS["Ti"])
    ct = np.sqrt(rng.uniform(0.0, 1.0, n_ti))
    st = np.sqrt(1.0 - ct * ct)
    ph = rng.uniform(0.0, 2.0 * np.pi, n_ti)
    vel_ti = speed[:, None] * np.stack([st * np.cos(ph), st * np.sin(ph), -ct], axis=1)
    # Ar+ uniform in the volume, Maxwellian
    ra = 0.999 * ch.radius * np.sqrt(rng.uniform(0.0, 1.0, n_ar))
    aa = rng.uniform(0.0, 2.0 * np.pi, n_ar)
    pos_ar = np.stack([ra * np.cos(aa), ra * np.sin(aa), rng.uniform(1e-4, ch.height - 1e-4, n_ar)], axis=1)
    vel_ar = rng.normal(0.0, np.sqrt(1.0 * EV_J / MASS["Ar"]), size=(n_ar, 3))
    pos = np.concatenate([pos_ti, pos_ar])
    vel = np.concatenate([vel_ti, vel_ar])
    sp = np.concatenate([np.zeros(n_ti, dtype=np.int32), np.ones(n_ar, dtype=np.int32)])
    return ChamberWorkload("chamber", ch, [ti, ar], g, dt, steps, refresh, seed, pos, vel, sp)


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
