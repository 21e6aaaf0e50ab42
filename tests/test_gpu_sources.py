"""Particle sources drawn on the device (start_generate) against the oracle's
restatement (oracle/pysources.py), and the generated population against the
same state loaded from the host."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle as O
from oracle import pysources as PS
from start_amd import workloads
from start_amd.engine import ParticleInCellStep
from start_amd.species import MASS, species

from ._parity import log

pytestmark = pytest.mark.gpu


def box_and_sources():
    """The small box (0.1 m cube) with four sources: the C1 point source (Ti
    5 eV along -z within 5 degrees, the reference's fan, displaced 1 mm), the
    top face as a surface source (Thompson energies, Lambert law about the
    inward normal), a Maxwellian point source and a lognormal cone."""
    w = workloads.small_box(n_particles=10)
    m = w.mesh
    z = m.nodes[m.tris][:, :, 2]
    top = np.all(z == z.max(), axis=1)
    cen = m.nodes[m.tris[top]].mean(axis=1)
    cen[:, 2] -= 1e-4  # just inside
    nor = np.tile([0.0, 0.0, -3.0], (len(cen), 1))
    srcs = [
        {"kind": "point", "species": 0, "count": 3000, "energy": "mono", "angle": "fan", "energy_eV": 5.0,
         "theta": np.pi, "phi": 0.0, "expansion": np.deg2rad(5.0), "position": (0.05, 0.05, 0.095),
         "displace": 0.001},
        {"kind": "surface", "species": 1, "count": 5001, "energy": "thompson", "angle": "cosine", "p0": 4.89,
         "p1": 100.0, "centres": cen, "normals": nor},
        {"kind": "point", "species": 1, "count": 2000, "energy": "maxwellian", "p0": 1.0,
         "position": (0.031, 0.047, 0.052)},
        {"kind": "point", "species": 0, "count": 1500, "energy": "lognormal", "angle": "cone", "p0": 10.0, "p1": 0.5,
         "theta": 0.4, "phi": 2.0, "expansion": 0.3, "position": (0.02, 0.08, 0.03)},
    ]
    sp = [species("Ti", "Ar", charge_units=0.0), species("Ar", "Ar", charge_units=1.0)]
    return w, sp, srcs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_generate_matches_oracle_sampler(precision):
    """Positions and velocities against the oracle's sampler: FP64 equal to
    the bit (sin, cos, log, exp, acos and atan2 are the same code on both
    sides), the FP32 state within 1e-6; owners equal to the oracle's point
    location; everything active."""
    w, sp, srcs = box_and_sources()
    seed = 20241
    eng = ParticleInCellStep(w.mesh, sp, w.gas, w.dt, w.seed, precision=precision)
    eng.generate(srcs, seed=seed)
    g = eng.particles()
    pos, vel, spi = PS.generate(srcs, [s.mass for s in sp], seed)
    xo = np.stack([g["x"], g["y"], g["z"]], 1)
    vo = np.stack([g["vx"], g["vy"], g["vz"]], 1)
    tol = 0.0 if precision == "fp64" else 1e-6
    ex = (np.abs(xo - pos).max(axis=1) / np.linalg.norm(pos, axis=1)).max()
    ev = (np.abs(vo - vel).max(axis=1) / np.linalg.norm(vel, axis=1)).max()
    log(f"generate {precision}: {len(pos)} particles, max rel x {ex:.2e}, v {ev:.2e}")
    assert ex <= tol and ev <= tol
    om = O.Mesh(w.mesh.nodes, w.mesh.tets, w.mesh.tris)
    tet = om.locate_many(pos[:, 0], pos[:, 1], pos[:, 2])
    same = g["tet"] == tet
    if precision == "fp64":
        assert same.all()
    else:  # the FP32 local position may round onto the other side of a face
        assert same.mean() > 0.999
    assert (g["status"] == 0).all()
    eng.close()


def test_generated_population_steps_like_a_loaded_one():
    """The generated state read back and loaded into a second context from
    the host: owners, statuses and 20 FP64 steps (charge, collisions) agree
    bit for bit, so generation leaves flags, load indices and owners as a
    load does."""
    w, sp, srcs = box_and_sources()
    g1 = ParticleInCellStep(w.mesh, sp, w.gas, w.dt, w.seed)
    g1.generate(srcs, seed=5)
    a = g1.particles()
    pos = np.stack([a["x"], a["y"], a["z"]], 1)
    vel = np.stack([a["vx"], a["vy"], a["vz"]], 1)
    spi = np.concatenate([np.full(s["count"], s["species"], dtype=np.int32) for s in srcs])
    g2 = ParticleInCellStep(w.mesh, sp, w.gas, w.dt, w.seed)
    g2.load_particles(pos, vel, spi)
    b = g2.particles()
    assert np.array_equal(a["tet"], b["tet"]) and np.array_equal(a["status"], b["status"])
    E = O.Mesh(w.mesh.nodes, w.mesh.tets, w.mesh.tris).tet_field(w.phi)
    for e in (g1, g2):
        e.set_tet_field(E)
        e.process(20)
    a, b = g1.particles(), g2.particles()
    for k in ("x", "y", "z", "vx", "vy", "vz", "tet", "status", "hit"):
        assert np.array_equal(a[k], b[k]), k
    q1, q2 = g1.node_charge(), g2.node_charge()  # FP64 atomics: any order, so the charge bar
    assert (np.abs(q1 - q2) / np.maximum(np.abs(q2), 1.602e-19)).max() <= 1e-12
    assert g1.stats()["processed"] == g2.stats()["processed"]
    log(f"generated vs loaded: {len(pos)} particles, 20 steps identical, settled {int((a['status'] == 1).sum())}")
    g1.close()
    g2.close()


def test_generate_rejects_bad_sources():
    """The reference's generator errors (zero count, zero energy, empty
    surface, zero normal) come back as START_ERR_ARG."""
    w, sp, _ = box_and_sources()
    eng = ParticleInCellStep(w.mesh, sp, w.gas, w.dt, w.seed)
    bad = [
        {"kind": "point", "count": 0, "energy_eV": 1.0},
        {"kind": "point", "count": 5, "energy_eV": 0.0},
        {"kind": "surface", "count": 5, "energy_eV": 1.0, "centres": np.zeros((0, 3)), "normals": np.zeros((0, 3))},
        {"kind": "surface", "count": 5, "energy_eV": 1.0, "centres": np.zeros((1, 3)), "normals": np.zeros((1, 3))},
        {"kind": "point", "count": 5, "energy": "thompson", "p0": 0.0, "p1": 1.0},
    ]
    for s in bad:
        with pytest.raises(RuntimeError):
            eng.generate([s], seed=1)
    eng.close()
