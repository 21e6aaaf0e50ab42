"""Particles appended to a live population (start_append_particles; the
continuous-sputtering spawn, ModelingMainDriver.cpp:141-142) against the oracle
running the same appends: owners, settled flags, hit triangles and triangle
counts, x and v bit for bit, node charge within 1e-12."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle as O
from start_amd import _lib, workloads
from start_amd.engine import ParticleInCellStep

from ._parity import REL_XV, Pair, log
from .test_gpu_movements import oracle_run  # noqa: F401  (same record rules)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [1, 4])
def test_appended_particles_match_oracle(K):
    w = workloads.small_box(n_particles=4000, charge_units=2.0)
    pr = Pair(w, n_particles=1000, steps_per_launch=K, sort_interval=8)
    pr.advance(8)
    log(f"append K={K}: 400 particles after 8 steps (a sort has compacted the settled ones)")
    pr.append(w.pos[1000:1400], w.vel[1000:1400], w.sp_idx[1000:1400])
    pr.advance(12)
    pr.compare(f"append K={K}, first append")
    lo = 1400
    for j in range(6):  # small appends, some within the grown capacity
        pr.append(w.pos[lo:lo + 100], w.vel[lo:lo + 100], w.sp_idx[lo:lo + 100])
        lo += 100
        pr.advance(4)
    pr.compare(f"append K={K}, after 7 appends")
    pr.advance(60)  # the jet reaches the floor: early and appended particles settle
    res = pr.compare(f"append K={K}, 60 steps later")
    st = pr.st
    assert (st.status[:1000] == O.SETTLED).any() and (st.status[1000:] == O.SETTLED).any()
    s = pr.eng.stats()
    assert s["n_particles"] == lo and s["settled"] == res["settled"]
    pr.close()


def test_append_to_an_empty_context_then_step():
    """A context loaded with no particles (or none yet) takes appended ones."""
    w = workloads.small_box(n_particles=600, charge_units=2.0)
    pr = Pair(w, n_particles=0)
    pr.advance(3)
    pr.append(w.pos[:600], w.vel[:600], w.sp_idx[:600])
    pr.advance(30)
    pr.compare("append to an empty population")
    pr.close()


def test_appended_particles_join_the_movement_records():
    """The reference's map takes a particle while it holds fewer than
    maxParticles (ParticleMovementTracker.cpp:11-12): of the appended
    particles, the first ones that move take the remaining slots."""
    w = workloads.small_box(n_particles=2000, charge_units=2.0)
    pr = Pair(w, n_particles=1000)
    pr.eng.track_movements(max_particles=1250, max_points=64)
    pr.advance(5)
    pr.append(w.pos[1000:1500], w.vel[1000:1500], w.sp_idx[1000:1500])
    x_app = w.pos[1000:1500]
    pr.advance(10)
    pr.compare("append with movement records")
    mv = pr.eng.movements()
    ids = np.array(sorted(mv))
    assert len(ids) == 1250
    assert np.array_equal(ids, np.r_[np.arange(1000), np.arange(1000, 1250)])
    for i in range(1000, 1250):  # first point: where it was appended; one point per step moved since
        pts = mv[i]
        assert np.abs(pts[0] - x_app[i - 1000]).max() <= REL_XV * np.abs(x_app[i - 1000]).max()
        assert 1 <= len(pts) <= 11
    pr.close()


def test_append_mixed_precision_keeps_owners_and_positions():
    """Mixed precision: appended particles are stored in their owner's frame;
    owners equal the oracle's point location and read-back positions are the
    loaded ones to FP32 rounding in that frame."""
    w = workloads.small_box(n_particles=2000, charge_units=2.0)
    eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed, precision="mixed")
    E = O.Mesh(w.mesh.nodes, w.mesh.tets, w.mesh.tris).tet_field(w.phi)
    eng.set_tet_field(E)
    eng.load_particles(w.pos[:1000], w.vel[:1000], w.sp_idx[:1000])
    eng.process(5)
    eng.append_particles(w.pos[1000:], w.vel[1000:], w.sp_idx[1000:])
    g = eng.particles()
    om = O.Mesh(w.mesh.nodes, w.mesh.tets, w.mesh.tris)
    want = om.locate_many(w.pos[1000:, 0], w.pos[1000:, 1], w.pos[1000:, 2])
    assert np.array_equal(g["tet"][1000:], want)
    d = np.abs(np.stack([g["x"], g["y"], g["z"]], 1)[1000:] - w.pos[1000:]).max()
    assert d <= 0.1 * 2.0 ** -23, d  # |local| <= tet size 0.01 m
    eng.process(20)
    s = eng.stats()
    assert s["n_particles"] == 2000 and s["processed"] > 0
    eng.close()


def test_append_errors():
    w = workloads.small_box(n_particles=100)
    eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed)
    eng.load_particles(w.pos[:50], w.vel[:50], w.sp_idx[:50])
    with pytest.raises(_lib.StartError, match="species index out of range"):
        eng.append_particles(w.pos[50:60], w.vel[50:60], np.full(10, 99))
    runs = [{"n": 30, "concentration": 1e20}, {"n": 20, "concentration": 1e20}]
    eng.load_batch(runs, w.pos[:50], w.vel[:50], w.sp_idx[:50])
    with pytest.raises(_lib.StartError, match="single simulation"):
        eng.append_particles(w.pos[50:60], w.vel[50:60], w.sp_idx[50:60])
    eng.close()
