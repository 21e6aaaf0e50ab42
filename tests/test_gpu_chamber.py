"""The chamber workload at reduced size (12 x 12 x 12 hexes x 6 tets,
20,000 Ti neutrals and Ar+, VSS): the node potential refreshed every 10 steps
through the asynchronous hand-off, the device field against the oracle's, and
the step against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle as O
from start_amd import workloads
from start_amd.engine import ParticleInCellStep

from ._parity import Pair, log

pytestmark = pytest.mark.gpu


def small_chamber(n_particles=20_000):
    return workloads.chamber(n_particles=n_particles, n=12, nz=12)


def field_scale(mesh, phi):
    """Per tet: sum_i |phi_i| |grad lambda_i|_inf, the size of the terms E
    is summed from (the scale of its rounding)."""
    nd, tt = mesh.nodes, mesh.tets
    ed = np.stack([nd[tt[:, j]] - nd[tt[:, 0]] for j in (1, 2, 3)], axis=2)  # columns v_j - v0
    g = np.linalg.inv(ed)  # rows: grad lambda_1..3
    g0 = -g.sum(axis=1)
    grads = np.concatenate([g0[:, None, :], g], axis=1)  # (T, 4, 3)
    return (np.abs(phi[tt])[:, :, None] * np.abs(grads)).sum(axis=1).max(axis=1)


def test_field_from_potential_matches_oracle():
    """E = -grad(phi) per tet from a node potential (the FEM solve's hand-off):
    device against the oracle, |dE| <= 1e-12 x the scale of the summed terms;
    refreshed potentials replace the field, basis 1 is independent."""
    w = small_chamber(n_particles=100)
    om = O.Mesh(w.mesh.nodes, w.mesh.tets, w.mesh.tris)
    eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed, precision="fp64")
    eng.load_particles(w.pos, w.vel, w.sp_idx)
    rng = np.random.default_rng(5)
    for k, phi in enumerate([w.potential(0), w.potential(50), rng.normal(0.0, 100.0, w.mesh.n_nodes)]):
        eng.set_node_potential(phi)
        got = eng.tet_field(0)
        ref = om.tet_field(phi)
        err = np.abs(got - ref).max(axis=1) / np.maximum(field_scale(w.mesh, phi), 1e-300)
        log(f"field {k}: max scaled error {err.max():.2e}, max |E| {np.abs(ref).max():.3e}")
        assert err.max() <= 1e-12
    phi1 = rng.normal(0.0, 10.0, w.mesh.n_nodes)
    eng.set_field_basis(1, phi=phi1)
    assert np.abs(eng.tet_field(1) - om.tet_field(phi1)).max() <= 1e-12 * field_scale(w.mesh, phi1).max()
    assert np.array_equal(eng.tet_field(0), got)  # basis 0 untouched
    eng.close()


def refresh(pr, w, step):
    """New node potential on the device; the oracle steps with the field the
    device derived from it (checked against the oracle's own field above)."""
    pr.eng.set_node_potential(w.potential(step))
    pr.sim.E = np.ascontiguousarray(pr.eng.tet_field(0)).reshape(-1)


@pytest.mark.parametrize("K", [1, 10])
def test_chamber_fp64_parity_with_refresh(K):
    """FP64, charge every step, potential refreshed every 10 steps between
    launches: owners, hits, statuses, triangle counts, x and v bit-exact,
    node charge within 1e-12."""
    w = small_chamber()
    pr = Pair(w, E=np.zeros((w.mesh.n_tets, 3)), steps_per_launch=K)
    for s in range(0, 60, w.refresh):
        refresh(pr, w, s)
        pr.advance(w.refresh)
        if (s + w.refresh) % 30 == 0:
            r = pr.compare(f"chamber fp64 K={K} step {pr.step}")
    sub = w.chamber.substrate
    log(f"chamber fp64 K={K}: settled {r['settled']}, substrate deposits "
        f"{int(pr.sim.tri_count[: len(sub)][sub].sum())}")
    assert r["settled"] > 1000 and r["owner_changed"] > 1000
    pr.close()


def test_chamber_mixed_against_fp64_oracle():
    """Mixed precision on the chamber with refreshes, to the mixed bars
    (owner/status/hit for all but 1e-3 of the particles, x and v within 1e-5,
    node charge L1 within 1e-5)."""
    w = small_chamber()
    pr = Pair(w, E=np.zeros((w.mesh.n_tets, 3)), steps_per_launch=10, precision="mixed")
    for s in range(0, 60, w.refresh):
        refresh(pr, w, s)
        pr.advance(w.refresh)
    pr.compare_mixed("chamber")
    pr.close()
