"""The workloads at full size: the headline (1,000,000 Ar+ on the 750,000-tet
box) FP64 bit-exact against the oracle over fused launches, with charge
conservation over a long fused run in both precisions; the production-like
chamber (16M particles, 3M tets) and the 64-run batched sweep (24.6M
particles) FP64 bit-exact against the oracle over their first steps."""
from __future__ import annotations

import numpy as np
import pytest

from start_amd import workloads
from start_amd.engine import ParticleInCellStep

from ._parity import Pair, log

pytestmark = pytest.mark.gpu


def test_full_size_fp64_parity():
    w = workloads.large_box()
    assert len(w.pos) == 1_000_000 and w.mesh.n_tets == 750_000
    pr = Pair(w, steps_per_launch=10, sort_interval=20)
    for _ in range(2):
        pr.advance(10)
        r = pr.compare(f"full-size large box step {pr.step}")
    assert r["processed"] > 19_000_000
    pr.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_full_size_charge_conservation(precision, tol):
    """After 300 fused steps and one single-step launch, sum_n Q_n equals the
    charge of the particles that own a tet and have not settled (each one's
    barycentric weights sum to 1; in mixed precision to 1 within FP32
    rounding)."""
    w = workloads.large_box()
    eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed, charge_every_step=True,
                             steps_per_launch=100, sort_interval=200, precision=precision)
    eng.load_particles(w.pos, w.vel, w.sp_idx)
    eng.set_node_potential(w.phi)
    eng.process(300)
    eng.set_launch(steps_per_launch=1, sort_interval=200)
    eng.process(1)  # the node charge now holds this one step's scatter
    p = eng.particles()
    q = np.array([s.charge for s in w.species])[w.sp_idx]
    live = (p["status"] == 0) & (p["tet"] >= 0)
    total = eng.node_charge().sum()
    expect = q[live].sum()
    rel = abs(total - expect) / abs(expect)
    log(f"full-size {precision}: node charge {total:.9e} C, live charge {expect:.9e} C, rel {rel:.2e}, "
        f"settled {int((p['status'] == 1).sum())}")
    assert rel < tol
    eng.close()


def test_full_size_chamber_fp64_parity():
    """The production-like chamber at full size (16M Ti neutrals and Ar+,
    2,995,200 tets, VSS, charge every step), the node potential handed over
    at steps 0 and 10, 20 steps in launches of 10: FP64 bit-exact against
    the oracle."""
    from .test_gpu_chamber import refresh
    w = workloads.chamber()
    assert len(w.pos) == 16_000_000 and w.mesh.n_tets == 2_995_200
    log("full-size chamber: workload built")
    pr = Pair(w, E=np.zeros((w.mesh.n_tets, 3)), steps_per_launch=10, sort_interval=50)
    for s in (0, 10):
        refresh(pr, w, s)
        pr.advance(10)
        log(f"full-size chamber: step {pr.step}")
    r = pr.compare(f"full-size chamber step {pr.step}")
    assert r["processed"] > 300_000_000 and r["owner_changed"] > 10_000
    pr.close()


def test_full_size_sweep_fp64_parity():
    """The batched sweep at full size (64 runs of log-uniform size, 24.6M
    particles, 0-1000 V, 0.1-5 Pa VSS, on the 750,000-tet box), 10 steps in
    launches of 5 captured in graphs: every run bit-exact against its own
    oracle run (owners, statuses, hits, per-run triangle counts, x and v;
    per-run node charge within 1e-12)."""
    from oracle import pyoracle as O
    from ._parity import REL_Q, REL_XV
    from start_amd.species import E_CHARGE
    sw = workloads.sweep_box()
    om = O.Mesh(sw.mesh.nodes, sw.mesh.tets, sw.mesh.tris)
    E0 = om.tet_field(sw.phi_unit)
    g = sw.gas
    ks = [O.species_constants(s.mass, s.charge, s.diameter, g.mass, g.omega, g.alpha, O.MODEL[g.model])
          for s in sw.species]
    eng = ParticleInCellStep(sw.mesh, sw.species, g, sw.dt, sw.seed, precision="fp64", steps_per_launch=5,
                             sort_interval=10, graph=True)
    eng.set_field_basis(0, E=E0)
    eng.load_batch(sw.runs, sw.pos, sw.vel, sw.sp_idx)
    log(f"full-size sweep: {len(sw.runs)} runs, {len(sw.pos)} particles loaded")
    sims, states, qs, off = [], [], [], 0
    for run in sw.runs:
        n = run["n"]
        pos, vel = sw.pos[off:off + n], sw.vel[off:off + n]
        tet0 = om.locate_many(pos[:, 0], pos[:, 1], pos[:, 2])
        states.append(O.State.new(pos, vel, sw.sp_idx[off:off + n], np.arange(n, dtype=np.uint64), tet0))
        sims.append(O.Sim(om, ks, O.MODEL[g.model], run["concentration"], sw.dt, sw.seed, run=run["run"],
                          E=run["scale"][0] * E0))
        qs.append(np.zeros(sw.mesh.n_nodes))
        off += n
    for step in range(10):
        for r in range(len(sw.runs)):
            if step % 5 == 0:
                qs[r][:] = 0.0
            sims[r].step(states[r], step, qs[r], nthreads=8)
    eng.process(10)
    log("full-size sweep: 10 steps on both sides")
    got = eng.particles()
    qg = eng.node_charge_runs()
    tg = eng.tri_counts_runs()
    off = 0
    for r, run in enumerate(sw.runs):
        n, st = run["n"], states[r]
        sl = slice(off, off + n)
        for key, ref in (("tet", st.tet), ("status", st.status), ("hit", st.hit)):
            bad = np.nonzero(got[key][sl] != ref)[0]
            assert len(bad) == 0, f"run {r} {key} differs at {bad[:5]}"
        xo = np.stack([st.x, st.y, st.z], 1)
        vo = np.stack([st.vx, st.vy, st.vz], 1)
        xg = np.stack([got["x"][sl], got["y"][sl], got["z"][sl]], 1)
        vg = np.stack([got["vx"][sl], got["vy"][sl], got["vz"][sl]], 1)
        assert (np.abs(xo - xg).max(1) / np.linalg.norm(xo, axis=1)).max() <= REL_XV, f"run {r} x"
        assert (np.abs(vo - vg).max(1) / np.maximum(np.linalg.norm(vo, axis=1), 1e-300)).max() <= REL_XV, f"run {r} v"
        assert np.array_equal(tg[r], sims[r].tri_count[: sw.mesh.n_tris]), f"run {r} triangle counts"
        eq = (np.abs(qg[r] - qs[r]) / np.maximum(np.abs(qs[r]), E_CHARGE)).max()
        assert eq <= REL_Q, f"run {r} node charge {eq:.2e}"
        off += n
    processed = eng.stats()["processed"]
    log(f"full-size sweep: {len(sw.runs)} runs bit-exact over {processed} particle-steps")
    assert processed > 240_000_000
    eng.close()
