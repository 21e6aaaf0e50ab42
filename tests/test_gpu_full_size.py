"""The headline workload at its full size (1,000,000 Ar+ on the 750,000-tet
box): FP64 bit-exact against the oracle over fused launches, and charge
conservation (the node charge sums to the charge of the live particles) over
a long fused run in both precisions."""
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
