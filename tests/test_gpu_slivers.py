"""The step on a conforming mesh full of slivers (tests/_slivers.py: 216 of
1,494 tets with condition numbers 1e5 to 2.7e6, and the certification
threshold tau raised to 4.7e-9 by them): FP64 bit-exact against the oracle,
whose point location and crossings are exact, for fused and single-step
launches; mixed precision to the mixed bars (its FP32 threshold is infinite
here, so the FP64 map decides every step)."""
from __future__ import annotations

import numpy as np
import pytest

from start_amd.engine import mesh_maps

from ._parity import Pair, log
from ._slivers import condition_numbers, sliver_box
from .test_gpu_gmsh import mesh_workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sliver_workload():
    m = sliver_box()
    k = condition_numbers(m)
    mp = mesh_maps(m)
    log(f"sliver box: {m.n_tets} tets, {int((k > 1e5).sum())} with condition number > 1e5 (max {k.max():.2e}), "
        f"tau {mp['tau']:.2e}, tau32 {mp['tau32']}")
    assert (k > 1e5).sum() > 100 and np.isfinite(mp["tau"])
    w = mesh_workload(m, "slivers", n=6000, seed=23)
    # 600 particles start inside slivers (random barycentric points of them)
    rng = np.random.default_rng(5)
    sl = np.nonzero(k > 1e5)[0]
    t = sl[rng.integers(0, len(sl), 600)]
    lam = rng.dirichlet(np.ones(4), 600)
    w.pos[:600] = np.einsum("ij,ijk->ik", lam, m.nodes[m.tets[t]])
    return w


@pytest.mark.parametrize("K", [1, 10])
def test_fp64_parity_on_slivers(sliver_workload, K):
    pr = Pair(sliver_workload, steps_per_launch=K)
    for k in range(3):
        pr.advance(20)
        r = pr.compare(f"slivers K={K} step {pr.step}")
    log(f"slivers K={K}: settled {r['settled']}, owner changed {r['owner_changed']}, deferred {r['deferred']}, "
        f"exhaustive {r['exhaustive']}")
    assert r["settled"] > 0 and r["owner_changed"] > 100
    assert (pr.tet0[:600] >= 0).all()
    pr.close()


def test_mixed_on_slivers(sliver_workload):
    pr = Pair(sliver_workload, steps_per_launch=10, precision="mixed")
    pr.advance(60)
    pr.compare_mixed("slivers mixed")
    pr.close()
