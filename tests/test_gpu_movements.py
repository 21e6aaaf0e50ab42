"""Movement records on the GPU (start_track_movements, the particleMovementMap
of process) against the oracle: which particles are tracked, how many points
each records, every start-of-step position, and the hit point of each
settling segment on its triangle."""
from __future__ import annotations

import json

import numpy as np
import pytest

from oracle import pyoracle as O

from ._parity import Pair, log
from .test_gpu_gmsh import gmsh_workload

pytestmark = pytest.mark.gpu


def oracle_run(pr, steps, n_track):
    """Steps the oracle like Pair.advance and lists, per tracked particle, what
    the reference's map receives: ("start", x) each step it is moved and
    ("hit", x_a, x_b, triangle) in the step it settles."""
    st, n = pr.st, pr.n
    expect, tracked = {}, None
    for j in range(steps):
        if pr.charge and j % pr.K == 0:
            pr.qo[:] = 0.0
        x0 = np.stack([st.x, st.y, st.z], 1)
        settled0 = st.status == O.SETTLED
        moved = np.zeros(n, dtype=np.uint8)
        pr.processed_o += pr.sim.step(st, pr.step, pr.qo, nthreads=8, moved=moved)
        pr.step += 1
        if tracked is None:
            tracked = np.nonzero(moved)[0][:n_track]
            expect = {int(i): [] for i in tracked}
        for i in tracked:
            if moved[i]:
                expect[int(i)].append(("start", x0[i]))
                if not settled0[i] and st.status[i] == O.SETTLED:
                    expect[int(i)].append(("hit", x0[i], np.array([st.x[i], st.y[i], st.z[i]]), int(st.hit[i])))
    return expect


def check_hit(mesh, h, a, b, tri, tol):
    """h lies on the segment a -> b and on triangle `tri` (within tol)."""
    A, B, C = mesh.nodes[mesh.tris[tri]]
    d = b - a
    s = np.dot(h - a, d) / np.dot(d, d)
    scale = max(np.abs(a).max(), np.abs(b).max(), 1e-300)
    assert -tol <= s <= 1 + tol, s
    assert np.abs(h - (a + s * d)).max() <= tol * scale
    n = np.cross(B - A, C - A)
    assert abs(np.dot(n / np.linalg.norm(n), h - A)) <= tol * scale
    M = np.c_[B - A, C - A, n]
    lam = np.linalg.solve(M, h - A)
    assert lam[0] >= -tol and lam[1] >= -tol and lam[0] + lam[1] <= 1 + tol


@pytest.mark.parametrize("precision,K", [("fp64", 10), ("fp64", 1), ("mixed", 10)])
def test_movements_match_the_oracle(precision, K, tmp_path):
    w = gmsh_workload("cube")
    w.vel[:40] *= 1e-6  # slow (|v| < 10 m/s): never moved, so never tracked
    n_track, max_points, steps = 1500, 30, 60
    pr = Pair(w, steps_per_launch=K, precision=precision)
    pr.eng.track_movements(n_track, max_points)
    expect = oracle_run(pr, steps, n_track)
    pr.eng.process(steps)
    if precision == "fp64":
        pr.compare(f"movements {precision} K={K}")
    got = pr.eng.movements()
    ids, counts = pr.eng.movement_counts()
    assert sorted(got) == sorted(expect), "tracked particles differ"
    assert min(got) >= 40
    cnt = dict(zip(ids.tolist(), counts.tolist()))
    g, st = pr.eng.particles(), pr.st
    # (mixed: a collision decided the other way round in FP32 sends a particle elsewhere)
    agree = (g["tet"] == st.tet) & (g["status"] == st.status) & (g["hit"] == st.hit)
    same = [i for i in expect if cnt[i] == len(expect[i]) and agree[i]]
    n_trunc = sum(cnt[i] > max_points for i in expect)
    n_hits = sum(any(e[0] == "hit" for e in expect[i][:max_points]) for i in same)
    log(f"movements {precision} K={K}: {len(expect)} tracked, counts equal {len(same)}, "
        f"truncated {n_trunc}, hit points checked {n_hits}")
    assert n_trunc > 100 and n_hits > 300
    if precision == "fp64":
        assert len(same) == len(expect)
        tol_x, tol_h = 0.0, 1e-9
    else:
        assert len(same) >= 0.98 * len(expect)
        tol_x, tol_h = 1e-5, 1e-5
    for i in same:
        pts = got[i]
        assert len(pts) == min(cnt[i], max_points)
        for k, e in enumerate(expect[i][:max_points]):
            if e[0] == "start":
                assert np.abs(pts[k] - e[1]).max() <= tol_x * max(np.abs(e[1]).max(), 1.0), (i, k)
            else:
                check_hit(w.mesh, pts[k], e[1], e[2], e[3], tol_h)
    # the JSON file of the records reads back to the same points
    from start_amd import outputs
    p = tmp_path / "movements.json"
    outputs.write_movements_json(str(p), got)
    back = json.loads(p.read_text())
    assert set(back) == {str(i) for i, v in got.items() if len(v) > 1}
    k0 = next(iter(back))
    assert np.array_equal(np.array([[q["x"], q["y"], q["z"]] for q in back[k0]]), got[int(k0)])
    pr.close()


def test_tracking_stops_and_resets():
    w = gmsh_workload("sphere", n=2000)
    pr = Pair(w, steps_per_launch=5)
    pr.eng.track_movements(100, 8)
    pr.advance(5)
    ids, counts = pr.eng.movement_counts()
    assert len(ids) == 100 and counts.max() <= 6 and counts.min() >= 1
    pr.eng.track_movements(0, 0)  # stops
    pr.advance(5)
    assert len(pr.eng.movement_counts()[0]) == 0
    pr.eng.track_movements(50, 4)
    pr.advance(10)
    ids, counts = pr.eng.movement_counts()
    assert len(ids) == 50 and counts.max() > 4  # counted beyond the kept points
    assert all(len(v) <= 4 for v in pr.eng.movements().values())
    pr.compare("tracking stop/restart")
    pr.eng.load_particles(w.pos[:10], w.vel[:10], w.sp_idx[:10])  # a new population stops it
    assert len(pr.eng.movement_counts()[0]) == 0
    with pytest.raises(RuntimeError):
        pr.eng.track_movements(10, 0)
    pr.close()
