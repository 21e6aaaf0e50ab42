"""Edge cases on the GPU against the oracle: an empty population, a single
particle, a population frozen from the start (it still carries charge), a
batch with an empty run, movement records across a batch and through the
autotuner, and a device allocation that fails."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle as O
from start_amd import _lib, workloads
from start_amd.engine import ParticleInCellStep

from ._parity import REL_XV, Pair, log
from .test_gpu_batch import CONCS, KEYS, SCALES, BatchPair
from .test_gpu_movements import oracle_run

pytestmark = pytest.mark.gpu


def test_empty_population():
    w = workloads.small_box(n_particles=50, charge_units=1.0)
    eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed, charge_every_step=True)
    eng.set_tet_field(np.zeros((w.mesh.n_tets, 3)))
    eng.load_particles(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32))
    eng.process(7)
    s = eng.stats()
    assert s["processed"] == 0 and s["settled"] == 0 and not s["stop"]
    assert not eng.node_charge().any()
    assert eng.tri_counts().sum() == 0 and len(eng.settled_ids()) == 0
    assert all(len(v) == 0 for v in eng.particles().values())
    eng.track_movements(10, 4)
    assert len(eng.movement_counts()[0]) == 0 and eng.movements() == {}
    eng.close()


def test_single_particle():
    w = workloads.small_box(n_particles=500, charge_units=2.0)
    pr = Pair(w, n_particles=1, steps_per_launch=5)
    for _ in range(4):
        pr.advance(10)
        r = pr.compare(f"one particle step {pr.step}")
    assert r["processed"] > 0
    pr.close()


def test_frozen_population_still_deposits_charge():
    """|v| < 10 m/s from the start: no particle is processed or moves, and the
    node charge of every step is still its barycentric scatter."""
    w = workloads.small_box(n_particles=800, charge_units=1.0)
    w.vel = w.vel * (1.0 / np.linalg.norm(w.vel, axis=1, keepdims=True)) * 9.5
    pr = Pair(w, steps_per_launch=4)
    pr.advance(12)
    r = pr.compare("frozen population")
    assert r["processed"] == 0 and r["max_rel_q"] <= 1e-12
    assert np.abs(pr.qo).sum() > 0
    pr.close()


def test_out_of_device_memory_is_reported_and_the_context_stays_usable():
    """Movement records of 2e9 points for each of 1,000 particles (48 TB):
    the allocation fails with START_ERR_NOMEM, tracking stays off, and the
    context steps on bit-exact against the oracle and tracks again when asked
    for a size that fits."""
    w = workloads.small_box(n_particles=1000, charge_units=2.0)
    pr = Pair(w, steps_per_launch=1)
    pr.advance(3)
    log("edge: asking for 48 TB of movement records")
    with pytest.raises(_lib.StartError) as e:
        pr.eng.track_movements(max_particles=1000, max_points=2_000_000_000)
    assert e.value.code == _lib.START_ERR_NOMEM, e.value
    assert len(pr.eng.movement_counts()[0]) == 0
    pr.advance(10)
    pr.compare("after a failed allocation")
    pr.eng.track_movements(max_particles=100, max_points=8)
    pr.advance(3)
    ids, counts = pr.eng.movement_counts()
    assert len(ids) == 100 and (counts >= 3).all() and (counts <= 4).all()  # 3 starts, and a hit point
    pr.compare("tracking after a failed allocation")
    pr.close()


def test_batch_with_an_empty_run():
    bp = BatchPair([400, 0, 300], SCALES[:3], CONCS[:3], KEYS[:3], K=5, sort_interval=10)
    bp.advance(20)
    bp.compare("batch with an empty run")
    assert not bp.eng.node_charge_runs()[1].any() and bp.eng.tri_counts_runs()[1].sum() == 0
    bp.close()


def test_movements_in_a_batch():
    """Tracked particles are the first moving ones in load order across the
    runs; points per particle and every start point as the per-run oracles."""
    sizes, n_track = [500, 0, 400], 600
    bp = BatchPair(sizes, SCALES[:3], CONCS[:3], KEYS[:3], K=5, sort_interval=10)
    bp.eng.track_movements(n_track, 64)
    off = np.r_[0, np.cumsum(sizes)]
    tracked, expect = None, None
    for j in range(20):
        moved_all, x0_all, settled_before = [], [], []
        for r in range(len(sizes)):
            st = bp.states[r]
            if j % bp.K == 0:
                bp.qo[r][:] = 0.0
            x0_all.append(np.stack([st.x, st.y, st.z], 1))
            settled_before.append(st.status == O.SETTLED)
            m = np.zeros(len(st.x), dtype=np.uint8)
            bp.sims[r].step(st, bp.step, bp.qo[r], nthreads=8, moved=m)
            moved_all.append(m)
        bp.step += 1
        moved = np.concatenate(moved_all)
        x0 = np.concatenate(x0_all)
        s0 = np.concatenate(settled_before)
        s1 = np.concatenate([st.status == O.SETTLED for st in bp.states])
        if tracked is None:
            tracked = np.nonzero(moved)[0][:n_track]
            expect = {int(i): [] for i in tracked}
        for i in tracked:
            if moved[i]:
                expect[int(i)].append(x0[i])
                if s1[i] and not s0[i]:
                    expect[int(i)].append(None)  # the hit point
    bp.eng.process(20)
    bp.compare("batch movements")
    got = bp.eng.movements()
    ids, counts = bp.eng.movement_counts()
    assert sorted(got) == sorted(expect)
    cnt = dict(zip(ids.tolist(), counts.tolist()))
    assert all(cnt[i] == len(expect[i]) for i in expect)
    for i, e in expect.items():
        for k, p in enumerate(e):
            if p is not None:
                assert np.abs(got[i][k] - p).max() <= REL_XV * np.abs(p).max(), (i, k)
    log(f"batch movements: {len(expect)} tracked over runs of {sizes}, "
        f"{sum(len(e) for e in expect.values())} points")
    bp.close()


def test_autotune_records_no_movements():
    """The autotuner's trials record nothing; the steps after it record as
    the oracle does."""
    w = workloads.small_box(n_particles=3000, charge_units=2.0)
    pr = Pair(w, steps_per_launch=5, sort_interval=10)
    pr.eng.track_movements(1000, 32)
    pr.eng.autotune(trial_steps=20)
    pr.K = pr.eng.get_launch()["steps_per_launch"]  # the oracle's charge window follows the launches
    ids, counts = pr.eng.movement_counts()
    assert len(ids) == 1000 and counts.max() == 0
    expect = oracle_run(pr, 20, 1000)
    pr.eng.process(20)
    pr.compare("after autotune")
    ids, counts = pr.eng.movement_counts()
    cnt = dict(zip(ids.tolist(), counts.tolist()))
    assert sorted(cnt) == sorted(expect) and all(cnt[i] == len(expect[i]) for i in expect)
    pr.close()


def test_stats_and_settled_ids_in_one_call():
    """start_read_stats_settled against start_read_stats + start_read_settled_since
    when more particles settled than its one-sync window (1,024) holds, and more
    than its first staging buffer (4,096): the same statistics and ids."""
    import ctypes as C

    w = workloads.small_box(n_particles=20000, charge_units=2.0)
    eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed)
    eng.load_particles(w.pos, w.vel, w.sp_idx)
    eng.set_node_potential(w.phi)
    eng.process(300)
    L = _lib.lib()
    ref = eng.stats()
    total = int(ref["settled"])
    log(f"stats and settled ids: {total} settled")
    assert total > 5000
    for first in (0, total - 2000, total - 700, total):
        ids_a = np.zeros(total + 1, dtype=np.uint64)
        na = C.c_int64(len(ids_a))
        st = _lib.Stats()
        assert L.start_read_stats_settled(eng.ctx, C.byref(st), first, ids_a.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          C.byref(na)) == 0, first
        ids_b = np.zeros(total + 1, dtype=np.uint64)
        nb = C.c_int64(len(ids_b))
        assert L.start_read_settled_since(eng.ctx, first, ids_b.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          C.byref(nb)) == 0
        assert na.value == nb.value == total - first
        assert np.array_equal(ids_a[:na.value], ids_b[:nb.value]), first
        assert st.settled == total and st.processed == ref["processed"] and st.stop == int(total == len(w.pos))
    small = C.c_int64(10)  # a capacity below the count is an error, as in start_read_settled_since
    assert L.start_read_stats_settled(eng.ctx, C.byref(_lib.Stats()), 0, np.zeros(10, dtype=np.uint64).ctypes.data_as(
        C.POINTER(C.c_uint64)), C.byref(small)) == -1
    eng.close()
