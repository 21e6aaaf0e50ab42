"""End-to-end pins of the oracle's step (ParticleDynamicsProcessor.cpp:161-200
stage order) on single particles: ballistic flight to a wall against the exact
crossing, the frozen rules' strict bounds, the uniform-field closed form, the
hit on the pre-collision segment, and the degenerate-segment skip."""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

from start_amd import meshgen
from start_amd import species as S

AR = S.MASS["Ar"]


@pytest.fixture(scope="module")
def box(oracle_lib):
    m = meshgen.box_kuhn(10, 10, 10, 0.1, 0.1, 0.1)
    return m, oracle_lib.Mesh(m.nodes, m.tets, m.tris)


def _sim(O, om, charge=0.0, model="NONE", conc=0.0, dt=1e-6, E=None):
    k = O.species_constants(AR, charge, 4.17e-10, AR, S.VTI["Ar"], S.VSS_ALPHA["Ar"], O.MODEL[model])
    return O.Sim(om, [k], O.MODEL[model], conc, dt, 7, E=E if E is not None else np.zeros((om.n_tets, 3)))


def _one(O, om, p, v):
    p = np.asarray([p], dtype=np.float64)
    v = np.asarray([v], dtype=np.float64)
    return O.State.new(p, v, np.zeros(1, np.int32), np.zeros(1, np.uint64), om.locate_many(p[:, 0], p[:, 1], p[:, 2]))


def _floor_triangle(mesh, x, y):
    """Surface triangle on z = 0 that contains (x, y), by exact 2-D signs."""
    X, Y = Fraction(x), Fraction(y)
    hits = []
    for s, tri in enumerate(mesh.tris):
        P = mesh.nodes[tri]
        if np.any(P[:, 2] != 0.0):
            continue
        a, b, c = ([Fraction(float(P[i, 0])), Fraction(float(P[i, 1]))] for i in range(3))
        def side(p, q):
            return (q[0] - p[0]) * (Y - p[1]) - (q[1] - p[1]) * (X - p[0])
        d = [side(a, b), side(b, c), side(c, a)]
        if all(t > 0 for t in d) or all(t < 0 for t in d):
            hits.append(s)
    assert len(hits) == 1
    return hits[0]


def test_ballistic_flight_settles_on_the_exact_floor_triangle(box, oracle_lib):
    O = oracle_lib
    mesh, om = box
    sim = _sim(O, om)
    st = _one(O, om, (0.0512, 0.0537, 0.05), (0.0, 0.0, -1100.0))
    z = 0.05
    for k in range(60):
        sim.step(st, k)
        z = z + (-1100.0 * 1e-6)
        if z >= 0.0:
            assert st.status[0] == O.ACTIVE and st.z[0] == z, k
        else:
            break
    # the first step whose segment crosses z = 0 settles it, on that triangle,
    # at the end-of-step position (not clamped to the hit point)
    assert k == 45 and st.status[0] == O.SETTLED
    assert st.hit[0] == _floor_triangle(mesh, 0.0512, 0.0537)
    assert st.z[0] == z and st.x[0] == 0.0512 and st.y[0] == 0.0537
    assert sim.tri_count[st.hit[0]] == 1 and sim.tri_count.sum() == 1


def test_frozen_speed_bound_is_strict(box, oracle_lib):
    O = oracle_lib
    mesh, om = box
    sim = _sim(O, om)
    slow = _one(O, om, (0.05, 0.05, 0.05), (9.999999, 0.0, 0.0))
    assert sum(sim.step(slow, k) for k in range(5)) == 0 and slow.x[0] == 0.05
    edge = _one(O, om, (0.05, 0.05, 0.05), (10.0, 0.0, 0.0))  # |v| == 10 is not below the limit
    assert sum(sim.step(edge, k) for k in range(5)) == 5 and edge.x[0] > 0.05


def test_frozen_outside_expanded_bbox(box, oracle_lib):
    """bbox of the surface expanded x1.01 about its centre: [-0.0005, 0.1005]."""
    O = oracle_lib
    mesh, om = box
    sim = _sim(O, om)
    far = _one(O, om, (0.1006, 0.05, 0.05), (100.0, 0.0, 0.0))
    assert sim.step(far, 0) == 0 and far.x[0] == 0.1006
    near = _one(O, om, (0.1004, 0.05, 0.05), (100.0, 0.0, 0.0))
    assert near.tet[0] == -1
    assert sim.step(near, 0) == 1 and near.x[0] == 0.1004 + 100.0 * 1e-6


def test_uniform_field_closed_form(box, oracle_lib):
    O = oracle_lib
    mesh, om = box
    E = np.zeros((om.n_tets, 3))
    E[:, 2] = -1000.0
    q, dt = S.E_CHARGE, 1e-8
    sim = _sim(O, om, charge=q, dt=dt, E=E)
    v0 = np.array([300.0, -200.0, 500.0])
    st = _one(O, om, (0.05, 0.05, 0.05), v0)
    a = q * -1000.0 / AR
    x = np.array([0.05, 0.05, 0.05])
    for k in range(200):
        sim.step(st, k)
    vz = v0[2] + 200 * a * dt
    assert abs(st.vz[0] - vz) <= 1e-12 * abs(vz)
    assert st.vx[0] == v0[0] and st.vy[0] == v0[1]
    zk = x[2] + dt * sum(v0[2] + j * a * dt for j in range(1, 201))
    assert abs(st.z[0] - zk) <= 1e-12 * zk
    assert abs(st.x[0] - (x[0] + 200 * v0[0] * dt)) <= 1e-12 * x[0]


def test_hit_uses_the_pre_collision_segment(box, oracle_lib):
    """With P >= 1 the particle collides in the step it reaches the floor: the
    collision changes v, but the hit test uses the segment of the move."""
    O = oracle_lib
    mesh, om = box
    sim = _sim(O, om, model="HS", conc=1e40)
    st = _one(O, om, (0.0512, 0.0537, 0.0005), (0.0, 0.0, -1100.0))
    sim.step(st, 0)
    assert st.status[0] == O.SETTLED
    assert st.hit[0] == _floor_triangle(mesh, 0.0512, 0.0537)
    assert (st.vx[0], st.vy[0], st.vz[0]) != (0.0, 0.0, -1100.0)  # it did collide
    assert st.z[0] == 0.0005 + (-1100.0 * 1e-6)


def test_degenerate_segment_skips_collision(box, oracle_lib):
    """x + v dt == x: the step ends before the gas collision (:188-190)."""
    O = oracle_lib
    mesh, om = box
    sim = _sim(O, om, model="HS", conc=1e45, dt=1e-25)
    st = _one(O, om, (0.05, 0.05, 0.05), (11.0, 0.0, 0.0))
    assert sim.step(st, 0) == 1
    assert (st.vx[0], st.vy[0], st.vz[0]) == (11.0, 0.0, 0.0) and st.x[0] == 0.05
    moving = _one(O, om, (0.05, 0.05, 0.05), (11.0, 0.0, 0.0))
    sim2 = _sim(O, om, model="HS", conc=1e45, dt=1e-15)
    sim2.step(moving, 0)
    assert (moving.vx[0], moving.vy[0], moving.vz[0]) != (11.0, 0.0, 0.0)


def test_movement_record_mask(box, oracle_lib):
    """The particles whose start position a step records (pushed and moved,
    ParticleDynamicsProcessor.cpp:182-184): not the settled, the slow
    (|v| < 10 m/s) or those outside the expanded bbox; a degenerate segment is
    still recorded (the record comes before the move check)."""
    O = oracle_lib
    mesh, om = box
    p = np.array([[0.05, 0.05, 0.05], [0.05, 0.05, 0.05], [0.1006, 0.05, 0.05], [0.03, 0.04, 0.0005],
                  [0.06, 0.02, 0.07]])
    v = np.array([[11.0, 0.0, 0.0], [9.0, 0.0, 0.0], [100.0, 0.0, 0.0], [0.0, 0.0, -1100.0],
                  [200.0, 50.0, 0.0]])
    st = O.State.new(p, v, np.zeros(5, np.int32), np.arange(5, dtype=np.uint64),
                     om.locate_many(p[:, 0], p[:, 1], p[:, 2]))
    sim = _sim(O, om)
    moved = np.full(5, 7, dtype=np.uint8)
    assert sim.step(st, 0, moved=moved) == 3
    assert list(moved) == [1, 0, 0, 1, 1]
    assert st.status[3] == O.SETTLED  # particle 3 settled in step 0 ...
    assert sim.step(st, 1, moved=moved) == 2
    assert list(moved) == [1, 0, 0, 0, 1]  # ... and is not recorded again
    tiny = _one(O, om, (0.05, 0.05, 0.05), (11.0, 0.0, 0.0))
    m1 = np.zeros(1, dtype=np.uint8)
    assert _sim(O, om, dt=1e-25).step(tiny, 0, moved=m1) == 1 and m1[0] == 1 and tiny.x[0] == 0.05
