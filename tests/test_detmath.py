"""The shared FP64 elementary functions of the collision path
(start_amd/csrc/detmath.h), compiled here into the oracle: their constants
regenerate from their definitions, they stay within 1 ulp (sin, cos, pow,
log) or 2 ulps (acos, atan2) of the C library's on the argument ranges the
step uses, they give the C library's special values, and the oracle run with
them agrees with the oracle run with the C library's on whole simulations
(owners, hits and settles equal; positions and velocities to 1e-13)."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from start_amd import meshgen, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = 400_000


def ulps(a, b):
    """|a - b| in units of the last place of b (inf where exactly one is NaN)."""
    a, b = np.asarray(a), np.asarray(b)
    u = np.spacing(np.abs(b))
    with np.errstate(invalid="ignore"):
        d = np.where(a == b, 0.0, np.abs(a - b) / u)
    return np.where(np.isnan(a) != np.isnan(b), np.inf, np.where(np.isnan(a) & np.isnan(b), 0.0, d))


def both(fn, x, y=None):
    return O.math_eval(fn, x, y, libm=False), O.math_eval(fn, x, y, libm=True)


def test_constants_regenerate():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "start_amd", "csrc", "gen_detmath.py"), "--check"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_sincos_within_one_ulp():
    rng = np.random.default_rng(1)
    # phi = u * 2 pi, beta in [0, pi], gamma in [-pi, pi], and the edges of the reduction
    x = np.concatenate([rng.uniform(-7.0, 7.0, N), rng.uniform(0.0, 2 * np.pi, N),
                        np.pi / 4 * np.arange(-9, 10), np.nextafter(np.pi / 4 * np.arange(-9, 10), 9.0),
                        [0.0, -0.0, 1e-300, -1e-20, np.pi, np.pi / 2, 2 * np.pi, 31.0]])
    (s, c), (ls, lc) = both("sincos", x)
    assert ulps(s, ls).max() <= 1.0 and ulps(c, lc).max() <= 1.0
    assert (s == ls).mean() > 0.95 and (c == lc).mean() > 0.95
    assert np.signbit(O.math_eval("sincos", np.array([-0.0]))[0][0])


def test_pow_within_one_ulp():
    rng = np.random.default_rng(2)
    # VSS: u2^(1/alpha), u2 in [0, 1); VHS: (2kT / (m_r v^2))^(omega - 1/2) over many decades
    u = np.concatenate([rng.uniform(0.0, 1.0, N), 2.0 ** -rng.uniform(0, 53, N // 4)])
    a = rng.choice([1 / 1.4, 1 / 1.0, 1 / 1.6, 0.5], len(u))
    b = 10.0 ** rng.uniform(-12, 12, N)
    e = rng.uniform(-0.6, 0.6, N)
    for x, y in ((u, a), (b, e)):
        p, lp = both("pow", x, y)
        assert ulps(p, lp).max() <= 1.0
        assert (p == lp).mean() > 0.85
    sx = np.array([0.0, 0.0, np.inf, np.inf, 1.0, 2.0, 0.25, 1e-300, np.nan, 3.0, -1.5, -0.0])
    sy = np.array([0.3, -0.3, 0.3, -0.3, 0.7, 10.0, 0.5, 0.7, 0.5, 0.0, 0.3, 0.7])
    p, lp = both("pow", sx, sy)
    assert np.array_equal(p, lp, equal_nan=True)


def test_log_within_one_ulp():
    rng = np.random.default_rng(3)
    x = np.concatenate([1.0 - rng.uniform(0.0, 1.0, N), 10.0 ** rng.uniform(-300, 300, N),
                        [1.0, 2.0, 0.5, 5e-324, np.inf, 0.0, np.sqrt(0.5), np.nextafter(np.sqrt(0.5), 0)]])
    v, lv = both("log", x)
    assert ulps(v, lv).max() <= 1.0
    assert np.array_equal(both("log", np.array([-1.0]))[0], [np.nan], equal_nan=True)


def test_acos_atan2_within_two_ulps():
    rng = np.random.default_rng(4)
    c = np.concatenate([rng.uniform(-1.0, 1.0, N), 1.0 - 10.0 ** -rng.uniform(1, 16, N // 4),
                        -1.0 + 10.0 ** -rng.uniform(1, 16, N // 4), [-1.0, 1.0, 0.0, -0.0, 0.5, -0.5]])
    v, lv = both("acos", c)
    assert ulps(v, lv).max() <= 2.0 and (v == lv).mean() > 0.85
    y = rng.standard_normal(N) * 10.0 ** rng.uniform(-5, 5, N)
    x = rng.standard_normal(N) * 10.0 ** rng.uniform(-5, 5, N)
    v, lv = both("atan2", x, y)
    assert ulps(v, lv).max() <= 2.0 and (v == lv).mean() > 0.85
    # signed zeros and the axes, as the C library gives them
    sy = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1.0, -1.0, 2.0, -3.0])
    sx = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 0.0, -0.0, -0.0, 2.0, -3.0])
    v, lv = both("atan2", sx, sy)
    assert np.array_equal(v, lv) and np.array_equal(np.signbit(v), np.signbit(lv))


def test_scatter_identities():
    """What the mathematics fixes: sin^2 + cos^2 = 1, cos(acos c) = c,
    atan2(sin t, cos t) = t on (-pi, pi], pow(pow(x, y), 1/y) = x."""
    rng = np.random.default_rng(5)
    t = rng.uniform(-np.pi, np.pi, N)
    s, c = O.math_eval("sincos", t)
    assert np.abs(s * s + c * c - 1.0).max() < 4e-16
    assert np.abs(O.math_eval("atan2", c, s) - t).max() <= 4 * np.spacing(np.pi)
    cc = rng.uniform(-1.0, 1.0, N)
    assert np.abs(O.math_eval("sincos", O.math_eval("acos", cc))[1] - cc).max() < 1e-15
    x = rng.uniform(0.01, 100.0, N)
    y = rng.uniform(0.2, 3.0, N)
    back = O.math_eval("pow", O.math_eval("pow", x, y), 1.0 / y)
    assert (np.abs(back - x) / x).max() < 2e-14


@pytest.mark.parametrize("model", ["HS", "VSS", "VHS"])
def test_oracle_with_libm_and_shared_math_agree(model):
    """Whole runs: the oracle with the shared functions against the oracle
    with the C library's (the reference's calls).  Collisions are frequent
    (high gas density), so many thousand scatterings go through both."""
    w = workloads.small_box(n_particles=4000, charge_units=2.0)
    mesh = w.mesh
    om = O.Mesh(mesh.nodes, mesh.tets, mesh.tris)
    E = om.tet_field(meshgen.linear_potential(mesh, 0.0, 200.0))
    g = w.gas
    ks = [O.species_constants(s.mass, s.charge, s.diameter, g.mass, g.omega, g.alpha, O.MODEL[model])
          for s in w.species]
    tet0 = om.locate_many(w.pos[:, 0], w.pos[:, 1], w.pos[:, 2])
    runs = []
    for libm in (False, True):
        O.set_libm(libm)
        try:
            st = O.State.new(w.pos, w.vel, w.sp_idx, np.arange(len(w.pos), dtype=np.uint64), tet0)
            sim = O.Sim(om, ks, O.MODEL[model], 2e21, w.dt, w.seed, E=E)
            q = np.zeros(mesh.n_nodes)
            for k in range(150):
                sim.step(st, k, q, nthreads=8)
            runs.append((st, q, sim.tri_count.copy()))
        finally:
            O.set_libm(False)
    (a, qa, ta), (b, qb, tb) = runs
    for key in ("tet", "status", "hit"):
        assert np.array_equal(getattr(a, key), getattr(b, key)), key
    assert np.array_equal(ta, tb)
    xa, xb = np.stack([a.x, a.y, a.z], 1), np.stack([b.x, b.y, b.z], 1)
    va, vb = np.stack([a.vx, a.vy, a.vz], 1), np.stack([b.vx, b.vy, b.vz], 1)
    assert (np.abs(xa - xb).max(1) / np.linalg.norm(xb, axis=1)).max() <= 1e-13
    assert (np.abs(va - vb).max(1) / np.linalg.norm(vb, axis=1)).max() <= 1e-13
    assert not np.array_equal(va, vb)  # the two libraries do differ in the last bits
    assert np.abs(qa - qb).max() <= 1e-12 * np.abs(qb).max()
    assert (a.status == 1).sum() > 200


def test_oracle_defaults_to_the_shared_functions():
    """The GPU parity tests compare against the oracle's default mode, the
    shared functions; or_set_libm switches to the C library and back."""
    assert O.get_libm() is False
    x = np.array([0.3, 1.7, 2.9])
    shared = O.math_eval("sincos", x)[0]
    O.set_libm(True)
    try:
        assert O.get_libm() is True
    finally:
        O.set_libm(False)
    assert O.get_libm() is False
    assert np.array_equal(O.math_eval("sincos", x)[0], shared)
