"""start_locate_noalloc (points and owners staged in a static device array)
against start_locate (allocating) through the C ABI: two calls in a row, the
second with fewer points, plus the size limit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from start_amd import _lib, workloads
from start_amd.engine import ParticleInCellStep

from ._parity import log

pytestmark = pytest.mark.gpu


def _points(mesh, n, rng):
    """Uniform points in (and a little beyond) the mesh box, with mesh nodes
    and edge midpoints mixed in (ties between tets)."""
    lo, hi = mesh.nodes.min(0), mesh.nodes.max(0)
    span = hi - lo
    p = rng.uniform(lo - 0.05 * span, hi + 0.05 * span, size=(n, 3))
    k = n // 4
    p[:k] = mesh.nodes[rng.integers(0, len(mesh.nodes), k)]
    a = mesh.nodes[mesh.tets[rng.integers(0, len(mesh.tets), k), 0]]
    b = mesh.nodes[mesh.tets[rng.integers(0, len(mesh.tets), k), 1]]
    p[k:2 * k] = 0.5 * (a + b)
    return p


def test_locate_noalloc_matches_locate():
    w = workloads.small_box(n_particles=100)
    eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed)
    rng = np.random.default_rng(7)
    for n in (_lib.NOALLOC_LOCATE_MAX, 3_000):  # the second call is smaller
        log(f"no-alloc locate: small_box, {n} points")
        pts = _points(w.mesh, n, rng)
        want = eng.locate(pts)
        got = eng.locate_noalloc(pts)
        assert np.array_equal(got, want), (n, int((got != want).sum()))
        assert (want >= 0).any() and (want < 0).any()
    eng.close()


def test_locate_noalloc_rejects_oversized_calls():
    w = workloads.small_box(n_particles=100)
    eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed)
    n = _lib.NOALLOC_LOCATE_MAX + 1
    x = np.zeros(n)
    out = np.empty(n, dtype=np.int32)
    dp = x.ctypes.data_as(C.POINTER(C.c_double))
    rc = _lib.lib().start_locate_noalloc(eng.ctx, n, dp, dp, dp, out.ctypes.data_as(C.POINTER(C.c_int32)))
    assert rc == -1 and b"at most 131072" in _lib.lib().start_last_error(eng.ctx)
    # the context still works after the rejection
    pts = _points(w.mesh, 500, np.random.default_rng(3))
    assert np.array_equal(eng.locate_noalloc(pts), eng.locate(pts))
    eng.close()
