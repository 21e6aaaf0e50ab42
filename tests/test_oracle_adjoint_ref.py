"""Pins the FP64 reference gradient (forward-mode tangents along the oracle's
discrete map) against central finite differences of the same map, per
particle, on a small parallel-plate case (no GPU).

The map is piecewise smooth: where V +- h settles a particle in a different
step than V, the hit point is continuous but its slope jumps, and a central
difference straddles the kink.  The comparison therefore uses the particles
that settle on the bottom plate in the same step in all three runs."""
from __future__ import annotations

import numpy as np

from start_amd import workloads

from . import _adjoint_ref as AR


def test_tangent_matches_central_differences(oracle_lib):
    w = workloads.plates(n_particles=3000, nx=11, ny=11, nz=5, steps=700, dt=1e-8)
    V = np.array([150.0, 30.0])
    h = 1e-3
    base = AR.run(w, V)
    sel = (base["tri"] >= 0) & w.bottom[np.clip(base["tri"], 0, None)]
    fd = np.zeros((len(w.pos), 3, 2))
    for j in range(2):
        e = np.zeros(2)
        e[j] = h
        up, dn = AR.run(w, V + e), AR.run(w, V - e)
        same = (up["tri"] >= 0) & (dn["tri"] >= 0) & w.bottom[np.clip(up["tri"], 0, None)] \
            & w.bottom[np.clip(dn["tri"], 0, None)] & (up["step"] == base["step"]) & (dn["step"] == base["step"])
        sel &= same
        fd[:, :, j] = (up["xh"] - dn["xh"]) / (2 * h)
    assert sel.sum() > 2500
    t = base["dxh"][sel]
    err = np.abs(fd[sel] - t).max() / np.abs(t).max()
    assert err < 1e-6, err
    # the z component of a point on the bottom plate does not move
    assert np.abs(t[:, 2, :]).max() <= 1e-12 * np.abs(t).max()
    # the potentials enter only through their difference for a uniform field
    assert np.allclose(t[:, :, 0], -t[:, :, 1], rtol=1e-9, atol=1e-12 * np.abs(t).max())
