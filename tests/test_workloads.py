"""Synthetic workload generators (CPU): the chamber mesh and the Thompson
energy sampler against closed forms."""
import numpy as np

from start_amd import meshgen, workloads
from oracle import pyoracle as O


def test_concentric_map_is_area_preserving():
    """Shirley-Chiu map: Jacobian pi/4 everywhere (area of the unit disk over
    the square [-1,1]^2), by central differences at random points."""
    rng = np.random.default_rng(3)
    a, b = rng.uniform(-0.99, 0.99, (2, 2000))
    h = 1e-6
    xa = np.stack(meshgen.concentric_disk(a + h, b)) - np.stack(meshgen.concentric_disk(a - h, b))
    xb = np.stack(meshgen.concentric_disk(a, b + h)) - np.stack(meshgen.concentric_disk(a, b - h))
    jac = (xa[0] * xb[1] - xa[1] * xb[0]) / (4 * h * h)
    keep = np.abs(np.abs(a) - np.abs(b)) > 1e-4  # the map is only piecewise smooth
    assert np.allclose(jac[keep], np.pi / 4, rtol=1e-6)
    r = np.hypot(*meshgen.concentric_disk(np.ones(50), np.linspace(-1, 1, 50)))
    assert np.allclose(r, 1.0)  # the square's edge goes to the circle


def test_chamber_mesh():
    """Reduced chamber: positive tets, volume close to pi r^2 h (inscribed
    polygonal section), boundary faces + substrate faces all distinct mesh
    faces, substrate triangles on the plane z = 0.1 within r <= 0.1325, target
    triangles on the top disk, and the oracle accepts the mesh."""
    ch = meshgen.cylinder_chamber(n=12, nz=12)
    m = ch.mesh
    p = m.nodes[m.tets]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])) / 6
    assert (vol > 0).all()
    exact = np.pi * 0.15 ** 2 * 0.3
    assert 0.97 * exact < vol.sum() < exact
    tri = m.nodes[m.tris]
    sub = tri[ch.substrate]
    assert np.allclose(sub[:, :, 2], 0.1, atol=1e-15)
    assert (np.hypot(sub[:, :, 0], sub[:, :, 1]).mean(axis=1) <= 0.1325).all()
    tgt = tri[ch.target]
    assert np.all(tgt[:, :, 2] == 0.3) and (np.hypot(tgt[:, :, 0], tgt[:, :, 1]).mean(axis=1) <= 0.05).all()
    key = np.sort(m.tris, axis=1)
    assert len(np.unique(key, axis=0)) == len(key)
    om = O.Mesh(m.nodes, m.tets, m.tris)
    assert om.n_tets == len(m.tets)


def test_thompson_energy_against_closed_form():
    """Thompson f(E) ~ E/(E+U)^3 truncated at M: the sample mean and the
    fraction below U against the closed forms
      int_0^M f = M^2 / (2U (M+U)^2),
      int_0^M E f = ln((M+U)/U) + 2U/(M+U) - U^2/(2(M+U)^2) - 3/2,
      P(E < U) = (U/(2U))^2 / (M/(M+U))^2."""
    U, M = 4.89, 100.0
    e = workloads.thompson_energy(2_000_000, np.random.default_rng(8), U, M)
    assert e.min() >= 0.0 and e.max() <= M
    norm = M * M / (2 * U * (M + U) ** 2)
    first = np.log((M + U) / U) + 2 * U / (M + U) - U * U / (2 * (M + U) ** 2) - 1.5
    mean = first / norm
    assert abs(e.mean() - mean) < 4 * e.std() / np.sqrt(len(e))
    p_below = 0.25 / (M / (M + U)) ** 2
    assert abs((e < U).mean() - p_below) < 4 * np.sqrt(p_below * (1 - p_below) / len(e))


def test_chamber_particles():
    """Ti start just under the target disk moving down (cosine law: E[cos] =
    2/3), Ar+ inside the cylinder; the potential is -300 V +- 20% at the
    target centre and 0 at the wall, refreshed in blocks of 10 steps."""
    w = workloads.chamber(n_particles=40_000, n=12, nz=12)
    ti = w.sp_idx == 0
    assert ti.sum() == 20_000
    assert np.all(np.hypot(w.pos[ti, 0], w.pos[ti, 1]) <= 0.05)
    assert np.all(w.vel[ti, 2] < 0)
    c = -w.vel[ti, 2] / np.linalg.norm(w.vel[ti], axis=1)
    assert abs(c.mean() - 2 / 3) < 0.01
    ar = ~ti
    assert np.all(np.hypot(w.pos[ar, 0], w.pos[ar, 1]) < 0.15) and np.all((w.pos[ar, 2] > 0) & (w.pos[ar, 2] < 0.3))
    phi = w.potential(0)
    top = np.argmin(np.hypot(w.mesh.nodes[:, 0], w.mesh.nodes[:, 1]) + np.abs(w.mesh.nodes[:, 2] - 0.3))
    assert phi[top] == -300.0
    wall = np.hypot(w.mesh.nodes[:, 0], w.mesh.nodes[:, 1]) >= 0.15 * (1 - 1e-12)
    assert np.all(phi[wall] == 0.0)
    assert np.array_equal(w.potential(10), w.potential(19))
    assert not np.array_equal(w.potential(10), w.potential(20))
