"""Pins of the oracle's particle-source restatement (oracle/pysources.py)
against the C oracle's Philox, the reference's generator rules and closed-form
laws."""
import numpy as np
import pytest

from oracle import pyoracle as O
from oracle import pysources as PS
from start_amd.species import EV_J, MASS

AR, TI, CU = MASS["Ar"], MASS["Ti"], MASS["Cu"]


def test_vectorised_philox_matches_the_c_oracle():
    """The numpy Philox against or_philox4x32_10 (pinned by the Random123
    known answers) on random counters and keys."""
    rng = np.random.default_rng(2)
    for _ in range(20):
        ctr = [int(v) for v in rng.integers(0, 2**32, 4)]
        key = [int(v) for v in rng.integers(0, 2**32, 2)]
        got = [int(w[0]) for w in PS.philox([ctr[0]], [ctr[1]], [ctr[2]], [ctr[3]], key[0], key[1])]
        assert got == [int(v) for v in O.philox(ctr, key)]


def test_point_source_reference_rules():
    """ParticleGeneratorHost::fromPointSource + ParticleDynamicUtils: every
    particle at the source point, |v| = sqrt(2 E eV_J / m), polar angle within
    theta +- expansion, azimuth fixed at phi (ParticleGeneratorTests.cpp:38-55:
    two sources, 50 Ar at 10 eV and 30 He at 20 eV)."""
    srcs = [{"kind": "point", "species": 0, "count": 50, "energy_eV": 10.0, "theta": 2.0, "phi": 1.0,
             "expansion": 1.0, "position": (0.0, 0.0, 0.0)},
            {"kind": "point", "species": 1, "count": 30, "energy_eV": 20.0, "theta": 3.0, "phi": 2.0,
             "expansion": 2.0, "position": (10.0, 10.0, 10.0)}]
    masses = [AR, MASS["He"]]
    pos, vel, sp = PS.generate(srcs, masses, 7)
    assert len(pos) == 80 and (sp[:50] == 0).all() and (sp[50:] == 1).all()
    assert (pos[:50] == 0.0).all() and (pos[50:] == 10.0).all()
    for sl, E, m, th0, ph0, a in ((slice(0, 50), 10.0, AR, 2.0, 1.0, 1.0), (slice(50, 80), 20.0, MASS["He"], 3.0, 2.0, 2.0)):
        v = vel[sl]
        speed = np.linalg.norm(v, axis=1)
        assert np.allclose(speed, np.sqrt(2 * (E * EV_J) / m), rtol=1e-14)
        e_ev = 0.5 * m * speed ** 2 / EV_J
        assert np.all((e_ev > 0.999999 * E) & (e_ev < 1.000001 * E))
        # azimuth fixed: (vx, vy) parallel to (cos phi, sin phi)
        assert np.allclose(v[:, 1] * np.cos(ph0), v[:, 0] * np.sin(ph0), atol=1e-9 * speed.max())
        # polar angle th0 + u a, u in [-1, 1): recovered from sin = vx / (|v| cos phi), cos = vz / |v|
        th = np.arctan2(v[:, 0] / (speed * np.cos(ph0)), v[:, 2] / speed)
        off = (th - (th0 - a)) % (2 * np.pi)
        assert np.all(off < 2 * a + 1e-12)
        assert off.min() < 0.2 * a and off.max() > 1.8 * a  # spread over the fan


def test_displacer_reference_example():
    """ParticleDisplacer::_calculateDisplacement: d / max|d_i| * 0.05; a point
    source along -z moves by (0, 0, -0.05); a zero direction does not move."""
    src = [{"kind": "point", "count": 3, "energy_eV": 5.0, "theta": np.pi, "phi": 0.0, "position": (0.05, 0.05, 0.095),
            "displace": 0.05}]
    pos, _, _ = PS.generate(src, [TI], 1)
    assert np.allclose(pos, [[0.05 + 0.05 * np.sin(np.pi) / 1.0, 0.05, 0.045]] * 3, atol=1e-17)
    src = [{"kind": "surface", "count": 4, "energy_eV": 5.0, "centres": [[1.0, 2.0, 3.0]], "normals": [[0.0, 2.0, 1.0]],
            "displace": 0.05}]
    pos, vel, _ = PS.generate(src, [TI], 1)
    assert np.allclose(pos, [[1.0, 2.05, 3.025]] * 4, rtol=1e-15)


def test_surface_source_counts_and_directions():
    """count / n_cells per cell, the remainder one each to count % n_cells
    cells; with no expansion every velocity is along its cell's normal."""
    rng = np.random.default_rng(4)
    nc = 13
    cen = rng.uniform(-1, 1, (nc, 3))
    nor = rng.normal(size=(nc, 3))
    src = [{"kind": "surface", "count": 100, "energy_eV": 7.0, "centres": cen, "normals": nor}]
    pos, vel, _ = PS.generate(src, [CU], 11)
    counts = np.array([np.sum(np.all(pos == c, axis=1)) for c in cen])
    assert counts.sum() == 100 and set(counts) == {7, 8} and (counts == 8).sum() == 100 % nc
    nn = nor / np.linalg.norm(nor, axis=1)[:, None]
    idx = np.array([np.nonzero(np.all(cen == p, axis=1))[0][0] for p in pos])
    vh = vel / np.linalg.norm(vel, axis=1)[:, None]
    assert np.allclose(vh, nn[idx], atol=1e-14)
    assert not np.array_equal(PS.cell_order(nc, 0, 11), PS.cell_order(nc, 0, 12))  # seeded


def _src(energy, angle="fan", **kw):
    s = {"kind": "point", "count": 400_000, "energy": energy, "angle": angle, "theta": 0.7, "phi": -1.1}
    s.update(kw)
    return [s]


def test_thompson_closed_form():
    U, M = 4.89, 100.0
    _, vel, _ = PS.generate(_src("thompson", p0=U, p1=M), [TI], 3)
    e = 0.5 * TI * np.sum(vel * vel, axis=1) / EV_J
    norm = M * M / (2 * U * (M + U) ** 2)
    mean = (np.log((M + U) / U) + 2 * U / (M + U) - U * U / (2 * (M + U) ** 2) - 1.5) / norm
    assert abs(e.mean() - mean) < 4 * e.std() / np.sqrt(len(e))
    assert e.max() <= M * (1 + 1e-12)
    p = 0.25 / (M / (M + U)) ** 2  # P(E < U)
    assert abs((e < U).mean() - p) < 4 * np.sqrt(p * (1 - p) / len(e))


def test_lognormal_closed_form():
    _, vel, _ = PS.generate(_src("lognormal", p0=10.0, p1=0.5), [CU], 5)
    e = 0.5 * CU * np.sum(vel * vel, axis=1) / EV_J
    le = np.log(e)
    assert abs(le.std() - 0.5) < 0.005
    assert abs(e.mean() - 10.0) < 4 * e.std() / np.sqrt(len(e))


def test_maxwellian_closed_form():
    _, vel, _ = PS.generate(_src("maxwellian", p0=1.0), [AR], 6)
    var = 1.0 * EV_J / AR
    for k in range(3):
        assert abs(vel[:, k].mean()) < 4 * np.sqrt(var / len(vel))
        assert abs(vel[:, k].var() / var - 1.0) < 4 * np.sqrt(2.0 / len(vel))


def test_direction_laws():
    """About d(theta, phi): Lambert <cos> = 2/3, <cos^2> = 1/2; isotropic
    <z> = 0, <z^2> = 1/3; cone within half-angle a, <cos> = (1 + cos a)/2."""
    d = np.array([np.sin(0.7) * np.cos(-1.1), np.sin(0.7) * np.sin(-1.1), np.cos(0.7)])
    n = 400_000
    _, vel, _ = PS.generate(_src("mono", "cosine", energy_eV=3.0), [AR], 8)
    c = vel @ d / np.linalg.norm(vel, axis=1)
    assert c.min() >= -1e-12 and abs(c.mean() - 2 / 3) < 4 * np.sqrt(1 / 18 / n)
    assert abs((c * c).mean() - 0.5) < 0.003
    _, vel, _ = PS.generate(_src("mono", "isotropic", energy_eV=3.0), [AR], 9)
    z = vel[:, 2] / np.linalg.norm(vel, axis=1)
    assert abs(z.mean()) < 4 * np.sqrt(1 / 3 / n) and abs((z * z).mean() - 1 / 3) < 0.003
    a = 0.3
    _, vel, _ = PS.generate(_src("mono", "cone", energy_eV=3.0, expansion=a), [AR], 10)
    c = vel @ d / np.linalg.norm(vel, axis=1)
    assert c.min() >= np.cos(a) - 1e-12
    assert abs(c.mean() - (1 + np.cos(a)) / 2) < 4 * (1 - np.cos(a)) / np.sqrt(12 * n)
    speed = np.linalg.norm(vel, axis=1)
    assert np.allclose(speed, np.sqrt(2 * (3.0 * EV_J) / AR), rtol=1e-14)


@pytest.mark.parametrize("count,ncells", [(5, 13), (13, 13), (27, 13)])
def test_surface_counts_small(count, ncells):
    cen = np.arange(3 * ncells, dtype=np.float64).reshape(-1, 3)
    nor = np.tile([0.0, 0.0, 1.0], (ncells, 1))
    pos, _, _ = PS.generate([{"kind": "surface", "count": count, "energy_eV": 1.0, "centres": cen, "normals": nor}],
                            [AR], 2)
    counts = np.array([np.sum(np.all(pos == c, axis=1)) for c in cen])
    q, r = divmod(count, ncells)
    assert counts.sum() == count and (counts == q + 1).sum() == r and (counts == q).sum() == ncells - r
