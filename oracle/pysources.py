"""Oracle restatement of the particle sources (test infrastructure: only tests
and the CPU baseline load it).

It follows the reference's point and surface generators
(ParticleGeneratorHost.cpp:33-159), its velocity construction
(ParticleDynamicUtils.hpp:52-96) and its displacer
(ParticleDisplacer.cpp:20-31, 48-134), plus the energy and direction samplers
start_generate offers.  The reference draws from random_device-seeded mt19937
and cuRAND (irreproducible); here, as on the device, every variate is
Philox4x32-10 with counter {particle, 1, source, draw} and key
{seed_lo, seed_hi}.

Pinned by tests/test_oracle_sources.py: the vectorised Philox below against
the C oracle's (itself pinned by the Random123 known answers); speeds against
sqrt(2 E / m); fan directions against the polar / azimuth rule; the reference
displacer example; per-cell counts; and the Thompson, lognormal, Maxwellian,
cosine, isotropic and cone laws against their closed forms.
"""
from __future__ import annotations

import numpy as np

from . import pyoracle as _O

EV_J = 1.60217733000001e-19  # Constants.hpp:15
PI = 3.14159265358979323846
M32 = np.uint64(0xFFFFFFFF)


# sin, cos, acos, atan2, log and exp through the C oracle: the shared
# deterministic functions the device also compiles, or the C library's in the
# oracle's libm mode (pyoracle.set_libm)
def _fn(name, x, y=None):
    scalar = np.ndim(x) == 0
    xa = np.atleast_1d(np.asarray(x, dtype=np.float64))
    ya = None if y is None else np.broadcast_to(np.asarray(y, dtype=np.float64), xa.shape)
    r = _O.math_eval(name, xa, ya, libm=_O.get_libm())
    return (r[0][0] if scalar else r[0], r[1][0] if scalar else r[1]) if name == "sincos" else \
        (r[0] if scalar else r)


def _sin(x):
    return _fn("sincos", x)[0]


def _cos(x):
    return _fn("sincos", x)[1]


def _log(x):
    return _fn("log", x)


def _exp(x):
    return _fn("exp", x)


def _acos(x):
    return _fn("acos", x)


def _atan2(y, x):
    return _fn("atan2", x, y)


def philox(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 on arrays of uint32 counters (Salmon et al., SC'11)."""
    c0, c1, c2, c3 = (np.asarray(a, dtype=np.uint64) & M32 for a in (c0, c1, c2, c3))
    k0 = np.uint64(k0) & M32
    k1 = np.uint64(k1) & M32
    for r in range(10):
        if r:
            k0 = (k0 + np.uint64(0x9E3779B9)) & M32
            k1 = (k1 + np.uint64(0xBB67AE85)) & M32
        p0 = np.uint64(0xD2511F53) * c0
        p1 = np.uint64(0xCD9E8D57) * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & M32
        hi1, lo1 = p1 >> np.uint64(32), p1 & M32
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return c0, c1, c2, c3


def uniforms(seed, i, s, draw):
    """Two 53-bit uniforms in [0,1) per particle index i (array)."""
    i = np.asarray(i, dtype=np.uint64)
    w0, w1, w2, w3 = philox(i, np.ones_like(i), np.full_like(i, s), np.full_like(i, draw),
                            seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    t = 1.0 / 9007199254740992.0
    ua = ((w0 << np.uint64(21)) | (w1 >> np.uint64(11))).astype(np.float64) * t
    ub = ((w2 << np.uint64(21)) | (w3 >> np.uint64(11))).astype(np.float64) * t
    return ua, ub


def cell_order(n_cells, s, seed):
    """Seeded Fisher-Yates order of a surface source's cells (counter {k, 2, s, 0})."""
    order = list(range(n_cells))
    for j in range(n_cells - 1, 0, -1):
        w0, w1, _, _ = philox([j], [2], [s], [0], seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
        u = float((int(w0[0]) << 21) | (int(w1[0]) >> 11)) / 9007199254740992.0
        r = min(int(u * (j + 1)), j)
        order[j], order[r] = order[r], order[j]
    return np.array(order, dtype=np.int64)


def generate(sources, masses, seed):
    """Positions (n, 3), velocities (n, 3), species (n,) in source order."""
    P, V, S = [], [], []
    first = 0
    for s, src in enumerate(sources):
        n = int(src["count"])
        m = masses[int(src.get("species", 0))]
        i = np.arange(first, first + n, dtype=np.uint64)
        first += n
        u0, u1 = uniforms(seed, i, s, 0)
        u2, u3 = uniforms(seed, i, s, 1)
        disp = float(src.get("displace", 0.0))
        if src.get("kind", "point") == "point":
            pos = np.tile(np.asarray(src.get("position", (0.0, 0.0, 0.0)), dtype=np.float64), (n, 1))
            t0 = np.full(n, float(src.get("theta", 0.0)))
            p0 = np.full(n, float(src.get("phi", 0.0)))
            st = _sin(t0)
            d = np.stack([st * _cos(p0), st * _sin(p0), _cos(t0)], axis=1)
        else:
            cen = np.asarray(src["centres"], dtype=np.float64).reshape(-1, 3)
            nor = np.asarray(src["normals"], dtype=np.float64).reshape(-1, 3)
            order = cell_order(len(cen), s, seed)
            cen, nor = cen[order], nor[order]
            nc = len(cen)
            q, r = n // nc, n % nc
            j = np.arange(n)
            cell = np.where(j < (q + 1) * r, j // (q + 1), r + (j - (q + 1) * r) // max(q, 1))
            pos = cen[cell].copy()
            nn = nor[cell]
            mag = np.sqrt(nn[:, 0] * nn[:, 0] + nn[:, 1] * nn[:, 1] + nn[:, 2] * nn[:, 2])
            t0 = _acos(nn[:, 2] / mag)
            p0 = _atan2(nn[:, 1] / mag, nn[:, 0] / mag)
            d = nn.copy()
        if disp != 0.0:  # ParticleDisplacer.cpp:20-31
            mc = np.max(np.abs(d), axis=1)
            ok = ~(mc < 2.220446049250313e-16)
            pos[ok] = pos[ok] + (d[ok] / mc[ok, None]) * disp
        energy = src.get("energy", "mono")
        if energy == "maxwellian":
            sig = np.sqrt((float(src["p0"]) * EV_J) / m)
            r1 = np.sqrt(-2.0 * _log(1.0 - u0))
            vel = np.stack([sig * (r1 * _cos(2.0 * PI * u1)), sig * (r1 * _sin(2.0 * PI * u1)),
                            sig * (np.sqrt(-2.0 * _log(1.0 - u2)) * _cos(2.0 * PI * u3))], axis=1)
        else:
            if energy == "mono":
                e = np.full(n, float(src["energy_eV"]))
            elif energy == "thompson":
                U, emax = float(src["p0"]), float(src["p1"])
                fr = emax / (emax + U)
                rr = np.sqrt(u0 * (fr * fr))
                e = U * rr / (1.0 - rr)
            else:
                mean, sl = float(src["p0"]), float(src["p1"])
                mu = _log(mean) - 0.5 * sl * sl
                e = _exp(mu + sl * (np.sqrt(-2.0 * _log(1.0 - u0)) * _cos(2.0 * PI * u1)))
            v = np.sqrt(2 * (e * EV_J) / m)
            angle = src.get("angle", "fan")
            if angle == "fan":
                a = float(src.get("expansion", 0.0))
                th = t0 + (-1.0 + 2.0 * u2) * a if a != 0.0 else t0
                vel = np.stack([v * _sin(th) * _cos(p0), v * _sin(th) * _sin(p0), v * _cos(th)], axis=1)
            else:
                if angle == "cosine":
                    ct = np.sqrt(u2)
                elif angle == "isotropic":
                    ct = 1.0 - 2.0 * u2
                else:
                    ct = 1.0 - u2 * (1.0 - _cos(float(src.get("expansion", 0.0))))
                ph = 2.0 * PI * u3
                stt = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
                a_, b_ = stt * _cos(ph), stt * _sin(ph)
                if angle == "isotropic":
                    vel = np.stack([v * a_, v * b_, v * ct], axis=1)
                else:
                    s0, c0, sp, cp = _sin(t0), _cos(t0), _sin(p0), _cos(p0)
                    vel = np.stack([v * ((a_ * (c0 * cp) - b_ * sp) + ct * (s0 * cp)),
                                    v * ((a_ * (c0 * sp) + b_ * cp) + ct * (s0 * sp)),
                                    v * (ct * c0 - a_ * s0)], axis=1)
        P.append(pos)
        V.append(vel)
        S.append(np.full(n, int(src.get("species", 0)), dtype=np.int32))
    return np.concatenate(P), np.concatenate(V), np.concatenate(S)
