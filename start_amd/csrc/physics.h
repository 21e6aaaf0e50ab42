// Per-particle physics shared by the FP64 and mixed-precision step kernels.
//
//  * Philox4x32-10 counter-based variates (Salmon et al., SC'11) keyed by
//    (seed, run, particle load index, step, draw).  They replace the
//    reference's per-call random_device-seeded mt19937 (HSModel.cpp:18,
//    RealNumberGeneratorHost.cpp:14) so CPU and GPU see identical variates.
//  * push: Particle::electroMagneticPush with B = 0 (Particle.cpp:196-234).
//  * collide: HSModel.cpp:6-42, VHSModel.cpp:6-52, VSSModel.cpp:6-54.
//
// Arithmetic follows the reference's evaluation order operation by operation;
// the translation unit is compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

// pow, sin, cos, acos and atan2 of the FP64 collision path, the same code as
// the oracle's (dm_* functions)
#define DM_FN __device__ inline
#define DM_TABLE static __constant__ const
#include "detmath.h"

namespace sg {

constexpr double kPi = 3.14159265358979323846;  // START_PI_NUMBER

struct SpeciesK {
    double mass, charge;
    double sigma_hs;   // PI * pow(d, 2)                      HSModel.cpp:12
    double vhs_c;      // pow(d_ref, 2) / tgamma(5/2 - omega)  VHSModel.cpp:19
    double two_kt;     // 2 * KT_reference                     VHSModel.cpp:20
    double m_r;        // p_mass*t_mass/(p_mass+t_mass)        VHSModel.cpp:15
    double vhs_exp;    // omega - 1/2                          VHSModel.cpp:16
    double mass_cp;    // p_mass/(t_mass+p_mass)               HSModel.cpp:26
    double mass_ct;    // t_mass/(t_mass+p_mass)               HSModel.cpp:27
    double inv_alpha;  // 1./alpha                             VSSModel.cpp:32
    // Push of a single simulation: >= 0 the slot of this species' table of
    // push increments per tet; -1 computed in the kernel; -2 charge 0 (the
    // increment is E*q, a signed zero)
    int32_t push_slot;
    int32_t pad;
};

__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Two 53-bit uniforms in [0,1): counter {pid, 0, step, draw}, key {seed_lo, seed_hi ^ run}.
__host__ __device__ inline void variates(uint64_t seed, uint32_t run, uint32_t pid, uint32_t step,
                                         uint32_t draw, double& ua, double& ub)
{
    uint32_t w[4];
    philox4x32_10(pid, 0u, step, draw, (uint32_t)seed, (uint32_t)(seed >> 32) ^ run, w);
    const double two_m53 = 1.0 / 9007199254740992.0;
    uint64_t a = ((uint64_t)w[0] << 21) | (w[1] >> 11);
    uint64_t b = ((uint64_t)w[2] << 21) | (w[3] >> 11);
    ua = (double)a * two_m53;
    ub = (double)b * two_m53;
}

// x rounded to FP32 stochastically: to the farther of the two neighbouring
// floats with probability |x - nearest| / gap, resolved to 2^-10 by r10 (a
// uniform 10-bit integer), so that the rounded value is x in expectation to
// 2^-11 of a gap.  Repeated nearest rounding of x + d with a slowly changing
// step d rounds the same way step after step and drifts linearly; this drifts
// as a random walk.  (x is a local coordinate, far from the FP32 subnormals;
// f == 0 with x != 0 gives a NaN g, and the comparison then keeps f.)
__device__ inline float round_stochastic(double x, uint32_t r10)
{
    const float f = (float)x;
    const double rem = x - (double)f;  // exact, |rem| <= gap / 2
    const uint32_t fb = __float_as_uint(f);
    const float g = __uint_as_float(((rem > 0.0) == (f > 0.0f)) ? fb + 1u : fb - 1u);  // next float toward x
    const double gap = fabs((double)g - (double)f);
    return ((double)r10 + 0.5) * gap < fabs(rem) * 1024.0 ? g : f;
}

// 32 hashed bits of (a, b, c): lowbias32 (C. Wellons' integer hash, public
// domain) of a linear mix.  The mixed path draws its rounding bits from it.
__device__ inline uint32_t hash3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t h = a * 0x9E3779B1u + b * 0x85EBCA77u + c;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// B = 0 push: v = (v + h) + h, h = (((E*q)/m)*dt)/2 per component; skipped when
// dt == 0 or E == 0 exactly (Particle.cpp:199-207).  h in R, the velocity
// (always FP64 state) accumulated in V.
template <typename R, typename V>
__device__ inline void push(R ex, R ey, R ez, R q, R m, R dt, V& vx, V& vy, V& vz)
{
    if (dt == R(0)) return;
    if (ex == R(0) && ey == R(0) && ez == R(0)) return;
    V hx = (V)((((ex * q) / m) * dt) / R(2));
    V hy = (V)((((ey * q) / m) * dt) / R(2));
    V hz = (V)((((ez * q) / m) * dt) / R(2));
    vx = (vx + hx) + hx;
    vy = (vy + hy) + hy;
    vz = (vz + hz) + hz;
}

template <typename R>
struct CollideK {
    R sigma_hs, vhs_c, two_kt, m_r, vhs_exp, mass_cp, mass_ct, inv_alpha;
};

template <typename R>
__device__ inline R m_sqrt(R x);
template <>
__device__ inline double m_sqrt<double>(double x) { return sqrt(x); }
template <>
__device__ inline float m_sqrt<float>(float x) { return sqrtf(x); }
// next representable value from x towards y (x itself when y == x)
__device__ inline double m_nextafter(double x, double y) { return nextafter(x, y); }
__device__ inline float m_nextafter(float x, float y) { return nextafterf(x, y); }
template <typename R>
__device__ inline R m_pow(R x, R y);
template <>
__device__ inline double m_pow<double>(double x, double y) { return dm_pow(x, y); }
template <>
__device__ inline float m_pow<float>(float x, float y) { return powf(x, y); }

// Collision probability P = ((sigma*|v|)*n)*dt (HSModel.cpp:15, VHSModel.cpp:24-25).
template <typename R>
__device__ inline R collision_probability(int model, const CollideK<R>& k, R n, R dt, R vmod)
{
    R sigma;
    if (model == 1) {
        sigma = k.sigma_hs;
    } else {
        R d_vhs_2 = k.vhs_c * m_pow<R>(k.two_kt / ((k.m_r * vmod) * vmod), k.vhs_exp);
        sigma = R(kPi) * d_vhs_2;
    }
    return ((sigma * vmod) * n) * dt;
}

// Post-collision velocity for variates u2 (cos chi) and u3 (phi), in FP64 in
// both precisions, following the reference expressions operation by operation
// (HSModel.cpp:22-34, VHSModel.cpp:31-43, VSSModel.cpp:32-46).
__device__ inline void scatter_velocity(int model, double mass_cp, double mass_ct, double inv_alpha, double u2,
                                        double u3, double& vx, double& vy, double& vz)
{
    double xi_cos;
    if (model == 1) xi_cos = (u2 * (1.0 - (-1.0))) + (-1.0);        // rng(-1, 1)
    else if (model == 2) xi_cos = (u2 * (1.0 - 0.0)) + 0.0;         // rng()
    else xi_cos = 2.0 * m_pow<double>(u2, inv_alpha) - 1.0;          // 2*pow(rng(),1/a)-1
    double xi_sin = m_sqrt<double>(1.0 - xi_cos * xi_cos);
    double phi = (u3 * (2.0 * kPi - 0.0)) + 0.0;                      // rng(0, 2*PI)
    double sphi, cphi;
    dm_sincos(phi, &sphi, &cphi);
    double x = xi_sin * cphi, y = xi_sin * sphi, z = xi_cos;
    double cmx = vx * mass_cp, cmy = vy * mass_cp, cmz = vz * mass_cp;
    double px = vx * mass_ct, py = vy * mass_ct, pz = vz * mass_ct;
    double mp = m_sqrt<double>(px * px + py * py + pz * pz);
    double dx = x * mp, dy = y * mp, dz = z * mp;
    if (model == 3) {
        // calcBetaGamma + rotation (GeometryVector.hpp:448-474, :66-90)
        double sb, cb, sgm, cg;
        double beta = dm_acos(pz / mp), gamma = dm_atan2(py, px);
        dm_sincos(beta, &sb, &cb);
        dm_sincos(gamma, &sgm, &cg);
        double tx = cb * dx + sb * dz;
        double tz = -sb * dx + cb * dz;
        dx = tx;
        dz = tz;
        tx = cg * dx - sgm * dy;
        double ty = sgm * dx + cg * dy;
        dx = tx;
        dy = ty;
    }
    vx = dx + cmx;
    vy = dy + cmy;
    vz = dz + cmz;
}

}  // namespace sg
