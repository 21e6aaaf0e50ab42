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

// B = 0 push: v = (v + h) + h, h = (((E*q)/m)*dt)/2 per component; skipped when
// dt == 0 or E == 0 exactly (Particle.cpp:199-207).
template <typename R>
__device__ inline void push(R ex, R ey, R ez, R q, R m, R dt, R& vx, R& vy, R& vz)
{
    if (dt == R(0)) return;
    if (ex == R(0) && ey == R(0) && ez == R(0)) return;
    R hx = (((ex * q) / m) * dt) / R(2);
    R hy = (((ey * q) / m) * dt) / R(2);
    R hz = (((ez * q) / m) * dt) / R(2);
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
__device__ inline double m_pow<double>(double x, double y) { return pow(x, y); }
template <>
__device__ inline float m_pow<float>(float x, float y) { return powf(x, y); }
template <typename R>
__device__ inline void m_sincos(R x, R* s, R* c);
template <>
__device__ inline void m_sincos<double>(double x, double* s, double* c) { *s = sin(x); *c = cos(x); }
template <>
__device__ inline void m_sincos<float>(float x, float* s, float* c) { *s = sinf(x); *c = cosf(x); }
template <typename R>
__device__ inline R m_acos(R x);
template <>
__device__ inline double m_acos<double>(double x) { return acos(x); }
template <>
__device__ inline float m_acos<float>(float x) { return acosf(x); }
template <typename R>
__device__ inline R m_atan2(R y, R x);
template <>
__device__ inline double m_atan2<double>(double y, double x) { return atan2(y, x); }
template <>
__device__ inline float m_atan2<float>(float y, float x) { return atan2f(y, x); }

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

// Post-collision velocity for variates u2 (cos chi) and u3 (phi).
// FP64 follows the reference expressions operation by operation.  FP32 uses
// algebraically identical forms that avoid cancellation: sin(chi) as
// sqrt((1-c)(1+c)), and the VSS rotation by beta = acos(pz/|p|),
// gamma = atan2(py, px) written with the direction cosines of p.
template <typename R>
__device__ inline void scatter_velocity(int model, const CollideK<R>& k, R u2, R u3, R& vx, R& vy, R& vz)
{
    constexpr bool kF64 = sizeof(R) == 8;
    R xi_cos;
    if (model == 1) xi_cos = (u2 * (R(1) - (R(-1)))) + (R(-1));       // rng(-1, 1)
    else if (model == 2) xi_cos = (u2 * (R(1) - R(0))) + R(0);         // rng()
    else xi_cos = R(2) * m_pow<R>(u2, k.inv_alpha) - R(1);             // 2*pow(rng(),1/a)-1
    R xi_sin = kF64 ? m_sqrt<R>(R(1) - xi_cos * xi_cos) : m_sqrt<R>((R(1) - xi_cos) * (R(1) + xi_cos));
    R phi = (u3 * (R(2) * R(kPi) - R(0))) + R(0);                       // rng(0, 2*PI)
    R sphi, cphi;
    m_sincos<R>(phi, &sphi, &cphi);
    R x = xi_sin * cphi, y = xi_sin * sphi, z = xi_cos;
    R cmx = vx * k.mass_cp, cmy = vy * k.mass_cp, cmz = vz * k.mass_cp;
    R px = vx * k.mass_ct, py = vy * k.mass_ct, pz = vz * k.mass_ct;
    R mp = m_sqrt<R>(px * px + py * py + pz * pz);
    R dx = x * mp, dy = y * mp, dz = z * mp;
    if (model == 3) {
        // calcBetaGamma + rotation (GeometryVector.hpp:448-474, :66-90)
        R sb, cb, sgm, cg;
        if (kF64) {
            R beta = m_acos<R>(pz / mp), gamma = m_atan2<R>(py, px);
            m_sincos<R>(beta, &sb, &cb);
            m_sincos<R>(gamma, &sgm, &cg);
        } else {
            R rho = m_sqrt<R>(px * px + py * py);
            cb = pz / mp;
            sb = rho / mp;
            if (rho > R(0)) {
                cg = px / rho;
                sgm = py / rho;
            } else {  // atan2(+-0, px): gamma = 0 or pi
                cg = px < R(0) ? R(-1) : R(1);
                sgm = R(0);
            }
        }
        R tx = cb * dx + sb * dz;
        R tz = -sb * dx + cb * dz;
        dx = tx;
        dz = tz;
        tx = cg * dx - sgm * dy;
        R ty = sgm * dx + cg * dy;
        dx = tx;
        dy = ty;
    }
    vx = dx + cmx;
    vy = dy + cmy;
    vz = dz + cmz;
}

}  // namespace sg
