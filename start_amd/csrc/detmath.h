/* Deterministic FP64 elementary functions: pow, log, sin and cos, acos and
 * atan2, written once and compiled into both the HIP kernels and the CPU
 * oracle, so that the collision scattering and the particle sources produce
 * the same bits on the GPU and in the oracle.  (The device math library and
 * the host's libm are each within an ulp or so of the true value, but not
 * always the same ulp.)
 *
 * The includer defines DM_FN (the qualifiers of a function: "static inline"
 * in C, "__device__ inline" in HIP) and DM_TABLE (those of a constant table:
 * "static const" in C, "static __constant__ const" in HIP), makes fma, rint,
 * frexp, ldexp, copysign, fabs, sqrt and signbit available, and compiles with
 * -ffp-contract=off: every + and * below rounds on its own, and fma is written
 * where one is meant (every Horner step is one).  All other operations used are exact or correctly
 * rounded IEEE operations, the same on both sides.
 *
 * Methods (constants from gen_detmath.py beside this file):
 *  - sin, cos: Cody-Waite reduction by pi/2 in three parts to r = rh + rl,
 *    |r| <= pi/4, then the Taylor series to r^17 (sin) and r^18 (cos);
 *    |x| <= 32.
 *  - atan2: t = min(|x|,|y|)/max(|x|,|y|) in [0, 1], c = round(8t)/8,
 *    atan t = atan c + atan s with s = (t - c)/(1 + t c) (|s| <= 1/16), the
 *    Taylor series of atan s to s^17, and atan c (and pi/2 -, pi -, pi/2 +
 *    it, for the other octants) from a table as hi + lo; finite arguments.
 *  - acos x = atan2(sqrt((1 - x)(1 + x)), x).
 *  - log: x = f 2^e with f in [sqrt(1/2), sqrt(2)), log f = 2 atanh q with
 *    q = (f - 1)/(f + 1) (|q| <= 0.172), the series to q^23, all as a
 *    double-double; e ln 2 with ln 2 split in two.
 *  - exp x: r = x - k ln 2 (|r| <= ln 2 / 2), the Taylor series to r^14, and
 *    2^k exactly.
 *  - pow x^y (x >= 0) = exp(y log x) with y log x as a double-double.
 * Measured against glibc by tests/test_detmath.py. */
#ifndef START_DETMATH_H
#define START_DETMATH_H

#if !defined(DM_FN) || !defined(DM_TABLE)
#error "define DM_FN and DM_TABLE before including detmath.h"
#endif

/* BEGIN generated constants */
#define DM_PIO2_1 (0x1.921fb54400000p+0)  /* pi/2, first 33 bits */
#define DM_PIO2_2 (0x1.0b4611a600000p-34)  /* pi/2, next 33 bits */
#define DM_PIO2_3 (0x1.3198a2e037073p-69)  /* pi/2, the rest */
#define DM_TWO_OVER_PI (0x1.45f306dc9c883p-1)  /* 2/pi */
#define DM_LN2_HI (0x1.62e42fefa3800p-1)  /* ln 2, first 42 bits */
#define DM_LN2_LO (0x1.ef35793c76730p-45)  /* ln 2, the rest */
#define DM_INV_LN2 (0x1.71547652b82fep+0)  /* 1/ln 2 */
#define DM_SQRT_HALF (0x1.6a09e667f3bcdp-1)  /* sqrt(1/2) */
/* atan(k/8) (row 0), pi/2 - atan(k/8), pi - atan(k/8), pi/2 + atan(k/8), k = 0..8 */
DM_TABLE double dm_atan_hi[4][9] = {
    {0x0.0p+0, 0x1.fd5ba9aac2f6ep-4, 0x1.f5b75f92c80ddp-3, 0x1.6f61941e4def1p-2, 0x1.dac670561bb4fp-2, 0x1.1e00babdefeb4p-1, 0x1.4978fa3269ee1p-1, 0x1.700a7c5784634p-1, 0x1.921fb54442d18p-1},
    {0x1.921fb54442d18p+0, 0x1.7249faa996a21p+0, 0x1.5368c951e9cfdp+0, 0x1.3647503caf55cp+0, 0x1.1b6e192ebbe44p+0, 0x1.031f57e54adbep+0, 0x1.dac670561bb4fp-1, 0x1.b434ee31013fdp-1, 0x1.921fb54442d18p-1},
    {0x1.921fb54442d18p+1, 0x1.8234d7f6ecb9dp+1, 0x1.72c43f4b1650ap+1, 0x1.643382c07913ap+1, 0x1.56c6e7397f5aep+1, 0x1.4a9f8694c6d6bp+1, 0x1.3fc176b7a8560p+1, 0x1.361d162e61b8bp+1, 0x1.2d97c7f3321d2p+1},
    {0x1.921fb54442d18p+0, 0x1.b1f56fdeef00fp+0, 0x1.d0d6a1369bd34p+0, 0x1.edf81a4bd64d4p+0, 0x1.0468a8ace4df6p+1, 0x1.109009519d639p+1, 0x1.1b6e192ebbe44p+1, 0x1.251279b802819p+1, 0x1.2d97c7f3321d2p+1},
};
DM_TABLE double dm_atan_lo[4][9] = {
    {0x0.0p+0, -0x1.cd37686760c17p-59, 0x1.8ab6e3cf7afbdp-57, -0x1.c63aae6f6e918p-56, 0x1.a2b7f222f65e2p-56, -0x1.928df287a668fp-58, 0x1.2419a87f2a458p-56, -0x1.8c34d25aadef6p-56, 0x1.1a62633145c07p-55},
    {0x1.1a62633145c07p-54, 0x1.a8cc1e7480c68p-54, -0x1.96f47948a99f1p-54, 0x1.17e21d9a42c9ap-55, 0x1.b1b466a88828ep-54, 0x1.338b4259c0270p-54, 0x1.a2b7f222f65e2p-55, -0x1.0520d0701d877p-55, 0x1.1a62633145c07p-55},
    {0x1.1a62633145c07p-53, -0x1.3cd17e5a39792p-54, 0x1.c1b6f4f44e10bp-53, 0x1.a65371fe67254p-54, 0x1.660b64ece6f4bp-53, 0x1.26f6d2c582f3bp-53, -0x1.441a3bd3f1083p-58, 0x1.4be8fd7c9b7e6p-53, 0x1.a79394c9e8a0ap-54},
    {0x1.1a62633145c07p-54, 0x1.17f14fdc1574cp-55, -0x1.a23602a65700cp-57, 0x1.a8d3b7956a1c1p-54, 0x1.0620bf7406affp-55, 0x1.01398408cb59ep-54, 0x1.b1b466a88828ep-53, 0x1.6eaa5d3534893p-55, 0x1.a79394c9e8a0ap-54},
};
/* Taylor coefficients: sin (-1)^k/(2k+1)!, cos (-1)^k/(2k)!, exp 1/k!, atan (-1)^k/(2k+1), atanh 1/(2k+1) */
#define DM_SIN0 (-0x1.5555555555555p-3)
#define DM_SIN1 (0x1.1111111111111p-7)
#define DM_SIN2 (-0x1.a01a01a01a01ap-13)
#define DM_SIN3 (0x1.71de3a556c734p-19)
#define DM_SIN4 (-0x1.ae64567f544e4p-26)
#define DM_SIN5 (0x1.6124613a86d09p-33)
#define DM_SIN6 (-0x1.ae7f3e733b81fp-41)
#define DM_SIN7 (0x1.952c77030ad4ap-49)
#define DM_COS0 (0x1.5555555555555p-5)
#define DM_COS1 (-0x1.6c16c16c16c17p-10)
#define DM_COS2 (0x1.a01a01a01a01ap-16)
#define DM_COS3 (-0x1.27e4fb7789f5cp-22)
#define DM_COS4 (0x1.1eed8eff8d898p-29)
#define DM_COS5 (-0x1.93974a8c07c9dp-37)
#define DM_COS6 (0x1.ae7f3e733b81fp-45)
#define DM_COS7 (-0x1.6827863b97d97p-53)
#define DM_EXP0 (0x1.0000000000000p-1)
#define DM_EXP1 (0x1.5555555555555p-3)
#define DM_EXP2 (0x1.5555555555555p-5)
#define DM_EXP3 (0x1.1111111111111p-7)
#define DM_EXP4 (0x1.6c16c16c16c17p-10)
#define DM_EXP5 (0x1.a01a01a01a01ap-13)
#define DM_EXP6 (0x1.a01a01a01a01ap-16)
#define DM_EXP7 (0x1.71de3a556c734p-19)
#define DM_EXP8 (0x1.27e4fb7789f5cp-22)
#define DM_EXP9 (0x1.ae64567f544e4p-26)
#define DM_EXP10 (0x1.1eed8eff8d898p-29)
#define DM_EXP11 (0x1.6124613a86d09p-33)
#define DM_EXP12 (0x1.93974a8c07c9dp-37)
#define DM_ATAN0 (-0x1.5555555555555p-2)
#define DM_ATAN1 (0x1.999999999999ap-3)
#define DM_ATAN2 (-0x1.2492492492492p-3)
#define DM_ATAN3 (0x1.c71c71c71c71cp-4)
#define DM_ATAN4 (-0x1.745d1745d1746p-4)
#define DM_ATAN5 (0x1.3b13b13b13b14p-4)
#define DM_ATAN6 (-0x1.1111111111111p-4)
#define DM_ATAN7 (0x1.e1e1e1e1e1e1ep-5)
#define DM_ATANH0 (0x1.5555555555555p-2)
#define DM_ATANH1 (0x1.999999999999ap-3)
#define DM_ATANH2 (0x1.2492492492492p-3)
#define DM_ATANH3 (0x1.c71c71c71c71cp-4)
#define DM_ATANH4 (0x1.745d1745d1746p-4)
#define DM_ATANH5 (0x1.3b13b13b13b14p-4)
#define DM_ATANH6 (0x1.1111111111111p-4)
#define DM_ATANH7 (0x1.e1e1e1e1e1e1ep-5)
#define DM_ATANH8 (0x1.af286bca1af28p-5)
#define DM_ATANH9 (0x1.8618618618618p-5)
#define DM_ATANH10 (0x1.642c8590b2164p-5)
/* END generated constants */

/* a + b = *s + *e exactly (Knuth's two-sum) */
DM_FN void dm_two_sum(double a, double b, double* s, double* e)
{
    const double t = a + b;
    const double bb = t - a;
    *e = (a - (t - bb)) + (b - bb);
    *s = t;
}

/* a + b = *s + *e exactly when |a| >= |b| or a == 0 (Dekker's fast two-sum) */
DM_FN void dm_fast_two_sum(double a, double b, double* s, double* e)
{
    const double t = a + b;
    *e = b - (t - a);
    *s = t;
}

DM_FN void dm_sincos(double x, double* s, double* c)
{
    if (x == 0.0) {  /* keeps the sign of a zero */
        *s = x;
        *c = 1.0;
        return;
    }
    const double k = rint(x * DM_TWO_OVER_PI);
    /* k * DM_PIO2_1 and k * DM_PIO2_2 are exact (33-bit constants, |k| < 32)
     * and x - k * DM_PIO2_1 is exact (the two are within a factor 2) */
    const double r1 = x - k * DM_PIO2_1;
    double h, l, rh, rl;
    dm_two_sum(r1, -(k * DM_PIO2_2), &h, &l);
    l = l - k * DM_PIO2_3;
    dm_fast_two_sum(h, l, &rh, &rl);
    const double z = rh * rh;
    double ps = fma(z, DM_SIN7, DM_SIN6);
    ps = fma(z, ps, DM_SIN5);
    ps = fma(z, ps, DM_SIN4);
    ps = fma(z, ps, DM_SIN3);
    ps = fma(z, ps, DM_SIN2);
    ps = fma(z, ps, DM_SIN1);
    ps = fma(z, ps, DM_SIN0);
    double pc = fma(z, DM_COS7, DM_COS6);
    pc = fma(z, pc, DM_COS5);
    pc = fma(z, pc, DM_COS4);
    pc = fma(z, pc, DM_COS3);
    pc = fma(z, pc, DM_COS2);
    pc = fma(z, pc, DM_COS1);
    pc = fma(z, pc, DM_COS0);
    /* sin(rh + rl) = sin rh + rl cos rh, cos(rh + rl) = cos rh - rl sin rh */
    const double sr = rh + (rh * (z * ps) + rl * (1.0 - 0.5 * z));
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;  /* 1 - w is exact, and (1 - w) - hz is the rounding error of w */
    const double cr = w + (((1.0 - w) - hz) + ((z * z) * pc - rh * rl));
    const int n = (int)k & 3;
    if (n == 0) {
        *s = sr;
        *c = cr;
    } else if (n == 1) {
        *s = cr;
        *c = -sr;
    } else if (n == 2) {
        *s = -sr;
        *c = -cr;
    } else {
        *s = -cr;
        *c = sr;
    }
}

DM_FN double dm_atan2(double y, double x)
{
    const double ax = fabs(x), ay = fabs(y);
    if (ax != ax || ay != ay) return x + y;
    if (ay == 0.0) return copysign(signbit(x) ? dm_atan_hi[2][0] : 0.0, y);
    const int swap = ay > ax;
    const double num = swap ? ax : ay, den = swap ? ay : ax;
    const int k = (int)rint(8.0 * (num / den));
    const double ck = 0.125 * (double)k;
    const double sv = fma(-ck, den, num) / fma(ck, num, den);  /* (t - c)/(1 + t c), t = num/den */
    const double z = sv * sv;
    double pa = fma(z, DM_ATAN7, DM_ATAN6);
    pa = fma(z, pa, DM_ATAN5);
    pa = fma(z, pa, DM_ATAN4);
    pa = fma(z, pa, DM_ATAN3);
    pa = fma(z, pa, DM_ATAN2);
    pa = fma(z, pa, DM_ATAN1);
    pa = fma(z, pa, DM_ATAN0);
    const double p = fma(sv, z * pa, sv);
    /* rows: atan t, pi/2 - atan t (|y| > |x|), pi - atan t (x < 0), pi/2 + atan t (both) */
    const int row = swap + 2 * (x < 0.0);
    const double r = (row == 0 || row == 3) ? dm_atan_hi[row][k] + (dm_atan_lo[row][k] + p)
                                            : dm_atan_hi[row][k] + (dm_atan_lo[row][k] - p);
    return copysign(r, y);
}

DM_FN double dm_acos(double x)
{
    return dm_atan2(sqrt((1.0 - x) * (1.0 + x)), x);
}

/* log x = *hi + *lo for finite x > 0 */
DM_FN void dm_log_dd(double x, double* hi, double* lo)
{
    int e;
    double f = frexp(x, &e);
    if (f < DM_SQRT_HALF) {
        f = 2.0 * f;
        e = e - 1;
    }
    const double a = f - 1.0;  /* exact */
    double b, bl;
    dm_two_sum(f, 1.0, &b, &bl);
    const double rb = 1.0 / b;
    const double q = a * rb;
    const double ql = (fma(-q, b, a) - q * bl) * rb;  /* a / (b + bl) = q + ql */
    const double z = q * q;
    double pt = fma(z, DM_ATANH10, DM_ATANH9);
    pt = fma(z, pt, DM_ATANH8);
    pt = fma(z, pt, DM_ATANH7);
    pt = fma(z, pt, DM_ATANH6);
    pt = fma(z, pt, DM_ATANH5);
    pt = fma(z, pt, DM_ATANH4);
    pt = fma(z, pt, DM_ATANH3);
    pt = fma(z, pt, DM_ATANH2);
    pt = fma(z, pt, DM_ATANH1);
    pt = fma(z, pt, DM_ATANH0);
    const double ed = (double)e;
    double s, sl;
    dm_two_sum(ed * DM_LN2_HI, 2.0 * q, &s, &sl);  /* ed * DM_LN2_HI is exact (42-bit constant) */
    sl = sl + (ed * DM_LN2_LO + 2.0 * (ql + q * (z * pt)));
    dm_fast_two_sum(s, sl, hi, lo);
}

DM_FN double dm_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    double h, l;
    dm_log_dd(x, &h, &l);
    return h + l;
}

/* exp(h + l) for |l| <= 2^-40 |h| or so, h not NaN */
DM_FN double dm_exp_dd(double h, double l)
{
    if (h > 710.0) return __builtin_inf();
    if (h < -746.0) return 0.0;
    const double k = rint(h * DM_INV_LN2);
    const double rh = h - k * DM_LN2_HI;  /* exact: k * DM_LN2_HI is exact and within a factor 2 of h */
    const double rl = l - k * DM_LN2_LO;
    double pe = fma(rh, DM_EXP12, DM_EXP11);
    pe = fma(rh, pe, DM_EXP10);
    pe = fma(rh, pe, DM_EXP9);
    pe = fma(rh, pe, DM_EXP8);
    pe = fma(rh, pe, DM_EXP7);
    pe = fma(rh, pe, DM_EXP6);
    pe = fma(rh, pe, DM_EXP5);
    pe = fma(rh, pe, DM_EXP4);
    pe = fma(rh, pe, DM_EXP3);
    pe = fma(rh, pe, DM_EXP2);
    pe = fma(rh, pe, DM_EXP1);
    pe = fma(rh, pe, DM_EXP0);
    const double u = (rh * rh) * pe;             /* exp rh - 1 - rh */
    const double corr = rl + rl * (rh + u);      /* rl exp rh - rl */
    return ldexp(1.0 + (rh + (u + corr)), (int)k);
}

DM_FN double dm_exp(double x)
{
    if (x != x) return x;
    return dm_exp_dd(x, 0.0);
}

DM_FN double dm_sin(double x)
{
    double s, c;
    dm_sincos(x, &s, &c);
    return s;
}

DM_FN double dm_cos(double x)
{
    double s, c;
    dm_sincos(x, &s, &c);
    return c;
}

/* x^y for x >= 0 (and NaN); NaN for x < 0, where the C library returns a
 * value only for integral y (the step never passes a negative base) */
DM_FN double dm_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x != x || y != y) return x + y;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return y > 0.0 ? 0.0 : __builtin_inf();
    if (x == __builtin_inf()) return y > 0.0 ? x : 0.0;
    double lh, ll;
    dm_log_dd(x, &lh, &ll);
    const double ph = y * lh;
    const double pl = fma(y, lh, -ph) + y * ll;
    double h, l;
    dm_fast_two_sum(ph, pl, &h, &l);
    return dm_exp_dd(h, l);
}

#endif
