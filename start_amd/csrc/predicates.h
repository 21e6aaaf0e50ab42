// Exact orientation predicates for host and device code.
//
// orient3d(a,b,c,d) = sign of det[b-a; c-a; d-a].  It reproduces the exact signs
// of CGAL's filtered Epick predicates used by the reference for containment
// (VolumetricMesh.cpp:23-33) and segment/triangle tests
// (SegmentTriangleIntersection.cpp:7-10).
//
// Evaluation: the double-precision determinant with Shewchuk's a-priori error
// bound (J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
// Robust Geometric Predicates", DCG 18, 1997: o3derrboundA = (7+56e)e); if the
// bound cannot certify the sign, the determinant is evaluated exactly as a
// floating-point expansion (TwoDiff / TwoProduct / Grow-Expansion with zero
// elimination, same paper).  TwoProduct uses fma, which is exact.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#define SG_HD __host__ __device__

namespace sg {

constexpr double kEps = 1.1102230246251565e-16;  // 2^-53
constexpr int kExpMax = 96;                      // expansion capacity

SG_HD inline void two_sum(double a, double b, double& x, double& y)
{
    x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    y = ar + br;
}

SG_HD inline void two_diff(double a, double b, double& x, double& y)
{
    x = a - b;
    double bv = a - x;
    double av = x + bv;
    double br = bv - b;
    double ar = a - av;
    y = ar + br;
}

SG_HD inline void two_prod(double a, double b, double& x, double& y)
{
    x = a * b;
    y = fma(a, b, -x);
}

// Adds b to the nonoverlapping expansion h[0..n) (increasing magnitude) in
// place; returns the new length, or -1 when the capacity would be exceeded.
SG_HD inline int grow_expansion(double* h, int n, double b)
{
    double q = b;
    int k = 0;
    for (int i = 0; i < n; ++i) {
        double s, e;
        two_sum(q, h[i], s, e);
        q = s;
        if (e != 0.0) h[k++] = e;
    }
    if (q != 0.0 || k == 0) {
        if (k >= kExpMax) return -1;
        h[k++] = q;
    }
    return k;
}

SG_HD inline double orient3d_fast(const double* a, const double* b, const double* c, const double* d,
                                  double& permanent)
{
    double bax = b[0] - a[0], bay = b[1] - a[1], baz = b[2] - a[2];
    double cax = c[0] - a[0], cay = c[1] - a[1], caz = c[2] - a[2];
    double dax = d[0] - a[0], day = d[1] - a[1], daz = d[2] - a[2];
    double m1a = cay * daz, m1b = caz * day;
    double m2a = caz * dax, m2b = cax * daz;
    double m3a = cax * day, m3b = cay * dax;
    double det = bax * (m1a - m1b) + bay * (m2a - m2b) + baz * (m3a - m3b);
    permanent = fabs(bax) * (fabs(m1a) + fabs(m1b)) + fabs(bay) * (fabs(m2a) + fabs(m2b)) +
                fabs(baz) * (fabs(m3a) + fabs(m3b));
    return det;
}

// Exact sign; returns 2 if the expansion capacity was exceeded (never for
// inputs whose exponents span less than ~1500 bits).
#ifdef __HIP_DEVICE_COMPILE__
__attribute__((noinline))
#endif
SG_HD inline int orient3d_exact(const double* a, const double* b, const double* c, const double* d)
{
    double u[3][2], v[3][2], w[3][2];
    for (int k = 0; k < 3; ++k) {
        two_diff(b[k], a[k], u[k][0], u[k][1]);
        two_diff(c[k], a[k], v[k][0], v[k][1]);
        two_diff(d[k], a[k], w[k][0], w[k][1]);
    }
    // det = sum over permutations (i,j,k) of sgn * u_i v_j w_k
    const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}, {1, 0, 2}, {2, 0, 1}, {2, 1, 0}};
    const double S[6] = {1.0, -1.0, 1.0, -1.0, 1.0, -1.0};
    double h[kExpMax];
    int n = 0;
    for (int m = 0; m < 6; ++m) {
        for (int sel = 0; sel < 8; ++sel) {
            double f1 = u[P[m][0]][sel & 1];
            double f2 = v[P[m][1]][(sel >> 1) & 1];
            double f3 = w[P[m][2]][(sel >> 2) & 1];
            if (f1 == 0.0 || f2 == 0.0 || f3 == 0.0) continue;
            double p, pe, q0, q1, q2, q3;
            two_prod(f1, f2, p, pe);
            two_prod(p, f3, q0, q1);
            two_prod(pe, f3, q2, q3);
            double t[4] = {q0, q1, q2, q3};
            for (int z = 0; z < 4; ++z) {
                if (t[z] == 0.0) continue;
                n = grow_expansion(h, n, S[m] * t[z]);
                if (n < 0) return 2;
            }
        }
    }
    if (n == 0) return 0;
    double top = h[n - 1];
    return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

// Filtered orientation: 1 / -1 certain, 0 means "ask the exact version".
SG_HD inline int orient3d_filter(const double* a, const double* b, const double* c, const double* d)
{
    double perm;
    double det = orient3d_fast(a, b, c, d, perm);
    double bound = (7.0 + 56.0 * kEps) * kEps * perm;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return 0;
}

SG_HD inline int orient3d(const double* a, const double* b, const double* c, const double* d)
{
    int s = orient3d_filter(a, b, c, d);
    if (s != 0) return s;
    return orient3d_exact(a, b, c, d);
}

// orient2d exact sign (2x2 determinant of differences), for coplanar cases.
SG_HD inline int orient2d(const double* a, const double* b, const double* c)
{
    double l = (b[0] - a[0]) * (c[1] - a[1]);
    double r = (b[1] - a[1]) * (c[0] - a[0]);
    double det = l - r;
    double bound = (3.0 + 16.0 * kEps) * kEps * (fabs(l) + fabs(r));
    if (det > bound) return 1;
    if (-det > bound) return -1;
    double bx[2], by[2], cx[2], cy[2];
    two_diff(b[0], a[0], bx[0], bx[1]);
    two_diff(b[1], a[1], by[0], by[1]);
    two_diff(c[0], a[0], cx[0], cx[1]);
    two_diff(c[1], a[1], cy[0], cy[1]);
    double h[kExpMax];
    int n = 0;
    for (int sel = 0; sel < 4; ++sel) {
        double f[2][2] = {{bx[sel & 1], cy[(sel >> 1) & 1]}, {by[sel & 1], cx[(sel >> 1) & 1]}};
        for (int m = 0; m < 2; ++m) {
            if (f[m][0] == 0.0 || f[m][1] == 0.0) continue;
            double p, e;
            two_prod(f[m][0], f[m][1], p, e);
            double sg = m ? -1.0 : 1.0;
            n = grow_expansion(h, n, sg * p);
            if (n < 0) return 2;
            if (e != 0.0) {
                n = grow_expansion(h, n, sg * e);
                if (n < 0) return 2;
            }
        }
    }
    if (n == 0) return 0;
    double top = h[n - 1];
    return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

}  // namespace sg
