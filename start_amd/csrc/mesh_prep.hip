// Host-side mesh preparation (see mesh_prep.h).
#include "mesh_prep.h"

#include <math.h>

#include <cmath>

#include <algorithm>
#include <array>
#include <cstring>
#include <limits>

namespace sg {

namespace {

const double* node_ptr(const HostMesh& m, int32_t i) { return &m.nodes[3 * (size_t)i]; }

bool tri_is_degenerate(const double* a, const double* b, const double* c)
{
    // Triangle_3::is_degenerate: collinear vertices (exact)
    // collinear iff orient2d is zero in all three coordinate projections
    const int P[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    for (int k = 0; k < 3; ++k) {
        double A[2] = {a[P[k][0]], a[P[k][1]]}, B[2] = {b[P[k][0]], b[P[k][1]]}, C[2] = {c[P[k][0]], c[P[k][1]]};
        if (orient2d(A, B, C) != 0) return false;
    }
    return true;
}

void build_grid(const HostMesh& m, int64_t nitems, int nv, const std::vector<int32_t>& conn,
                std::vector<int32_t>& start, std::vector<int32_t>& list)
{
    const GridInfo& g = m.grid;
    int64_t ncell = (int64_t)g.div[0] * g.div[1] * g.div[2];
    start.assign((size_t)ncell + 1, 0);
    std::vector<std::array<int, 6>> range((size_t)nitems);
    for (int64_t t = 0; t < nitems; ++t) {
        auto& r = range[(size_t)t];
        for (int k = 0; k < 3; ++k) {
            double mn = std::numeric_limits<double>::infinity(), mx = -mn;
            for (int v = 0; v < nv; ++v) {
                double c = m.nodes[3 * (size_t)conn[(size_t)(nv * t + v)] + k];
                mn = std::min(mn, c);
                mx = std::max(mx, c);
            }
            // identical expression to the query: monotone rounding keeps every
            // point of the closed bbox inside this cell range
            r[k] = grid_coord(g, k, mn);
            r[3 + k] = grid_coord(g, k, mx);
        }
        for (int a = r[0]; a <= r[3]; ++a)
            for (int b = r[1]; b <= r[4]; ++b)
                for (int c = r[2]; c <= r[5]; ++c) start[(size_t)(((int64_t)a * g.div[1] + b) * g.div[2] + c) + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) start[(size_t)c + 1] += start[(size_t)c];
    list.assign((size_t)start[(size_t)ncell] + 1, 0);
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int64_t t = 0; t < nitems; ++t) {  // ascending id within each cell
        auto& r = range[(size_t)t];
        for (int a = r[0]; a <= r[3]; ++a)
            for (int b = r[1]; b <= r[4]; ++b)
                for (int c = r[2]; c <= r[5]; ++c)
                    list[(size_t)fill[(size_t)(((int64_t)a * g.div[1] + b) * g.div[2] + c)]++] = (int32_t)t;
    }
}

// Error-free transforms (exact on IEEE doubles without contraction):
// a - b = s + e and a * b = p + e.
inline void two_diff(double a, double b, double& s, double& e)
{
    const double nb = -b;  // Knuth's TwoSum of a and -b
    s = a + nb;
    const double bv = s - a;
    const double av = s - bv;
    e = (a - av) + (nb - bv);
}
inline void two_prod(double a, double b, double& p, double& e)
{
    p = a * b;
    e = std::fma(a, b, -p);
}

// R = A E - I with E = [v1 - v0, v2 - v0, v3 - v0] exactly (each difference a
// double-double hi + lo), and a bound err >= |R_computed - R_exact| per entry.
// A*hi is split exactly; A*lo (|lo| <= eps |hi|) is rounded once; the nine
// terms of an entry are summed in long double (64-bit significand).
void map_residual(const double* A, const double hi[3][3], const double lo[3][3], long double R[3][3],
                  double err[3][3])
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            long double acc = r == c ? -1.0L : 0.0L;
            double mag = 0.0;
            for (int k = 0; k < 3; ++k) {
                double p, pe;
                two_prod(A[3 * r + k], hi[k][c], p, pe);
                acc += (long double)p;
                acc += (long double)pe;
                acc += (long double)(A[3 * r + k] * lo[k][c]);
                mag += fabs(A[3 * r + k]) * (fabs(hi[k][c]) + fabs(lo[k][c]));
            }
            // 10 long-double additions (2^-64 each, relative to the running
            // magnitude <= mag + 1), the A*lo rounding (2^-53 * 2^-53 mag) and
            // the rounding of mag itself: 2^-58 (mag + 1) covers all of it
            R[r][c] = acc;
            err[r][c] = 0x1p-58 * (mag + 1.0);
        }
}

// The barycentric map of a positively oriented tet with vertices P[0..3]:
// g = v0 (3) then A (3x3, row-major) with lambda_{1,2,3}(x) = A (x - v0).
// A starts from the long-double cofactor inverse of the edge matrix E, is
// refined by one Newton step A <- A - R A with the residual R = A E - I taken
// nearly exactly, and is rounded to double.  Returns the largest row sum of
// |A|, the largest vertex coordinate difference (diam) and rho, a rigorous
// bound on the largest row sum of |A E - I| for the stored double A.
void map_of(const double* const P[4], double* g, double& rowsum, double& diam, double& rho)
{
    double hi[3][3], lo[3][3];
    long double e[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            two_diff(P[c + 1][r], P[0][r], hi[r][c], lo[r][c]);
            e[r][c] = (long double)hi[r][c] + (long double)lo[r][c];
        }
    diam = 0.0;
    for (int a = 0; a < 4; ++a)
        for (int b = a + 1; b < 4; ++b)
            for (int k = 0; k < 3; ++k) diam = std::max(diam, fabs(P[a][k] - P[b][k]));
    long double det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                      e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                      e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
    long double inv[3][3];
    inv[0][0] = (e[1][1] * e[2][2] - e[1][2] * e[2][1]) / det;
    inv[0][1] = (e[0][2] * e[2][1] - e[0][1] * e[2][2]) / det;
    inv[0][2] = (e[0][1] * e[1][2] - e[0][2] * e[1][1]) / det;
    inv[1][0] = (e[1][2] * e[2][0] - e[1][0] * e[2][2]) / det;
    inv[1][1] = (e[0][0] * e[2][2] - e[0][2] * e[2][0]) / det;
    inv[1][2] = (e[0][2] * e[1][0] - e[0][0] * e[1][2]) / det;
    inv[2][0] = (e[1][0] * e[2][1] - e[1][1] * e[2][0]) / det;
    inv[2][1] = (e[0][1] * e[2][0] - e[0][0] * e[2][1]) / det;
    inv[2][2] = (e[0][0] * e[1][1] - e[0][1] * e[1][0]) / det;
    double A[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) A[3 * r + c] = (double)inv[r][c];
    // one Newton step: the cofactor inverse errs by ~2^-64 kappa relative, more
    // than the double rounding once kappa > 2^11; afterwards only that rounding
    long double R[3][3];
    double err[3][3];
    map_residual(A, hi, lo, R, err);
    long double A1[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            long double corr = 0.0L;
            for (int k = 0; k < 3; ++k) corr += R[r][k] * (long double)A[3 * k + c];
            A1[3 * r + c] = (long double)A[3 * r + c] - corr;
        }
    for (int q = 0; q < 9; ++q) A[q] = (double)A1[q];
    map_residual(A, hi, lo, R, err);
    rowsum = 0.0;
    rho = 0.0;
    for (int r = 0; r < 3; ++r) {
        double srow = 0.0, rrow = 0.0;
        for (int c = 0; c < 3; ++c) {
            srow += fabs(A[3 * r + c]);
            rrow += fabsl(R[r][c]) + err[r][c];
        }
        rowsum = std::max(rowsum, srow);
        rho = std::max(rho, rrow * (1.0 + 1e-12));
    }
    g[0] = P[0][0];
    g[1] = P[0][1];
    g[2] = P[0][2];
    for (int q = 0; q < 9; ++q) g[3 + q] = A[q];
}

}  // namespace

// The step kernels evaluate lambda at a point x (double) as
//   d = x - v0;  l_r = (A_r0 d0 + A_r1 d1) + A_r2 d2;  l_0 = 1 - ((l1 + l2) + l3)
// (no contraction).  Against the exact mu = E^-1 (x - v0), with
// |x - v0|_inf <= D and R = A E - I:
//   |l_r - mu_r| <= (eps + gamma_3) sum_k |A_rk| |x_k - v0_k| + |(R mu)_r|
//                <= 4.01 eps rowsum D + rho max|mu|,
//   |l_0 - mu_0| <= sum_r |l_r - mu_r| + gamma_3 (1 + sum |l|).
// The kernels classify only points with sum |l| <= 2, i.e. within
// (2 + O(tau)) diam of v0 and with max|mu| <= 2 + O(tau): with D = 2 diam the
// worst case is 3 (8.02 eps R' + 2 rho) + 9 eps, R' = rowsum diam.  tau =
// 4 eps (12 R' + 8) + 8 rho is 2x the eps part and 1.33x the rho part.  Both
// grow linearly in |x - v0|, which the multi-tet walk's scaled tolerance
// relies on.
double tau_bound(double rowsum, double diam, double rho)
{
    return 4.0 * kEps * (12.0 * rowsum * diam + 8.0) + 8.0 * rho;
}

std::string prepare_mesh(HostMesh& m, int64_t n_nodes, const double* nodes, int64_t n_tets,
                         const int32_t* tets, int64_t n_tris, const int32_t* tris, double bbox_expansion)
{
    if (n_nodes <= 0 || n_tets <= 0 || !nodes || !tets) return "mesh needs nodes and tets";
    if (n_tets > (int64_t)std::numeric_limits<int32_t>::max() / 2 ||
        n_tris > (int64_t)std::numeric_limits<int32_t>::max() / 2)
        return "mesh too large for 32-bit ids";
    m.n_nodes = n_nodes;
    m.n_tets = n_tets;
    m.n_tris = n_tris;
    m.nodes.assign(nodes, nodes + 3 * n_nodes);
    m.tets.assign(tets, tets + 4 * n_tets);
    m.tris.assign(tris ? tris : tets, tris ? tris + 3 * n_tris : tets);
    for (int64_t i = 0; i < 4 * n_tets; ++i)
        if (m.tets[(size_t)i] < 0 || m.tets[(size_t)i] >= n_nodes) return "tet node id out of range";
    for (int64_t i = 0; i < 3 * n_tris; ++i)
        if (m.tris[(size_t)i] < 0 || m.tris[(size_t)i] >= n_nodes) return "triangle node id out of range";

    // 1. orientation
    for (int64_t t = 0; t < n_tets; ++t) {
        int32_t* v = &m.tets[4 * (size_t)t];
        int s = orient3d(node_ptr(m, v[0]), node_ptr(m, v[1]), node_ptr(m, v[2]), node_ptr(m, v[3]));
        if (s == 2) return "exact predicate capacity exceeded";
        if (s == 0) return "degenerate (zero-volume) tetrahedron " + std::to_string(t);
        if (s < 0) std::swap(v[2], v[3]);
    }

    // 2. face adjacency by sorted node triples
    struct FaceKey {
        int32_t a, b, c;
        int64_t tf;
    };
    std::vector<FaceKey> fk((size_t)(4 * n_tets));
    for (int64_t t = 0; t < n_tets; ++t)
        for (int f = 0; f < 4; ++f) {
            int32_t k[3] = {m.tets[4 * (size_t)t + kFace[f][0]], m.tets[4 * (size_t)t + kFace[f][1]],
                            m.tets[4 * (size_t)t + kFace[f][2]]};
            std::sort(k, k + 3);
            fk[(size_t)(4 * t + f)] = {k[0], k[1], k[2], 4 * t + f};
        }
    auto less = [](const FaceKey& x, const FaceKey& y) {
        if (x.a != y.a) return x.a < y.a;
        if (x.b != y.b) return x.b < y.b;
        if (x.c != y.c) return x.c < y.c;
        return x.tf < y.tf;
    };
    std::sort(fk.begin(), fk.end(), less);
    m.nbr.assign((size_t)(4 * n_tets), -1);
    for (size_t i = 0; i < fk.size();) {
        size_t j = i + 1;
        while (j < fk.size() && fk[j].a == fk[i].a && fk[j].b == fk[i].b && fk[j].c == fk[i].c) ++j;
        if (j - i > 2) return "non-manifold face shared by more than two tets";
        if (j - i == 2) {
            m.nbr[(size_t)fk[i].tf] = (int32_t)(fk[i + 1].tf / 4);
            m.nbr[(size_t)fk[i + 1].tf] = (int32_t)(fk[i].tf / 4);
        }
        i = j;
    }

    // 3. surface triangles onto faces; the first (smallest) id wins on duplicates
    m.tri_degenerate.assign((size_t)std::max<int64_t>(n_tris, 1), 0);
    std::vector<uint8_t> face_is_surface((size_t)(4 * n_tets), 0);
    for (int64_t s = 0; s < n_tris; ++s) {
        const int32_t* v = &m.tris[3 * (size_t)s];
        if (tri_is_degenerate(node_ptr(m, v[0]), node_ptr(m, v[1]), node_ptr(m, v[2]))) {
            m.tri_degenerate[(size_t)s] = 1;  // skipped like SurfaceMesh.cpp:16-21
            continue;
        }
        int32_t k[3] = {v[0], v[1], v[2]};
        std::sort(k, k + 3);
        FaceKey probe{k[0], k[1], k[2], -1};
        auto it = std::lower_bound(fk.begin(), fk.end(), probe, less);
        bool found = false;
        for (; it != fk.end() && it->a == k[0] && it->b == k[1] && it->c == k[2]; ++it) {
            found = true;
            if (!face_is_surface[(size_t)it->tf]) {
                face_is_surface[(size_t)it->tf] = 1;
                m.nbr[(size_t)it->tf] = -2 - (int32_t)s;
            }
        }
        if (!found) return "surface triangle " + std::to_string(s) + " is not a face of the tet mesh";
    }

    // 4. barycentric maps and the certification threshold tau (see map_of)
    m.geo.assign((size_t)(12 * n_tets), 0.0);
    m.rho.assign((size_t)n_tets, 0.0);
    double tau = 0.0;
    for (int64_t t = 0; t < n_tets; ++t) {
        const int32_t* v = &m.tets[4 * (size_t)t];
        const double* P[4] = {node_ptr(m, v[0]), node_ptr(m, v[1]), node_ptr(m, v[2]), node_ptr(m, v[3])};
        double rowsum = 0.0, diam = 0.0, rho = 0.0;
        map_of(P, &m.geo[12 * (size_t)t], rowsum, diam, rho);
        m.rho[(size_t)t] = rho;
        if (!(rho < 1e-3)) {  // the inverse is not even approximate: never certify with it
            tau = INFINITY;
            continue;
        }
        tau = std::max(tau, tau_bound(rowsum, diam, rho));
    }
    if (std::isnan(tau)) return "non-finite barycentric map (degenerate tetrahedron)";
    m.tau = tau;

    // 5. bucket grids
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n_nodes; ++i)
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], m.nodes[3 * (size_t)i + k]);
            mx[k] = std::max(mx[k], m.nodes[3 * (size_t)i + k]);
        }
    double vol = std::max((mx[0] - mn[0]) * (mx[1] - mn[1]) * (mx[2] - mn[2]), 1e-300);
    double edge = cbrt(vol / (double)n_tets * 6.0);
    for (int k = 0; k < 3; ++k) {
        m.node_lo[k] = mn[k];
        m.node_hi[k] = mx[k];
        m.grid.lo[k] = mn[k];
        int d = (int)ceil((mx[k] - mn[k]) / edge);
        m.grid.div[k] = std::max(1, std::min(d, 1024));
    }
    m.grid.edge = edge;
    build_grid(m, n_tets, 4, m.tets, m.tgrid_start, m.tgrid_list);
    build_grid(m, n_tris, 3, m.tris, m.sgrid_start, m.sgrid_list);

    // 6. surface bbox with the reference arithmetic (SurfaceMesh.cpp:92-134)
    double bmn[3] = {INFINITY, INFINITY, INFINITY}, bmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t s = 0; s < n_tris; ++s) {
        if (m.tri_degenerate[(size_t)s]) continue;
        for (int q = 0; q < 3; ++q) {
            const double* p = node_ptr(m, m.tris[3 * (size_t)s + q]);
            for (int k = 0; k < 3; ++k) {
                bmn[k] = std::min(bmn[k], p[k]);
                bmx[k] = std::max(bmx[k], p[k]);
            }
        }
    }
    for (int k = 0; k < 3; ++k) {
        if (n_tris == 0) {  // no surface: nothing is ever outside the modelling space
            m.bbox_lo[k] = -INFINITY;
            m.bbox_hi[k] = INFINITY;
            continue;
        }
        double centre = (bmn[k] + bmx[k]) / 2.0;
        double half = (bmx[k] - bmn[k]) / 2.0;
        half *= bbox_expansion;
        m.bbox_lo[k] = centre - half;
        m.bbox_hi[k] = centre + half;
    }
    return std::string();
}

}  // namespace sg
