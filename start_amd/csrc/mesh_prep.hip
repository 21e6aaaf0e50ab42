// Host-side mesh preparation (see mesh_prep.h).
#include "mesh_prep.h"

#include <math.h>

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

}  // namespace

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

    // 4. barycentric maps (long double inverse of the edge matrix) and the
    //    certification threshold tau
    m.geo.assign((size_t)(12 * n_tets), 0.0);
    double tau = 0.0;
    for (int64_t t = 0; t < n_tets; ++t) {
        const int32_t* v = &m.tets[4 * (size_t)t];
        const double* p0 = node_ptr(m, v[0]);
        long double e[3][3];  // columns: v1-v0, v2-v0, v3-v0
        double diam = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double* pc = node_ptr(m, v[c + 1]);
            for (int r = 0; r < 3; ++r) e[r][c] = (long double)pc[r] - (long double)p0[r];
        }
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                const double* pa = node_ptr(m, v[a]);
                const double* pb = node_ptr(m, v[b]);
                for (int k = 0; k < 3; ++k) diam = std::max(diam, fabs(pa[k] - pb[k]));
            }
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
        double* g = &m.geo[12 * (size_t)t];
        g[0] = p0[0];
        g[1] = p0[1];
        g[2] = p0[2];
        double rowsum = 0.0;
        for (int r = 0; r < 3; ++r) {
            double s = 0.0;
            for (int c = 0; c < 3; ++c) {
                g[3 + 3 * r + c] = (double)inv[r][c];
                s += fabs((double)inv[r][c]);
            }
            rowsum = std::max(rowsum, s);
        }
        // |computed - exact| of lambda at a point x with |x - v0|_inf <= D, to
        // first order in eps: map rounding (eps rowsum D) + (x - v0) rounding
        // (eps rowsum D) + 3-term dot product (3 eps rowsum D) = 5 eps rowsum D
        // for each of lambda_1..3; lambda_0 = 1 - ((l1 + l2) + l3) adds the
        // three and 3 eps (1 + sum |lambda_j|) <= 9 eps.  So <= 15 eps R + 9 eps
        // with R = rowsum D.  The step kernels classify only points with
        // sum |lambda| <= 2, hence |x - v0|_inf <= 2 diam (+ O(tau)): D = 2 diam,
        // and tau = 4 eps (6 R + 8) is 1.6x the bound.
        double bound = 4.0 * kEps * (12.0 * rowsum * diam + 8.0);
        tau = std::max(tau, bound);
    }
    if (tau > 1e-6) return "mesh too ill-conditioned for the certified barycentric fast path";
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
