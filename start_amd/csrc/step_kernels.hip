// Step kernels of the particle path (gfx950).
//
// One step = prologue (clear node charge, reset work list)
//          -> step_main   one thread per particle: skip rules, push, move,
//                         collision, certified same-tet relocation, charge
//          -> step_resolve one thread per deferred particle: exact tet walk
//                         along the step segment, surface hit and settle,
//                         grid search for degenerate cases, charge.
// The rare exact work (tet crossings, faces, hits) is compacted into a work
// list so the streaming kernel stays short and free of divergent branches.
#include <type_traits>

#include "kernels.h"
#include "predicates.h"

namespace sg {

// ---------------------------------------------------------------------------
// exact geometry on the device
// ---------------------------------------------------------------------------
__device__ inline void load_node(const DevMesh& M, int32_t n, double* p)
{
    p[0] = M.nodes[3 * (int64_t)n];
    p[1] = M.nodes[3 * (int64_t)n + 1];
    p[2] = M.nodes[3 * (int64_t)n + 2];
}

__device__ inline int nth(const int4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// closed containment: 0 outside, 1 on the boundary, 2 strictly inside, -1 error
__device__ int tet_contains(const DevMesh& M, int32_t t, const double* p)
{
    int4 nv = M.tnode[t];
    double P[4][3];
    for (int k = 0; k < 4; ++k) load_node(M, nth(nv, k), P[k]);
    int strict = 1;
    for (int f = 0; f < 4; ++f) {
        int s = orient3d(P[kFace[f][0]], P[kFace[f][1]], P[kFace[f][2]], p);
        if (s == 2) return -1;
        if (s < 0) return 0;
        if (s == 0) strict = 0;
    }
    return strict ? 2 : 1;
}

// minimum-id tet whose closed volume holds p (CubicGrid.cpp:94-104, :149-159)
__device__ int32_t locate_grid(const DevMesh& M, const double* p, int* strict, int* err)
{
    int a = grid_coord(M.grid, 0, p[0]), b = grid_coord(M.grid, 1, p[1]), c = grid_coord(M.grid, 2, p[2]);
    int64_t cell = ((int64_t)a * M.grid.div[1] + b) * M.grid.div[2] + c;
    for (int32_t j = M.tgrid_start[cell]; j < M.tgrid_start[cell + 1]; ++j) {
        int32_t t = M.tgrid_list[j];
        int r = tet_contains(M, t, p);
        if (r < 0) { *err = 1; continue; }
        if (r > 0) {
            *strict = (r == 2);
            return t;
        }
    }
    *strict = 0;
    return -1;
}

// closed segment/triangle intersection (SegmentTriangleIntersection.cpp:7-40);
// *t = parameter of the intersection point the reference ranks by distance
__device__ int seg_tri(const double* p, const double* q, const double* a, const double* b, const double* c,
                       double* t, int* err)
{
    int sp = orient3d(a, b, c, p), sq = orient3d(a, b, c, q);
    if (sp == 2 || sq == 2) { *err = 1; return 0; }
    if ((sp > 0 && sq > 0) || (sp < 0 && sq < 0)) return 0;
    if (sp == 0 && sq == 0) {
        double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
        double wx = c[0] - a[0], wy = c[1] - a[1], wz = c[2] - a[2];
        double nx = fabs(uy * wz - uz * wy), ny = fabs(uz * wx - ux * wz), nz = fabs(ux * wy - uy * wx);
        int i0, i1;
        if (nx >= ny && nx >= nz) { i0 = 1; i1 = 2; }
        else if (ny >= nz) { i0 = 0; i1 = 2; }
        else { i0 = 0; i1 = 1; }
        double P2[2] = {p[i0], p[i1]}, Q2[2] = {q[i0], q[i1]};
        double V[3][2] = {{a[i0], a[i1]}, {b[i0], b[i1]}, {c[i0], c[i1]}};
        auto in_tri = [&](const double* X) {
            int o1 = orient2d(V[0], V[1], X), o2 = orient2d(V[1], V[2], X), o3 = orient2d(V[2], V[0], X);
            if (o1 == 2 || o2 == 2 || o3 == 2) *err = 1;
            return (o1 >= 0 && o2 >= 0 && o3 >= 0) || (o1 <= 0 && o2 <= 0 && o3 <= 0);
        };
        auto on_seg = [](const double* s0, const double* s1, const double* r) {
            return fmin(s0[0], s1[0]) <= r[0] && r[0] <= fmax(s0[0], s1[0]) && fmin(s0[1], s1[1]) <= r[1] &&
                   r[1] <= fmax(s0[1], s1[1]);
        };
        auto seg_seg = [&](const double* p1, const double* p2, const double* q1, const double* q2) {
            int d1 = orient2d(q1, q2, p1), d2 = orient2d(q1, q2, p2);
            int d3 = orient2d(p1, p2, q1), d4 = orient2d(p1, p2, q2);
            if (((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0)) && ((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0))) return true;
            if (d1 == 0 && on_seg(q1, q2, p1)) return true;
            if (d2 == 0 && on_seg(q1, q2, p2)) return true;
            if (d3 == 0 && on_seg(p1, p2, q1)) return true;
            if (d4 == 0 && on_seg(p1, p2, q2)) return true;
            return false;
        };
        bool p_in = in_tri(P2);
        bool hit = p_in || in_tri(Q2) || seg_seg(P2, Q2, V[0], V[1]) || seg_seg(P2, Q2, V[1], V[2]) ||
                   seg_seg(P2, Q2, V[2], V[0]);
        if (!hit) return 0;
        if (p_in) {
            *t = 0.0;
        } else {
            double best = 1.0;
            for (int k = 0; k < 3; ++k) {
                const double* e0 = V[k];
                const double* e1 = V[(k + 1) % 3];
                double ex = e1[0] - e0[0], ey = e1[1] - e0[1];
                double dx = Q2[0] - P2[0], dy = Q2[1] - P2[1];
                double den = dx * ey - dy * ex;
                if (den == 0.0) continue;
                double tt = ((e0[0] - P2[0]) * ey - (e0[1] - P2[1]) * ex) / den;
                if (tt >= 0.0 && tt < best) best = tt;
            }
            *t = best;
        }
        return 1;
    }
    int o1 = orient3d(p, q, a, b), o2 = orient3d(p, q, b, c), o3 = orient3d(p, q, c, a);
    if (o1 == 2 || o2 == 2 || o3 == 2) { *err = 1; return 0; }
    bool hit = (o1 >= 0 && o2 >= 0 && o3 >= 0) || (o1 <= 0 && o2 <= 0 && o3 <= 0);
    if (!hit) return 0;
    if (sp == 0) *t = 0.0;
    else if (sq == 0) *t = 1.0;
    else {
        double pm;
        double fp = orient3d_fast(a, b, c, p, pm), fq = orient3d_fast(a, b, c, q, pm);
        *t = fp / (fp - fq);
    }
    return 1;
}

// closest surface triangle met by the closed segment p->q, ties to the
// smaller id (ParticleSurfaceCollisionHandler.cpp:104-123)
__device__ int32_t first_hit_grid(const DevMesh& M, const double* p, const double* q, int* err)
{
    int i0[3], i1[3];
    for (int k = 0; k < 3; ++k) {
        i0[k] = grid_coord(M.grid, k, fmin(p[k], q[k]));
        i1[k] = grid_coord(M.grid, k, fmax(p[k], q[k]));
    }
    int32_t best = -1;
    double best_t = INFINITY;
    for (int a = i0[0]; a <= i1[0]; ++a)
        for (int b = i0[1]; b <= i1[1]; ++b)
            for (int c = i0[2]; c <= i1[2]; ++c) {
                int64_t cell = ((int64_t)a * M.grid.div[1] + b) * M.grid.div[2] + c;
                for (int32_t j = M.sgrid_start[cell]; j < M.sgrid_start[cell + 1]; ++j) {
                    int32_t s = M.sgrid_list[j];
                    if (M.tri_degenerate[s]) continue;
                    double A[3], B[3], Cc[3];
                    load_node(M, M.tris[3 * (int64_t)s], A);
                    load_node(M, M.tris[3 * (int64_t)s + 1], B);
                    load_node(M, M.tris[3 * (int64_t)s + 2], Cc);
                    double t;
                    if (!seg_tri(p, q, A, B, Cc, &t, err)) continue;
                    if (t < best_t || (t == best_t && s < best)) {
                        best_t = t;
                        best = s;
                    }
                }
            }
    return best;
}

struct Resolved {
    int32_t hit;     // >= 0: settled on this triangle
    int32_t tet;     // owner of x1 (-1: outside the mesh)
    int32_t strict;  // x1 strictly inside tet
    int32_t exhaustive;
    int32_t err;
};

// Owner of p starting from a nearby tet: visibility walk (move across a face
// whose plane separates p from the tet) until p is strictly inside a tet,
// which is then the unique (so minimum-id) closed-containing tet.  A point on
// a face/edge/vertex, a walk that reaches the boundary or runs long goes to
// the grid search, which applies the minimum-id rule directly.
__device__ int32_t locate_near(const DevMesh& M, int32_t hint, const double* p, int* strict, int* err)
{
    int32_t t = hint, prev = -1;
    for (int it = 0; t >= 0 && it < 64; ++it) {
        int4 nv = M.tnode[t];
        double P[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k) load_node(M, nth(nv, k), P[k]);
        int s[4], nneg = 0, nzero = 0;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            s[f] = orient3d(P[kFace[f][0]], P[kFace[f][1]], P[kFace[f][2]], p);
            if (s[f] == 2) {
                *err = 1;
                s[f] = 0;
            }
            nneg += s[f] < 0;
            nzero += s[f] == 0;
        }
        if (nneg == 0) {
            if (nzero == 0) {
                *strict = 1;
                return t;
            }
            break;  // on the boundary of t: ties need the grid's minimum-id rule
        }
        // cross a separating face, preferring one that does not lead back
        int4 nb = M.tnbr[t];
        int32_t next = -1, back = -1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int f = (q + it) & 3;
            int32_t c = nth(nb, f);
            if (s[f] >= 0 || c < 0) continue;
            if (c != prev) {
                if (next < 0) next = c;
            } else {
                back = c;
            }
        }
        if (next < 0) next = back;
        if (next < 0) break;  // separated by a boundary face: outside or non-convex
        prev = t;
        t = next;
    }
    return locate_grid(M, p, strict, err);
}

__device__ Resolved resolve_exhaustive(const DevMesh& M, const double* x0, const double* x1, int32_t hint)
{
    Resolved r{-1, -1, 0, 1, 0};
    r.hit = first_hit_grid(M, x0, x1, &r.err);
    if (r.hit < 0) {
        int strict = 0;
        r.tet = hint >= 0 ? locate_near(M, hint, x1, &strict, &r.err) : locate_grid(M, x1, &strict, &r.err);
        r.strict = strict;
    }
    return r;
}

// Straight walk from t0 (x0 strictly inside) along x0->x1 through exit faces.
// Generic crossings are decided with exact signs; any zero sign (the segment
// meets an edge, a vertex or ends on a face) falls back to the grid search,
// which applies the reference rules directly.
__device__ __noinline__ Resolved resolve(const DevMesh& M, const double* x0, const double* x1, int32_t t0,
                                        bool strict0)
{
    if (t0 < 0 || !strict0) return resolve_exhaustive(M, x0, x1, t0);
    int32_t t = t0;
    int entry = -1;
    for (int iter = 0; iter < 256; ++iter) {
        int4 nv = M.tnode[t];
        double P[4][3];
        for (int k = 0; k < 4; ++k) load_node(M, nth(nv, k), P[k]);
        int s[4];
        int neg = 0, zero = 0;
        for (int f = 0; f < 4; ++f) {
            s[f] = orient3d(P[kFace[f][0]], P[kFace[f][1]], P[kFace[f][2]], x1);
            if (s[f] == 2) return resolve_exhaustive(M, x0, x1, t0);
            neg += s[f] < 0;
            zero += s[f] == 0;
        }
        if (neg == 0) {
            if (zero == 0) return Resolved{-1, t, 1, 0, 0};
            return resolve_exhaustive(M, x0, x1, t0);  // x1 on a face/edge/vertex of t
        }
        int exit_f = -1, nexit = 0;
        for (int f = 0; f < 4; ++f) {
            if (f == entry || s[f] >= 0) continue;
            const double* a = P[kFace[f][0]];
            const double* b = P[kFace[f][1]];
            const double* c = P[kFace[f][2]];
            int o1 = orient3d(x0, x1, a, b), o2 = orient3d(x0, x1, b, c), o3 = orient3d(x0, x1, c, a);
            if (o1 == 0 || o2 == 0 || o3 == 0 || o1 == 2 || o2 == 2 || o3 == 2)
                return resolve_exhaustive(M, x0, x1, t0);
            if (o1 == o2 && o2 == o3) {
                exit_f = f;
                ++nexit;
            }
        }
        if (nexit != 1) return resolve_exhaustive(M, x0, x1, t0);
        int4 nb = M.tnbr[t];
        int32_t code = nth(nb, exit_f);
        if (code <= -2) return Resolved{-2 - code, -1, 0, 0, 0};  // surface triangle crossed first
        if (code == -1) return Resolved{-1, -1, 0, 0, 0};         // left through an open face
        int4 back = M.tnbr[code];
        entry = -1;
        for (int g = 0; g < 4; ++g)
            if (nth(back, g) == t) entry = g;
        t = code;
    }
    return resolve_exhaustive(M, x0, x1, t0);
}

// ---------------------------------------------------------------------------
// barycentric map and charge
// ---------------------------------------------------------------------------
__device__ inline void bary64(const double* __restrict__ g, double x, double y, double z, double lam[4])
{
    double dx = x - g[0], dy = y - g[1], dz = z - g[2];
    lam[1] = (g[3] * dx + g[4] * dy) + g[5] * dz;
    lam[2] = (g[6] * dx + g[7] * dy) + g[8] * dz;
    lam[3] = (g[9] * dx + g[10] * dy) + g[11] * dz;
    lam[0] = 1.0 - ((lam[1] + lam[2]) + lam[3]);
}

// FP64 atomic adds with the address space stated: a pointer that reaches the
// add through a struct is generic to the compiler, which then emits a FLAT
// atomic (slower, and for LDS it waits on both memory counters).
typedef __attribute__((address_space(3))) double lds_f64;
typedef __attribute__((address_space(1))) double global_f64;
__device__ inline void lds_add_f64(double* p, double v)
{
    __hip_atomic_fetch_add((lds_f64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);  // ds_add_f64
}
__device__ inline void global_add_f64(double* p, double v)
{
    __hip_atomic_fetch_add((global_f64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // global_atomic_add_f64
}

__device__ inline void scatter4(double* __restrict__ node_charge, const int4& nv, double q, const double lam[4])
{
    global_add_f64(&node_charge[nv.x], q * lam[0]);
    global_add_f64(&node_charge[nv.y], q * lam[1]);
    global_add_f64(&node_charge[nv.z], q * lam[2]);
    global_add_f64(&node_charge[nv.w], q * lam[3]);
}

// Block-local reduction of the charge scatter: contributions to the same node
// from the block's particles are summed in an LDS hash table (open addressing,
// linear probing) and flushed with one FP64 atomic per distinct node.  With
// particles sorted by tet a 256-particle block touches ~100-150 nodes instead
// of issuing 1,024 global atomics.
template <int Bits>
__device__ inline void lds_charge_add(int32_t* hk, double* hv, double* __restrict__ node_charge, int32_t node,
                                      double val)
{
    uint32_t h = ((uint32_t)node * 2654435761u) >> (32 - Bits);
    // (left to the compiler to unroll: a rolled loop made the FP64 streaming
    // kernel 7% slower)
    // (reading the slot before the compare-and-swap, to skip it for nodes
    // already present, made the kernel 17-35% slower)
    for (int probe = 0; probe < 32; ++probe) {
        int32_t prev = atomicCAS(&hk[h], -1, node);
        if (prev == -1 || prev == node) {
            lds_add_f64(&hv[h], val);
            return;
        }
        h = (h + 1) & ((1u << Bits) - 1u);
    }
    global_add_f64(&node_charge[node], val);  // crowded table: straight to memory
}

// Node window: a block's tet-sorted particles touch nodes within a few
// thousand ids of the lowest one (the node numbering of a structured mesh
// puts a cell's nodes one layer, (nx+1)(ny+1) ids, apart), so nodes in
// [wbase, wbase + Win) are summed in a plain LDS array by an LDS FP64 add,
// without the hash's compare-and-swap; the others go to the hash.
template <int Bits, int Win>
__device__ inline void lds_window_add(int32_t* hk, double* hv, double* wv, int32_t wbase,
                                      double* __restrict__ node_charge, int32_t node, double val)
{
    if (Win > 0) {
        const uint32_t off = (uint32_t)(node - wbase);
        if (off < (uint32_t)Win) {
            lds_add_f64(&wv[off], val);
            return;
        }
    }
    lds_charge_add<Bits>(hk, hv, node_charge, node, val);
}

// base: offset of the run's node-charge array (0 for a single simulation)
template <int Bits, int Win>
__device__ inline void lds_scatter4(int32_t* hk, double* hv, double* wv, int32_t wbase,
                                    double* __restrict__ node_charge, const int4& nv, int32_t base, double q,
                                    const double lam[4])
{
    lds_window_add<Bits, Win>(hk, hv, wv, wbase, node_charge, base + nv.x, q * lam[0]);
    lds_window_add<Bits, Win>(hk, hv, wv, wbase, node_charge, base + nv.y, q * lam[1]);
    lds_window_add<Bits, Win>(hk, hv, wv, wbase, node_charge, base + nv.z, q * lam[2]);
    lds_window_add<Bits, Win>(hk, hv, wv, wbase, node_charge, base + nv.w, q * lam[3]);
}

// ---------------------------------------------------------------------------
// per-thread tet cache and the two arithmetic layouts
//   FP64:  position absolute (double), map = v0 + A (double), E (double)
//   mixed: position relative to v0 of the owning tet (float), map A (float),
//          E (float); absolute coordinates are v0 (double) + local
// ---------------------------------------------------------------------------
// lazily loaded pieces of the owning tet's record (keeps registers free across
// the collision code)
__device__ inline void load_E(const DevMesh& M, int32_t t, double& ex, double& ey, double& ez)
{
    const double* E = M.E + 3 * (int64_t)t;
    ex = E[0];
    ey = E[1];
    ez = E[2];
}
__device__ inline void load_E(const DevMesh& M, int32_t t, float& ex, float& ey, float& ez)
{
    float4 e = reinterpret_cast<const float4*>(M.E32)[t];
    ex = e.x;
    ey = e.y;
    ez = e.z;
}
__device__ inline void load_E1(const DevMesh& M, int32_t t, double& ex, double& ey, double& ez)
{
    const double* E = M.E1 + 3 * (int64_t)t;
    ex = E[0];
    ey = E[1];
    ez = E[2];
}
__device__ inline void load_E1(const DevMesh& M, int32_t t, float& ex, float& ey, float& ez)
{
    float4 e = reinterpret_cast<const float4*>(M.E1_32)[t];
    ex = e.x;
    ey = e.y;
    ez = e.z;
}
// field of the particle's run in tet t: s0 * E_0 + s1 * E_1 (s0 = 1 and no
// second basis for a single simulation, where this is E_0 exactly)
template <typename R, bool kBatch>
__device__ inline void run_field(const DevMesh& M, const RunK* __restrict__ rk, int32_t t, R& ex, R& ey, R& ez)
{
    load_E(M, t, ex, ey, ez);
    if (!kBatch) return;
    const R s0 = (R)rk->s0;
    ex = s0 * ex;
    ey = s0 * ey;
    ez = s0 * ez;
    if (M.E1) {
        R fx, fy, fz;
        load_E1(M, t, fx, fy, fz);
        const R s1 = (R)rk->s1;
        ex = ex + s1 * fx;
        ey = ey + s1 * fy;
        ez = ez + s1 * fz;
    }
}
// push increment of tet t for the species in table `slot`
__device__ inline void load_push(const DevMesh& M, int slot, int32_t t, double& hx, double& hy, double& hz)
{
    const double* h = M.H + 3 * ((int64_t)slot * M.n_tets + t);
    hx = h[0];
    hy = h[1];
    hz = h[2];
}
__device__ inline void load_push(const DevMesh& M, int slot, int32_t t, float& hx, float& hy, float& hz)
{
    const float4 h = reinterpret_cast<const float4*>(M.H32)[(int64_t)slot * M.n_tets + t];
    hx = h.x;
    hy = h.y;
    hz = h.z;
}
__device__ inline void bary_at(const DevMesh& M, int32_t t, double x, double y, double z, double lam[4])
{
    bary64(M.geo + 12 * (int64_t)t, x, y, z, lam);
}
__device__ inline void bary_at(const DevMesh& M, int32_t t, float x, float y, float z, double lam[4])
{
    const float4* a = reinterpret_cast<const float4*>(M.geo32) + 3 * (int64_t)t;
    float4 a0 = a[0], a1 = a[1], a2 = a[2];
    float l1 = (a0.x * x + a0.y * y) + a0.z * z;
    float l2 = (a0.w * x + a1.x * y) + a1.y * z;
    float l3 = (a1.z * x + a1.w * y) + a2.x * z;
    float l0 = 1.0f - ((l1 + l2) + l3);
    lam[0] = l0;
    lam[1] = l1;
    lam[2] = l2;
    lam[3] = l3;
}
// bbox test of the frozen rule on the stored position
__device__ inline bool in_bbox_at(const DevMesh& M, const StepParams& S, int32_t t, double x, double y, double z)
{
    return x >= S.lo[0] && x <= S.hi[0] && y >= S.lo[1] && y <= S.hi[1] && z >= S.lo[2] && z <= S.hi[2];
}
__device__ inline bool in_bbox_at(const DevMesh& M, const StepParams& S, int32_t t, float x, float y, float z)
{
    float ax = x, ay = y, az = z;
    if (t >= 0) {
        float4 a2 = (reinterpret_cast<const float4*>(M.geo32) + 3 * (int64_t)t)[2];
        ax += a2.y;
        ay += a2.z;
        az += a2.w;
    }
    return ax >= (float)S.lo[0] && ax <= (float)S.hi[0] && ay >= (float)S.lo[1] && ay <= (float)S.hi[1] &&
           az >= (float)S.lo[2] && az <= (float)S.hi[2];
}

template <typename R>
__device__ inline double tau_of(const DevMesh& M);
template <>
__device__ inline double tau_of<double>(const DevMesh& M) { return M.tau; }
template <>
__device__ inline double tau_of<float>(const DevMesh& M) { return (double)M.tau32; }

__device__ inline void v0_of(const DevMesh& M, int32_t t, double* v0)
{
    const double* g = M.geo + 12 * (int64_t)t;
    v0[0] = g[0];
    v0[1] = g[1];
    v0[2] = g[2];
}

// absolute position of a stored state
__device__ inline void abs_pos(const DevMesh& M, int32_t t, double x, double y, double z, double* p)
{
    p[0] = x;
    p[1] = y;
    p[2] = z;
}
__device__ inline void abs_pos(const DevMesh& M, int32_t t, float x, float y, float z, double* p)
{
    double v0[3] = {0.0, 0.0, 0.0};
    if (t >= 0) v0_of(M, t, v0);
    p[0] = v0[0] + (double)x;
    p[1] = v0[1] + (double)y;
    p[2] = v0[2] + (double)z;
}

// store an absolute position in the frame of tet t
__device__ inline void to_frame(const DevMesh& M, int32_t t, const double* p, double& x, double& y, double& z)
{
    x = p[0];
    y = p[1];
    z = p[2];
}
__device__ inline void to_frame(const DevMesh& M, int32_t t, const double* p, float& x, float& y, float& z)
{
    double v0[3] = {0.0, 0.0, 0.0};
    if (t >= 0) v0_of(M, t, v0);
    x = (float)(p[0] - v0[0]);
    y = (float)(p[1] - v0[1]);
    z = (float)(p[2] - v0[2]);
}

// ---------------------------------------------------------------------------
// charge sinks: LDS hash (streaming kernel) or direct atomics (rare paths)
// ---------------------------------------------------------------------------
template <int Bits, int Win>
struct LdsSink {
    int32_t* hk;
    double* hv;
    double* wv;
    int32_t wbase;
    double* nc;
    __device__ void add4(const int4& nv, int32_t base, double q, const double lam[4])
    {
        lds_scatter4<Bits, Win>(hk, hv, wv, wbase, nc, nv, base, q, lam);
    }
};
struct GlobalSink {
    double* nc;
    __device__ void add4(const int4& nv, int32_t base, double q, const double lam[4])
    {
        scatter4(nc + base, nv, q, lam);
    }
};
// Per-particle accumulation across the steps of a launch: a particle that
// stays in one tet adds q*lambda into four registers, and its sums go to the
// inner sink only when it changes tet or the launch ends.  base is the offset
// of the particle's run in the node-charge buffer.
template <typename Inner>
struct RegSink {
    Inner inner;
    int32_t cur;
    int32_t base;
    double a0, a1, a2, a3;
    __device__ void flush(const DevMesh& M)
    {
        if (cur >= 0) {
            double v[4] = {a0, a1, a2, a3};
            inner.add4(M.tnode[cur], base, 1.0, v);
        }
        cur = -1;
        a0 = a1 = a2 = a3 = 0.0;
    }
    __device__ void add(const DevMesh& M, int32_t t, double q, const double lam[4])
    {
        if (t != cur) {
            flush(M);
            cur = t;
        }
        a0 += q * lam[0];
        a1 += q * lam[1];
        a2 += q * lam[2];
        a3 += q * lam[3];
    }
    // The flush at the end of a tile, where every lane of the wavefront is
    // active: lanes of one run of equal (tet, run offset) -- neighbours in the
    // tet-sorted order -- sum their four weights into the run's first lane
    // (segmented Hillis-Steele over the wavefront), which alone goes to the
    // inner sink.  One LDS insertion per run instead of one per particle.
    __device__ void flush_wave(const DevMesh& M)
    {
        const int lane = threadIdx.x & 63;
        const int32_t up_t = __shfl_up(cur, 1), up_b = __shfl_up(base, 1);
        const bool head = lane == 0 || up_t != cur || up_b != base;
        const unsigned long long hm = __ballot(head);
        const int seg = __popcll(hm & (~0ull >> (63 - lane)));  // runs starting at or below this lane
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int sd = __shfl_down(seg, d);
            const double b0 = __shfl_down(a0, d), b1 = __shfl_down(a1, d), b2 = __shfl_down(a2, d),
                         b3 = __shfl_down(a3, d);
            if (lane + d < 64 && sd == seg) {
                a0 += b0;
                a1 += b1;
                a2 += b2;
                a3 += b3;
            }
        }
        if (!head) cur = -1;  // summed into the run's first lane
        flush(M);
    }
    // flush_wave with the node window of an LDS sink, when no charge has gone
    // to LDS yet in this tile (single-step launches): the window starts at the
    // lowest node the tile's runs touch (a block minimum, s_wbase preset to
    // INT32_MAX); called by every thread of the workgroup.  The segmented sum
    // runs two rounds (shifts 1 and 2): lane i then holds the sum of lanes
    // i..i+3 of its run, and the lanes 0, 4, 8, ... after the start of their
    // run insert, so every lane is counted once.  With about one particle per
    // tet, runs longer than four are rare; the full six rounds cost more
    // cross-lane moves than the inserts they save (measured: 63.5 against
    // 65.8 us per single-step FP64 launch on the large box).
    __device__ void flush_wave_window(const DevMesh& M, int32_t* s_wbase)
    {
        const int lane = threadIdx.x & 63;
        const int32_t up_t = __shfl_up(cur, 1), up_b = __shfl_up(base, 1);
        const bool head = lane == 0 || up_t != cur || up_b != base;
        const unsigned long long hm = __ballot(head);
        const int seg = __popcll(hm & (~0ull >> (63 - lane)));
        const unsigned long long below = hm & (~0ull >> (63 - lane));
        const int start = 63 - __builtin_clzll(below);
#pragma unroll
        for (int d = 1; d < 4; d <<= 1) {
            const int sd = __shfl_down(seg, d);
            const double b0 = __shfl_down(a0, d), b1 = __shfl_down(a1, d), b2 = __shfl_down(a2, d),
                         b3 = __shfl_down(a3, d);
            if (lane + d < 64 && sd == seg) {
                a0 += b0;
                a1 += b1;
                a2 += b2;
                a3 += b3;
            }
        }
        if (((lane - start) & 3) != 0) cur = -1;
        int4 nv = make_int4(0, 0, 0, 0);
        int32_t lo = INT32_MAX;
        if (cur >= 0) {
            nv = M.tnode[cur];
            lo = base + min(min(nv.x, nv.y), min(nv.z, nv.w));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) lo = min(lo, __shfl_xor(lo, off));
        if (lane == 0) atomicMin(s_wbase, lo);
        __syncthreads();
        inner.wbase = __builtin_amdgcn_readfirstlane(*s_wbase);
        if (cur >= 0) {
            double v[4] = {a0, a1, a2, a3};
            inner.add4(nv, base, 1.0, v);
        }
        cur = -1;
        a0 = a1 = a2 = a3 = 0.0;
    }
};

// Collision decision u1 < P (HSModel.cpp:15-19, VHSModel.cpp:60-70).  For
// VHS/VSS the FP64 probability needs pow() every step; an FP32 estimate with a
// relative error far below 1e-4 decides all draws outside a +-1e-4 band
// around it, and only draws inside the band evaluate the FP64 expression, so
// the FP64 decision is reproduced exactly at FP32 cost.
__device__ __noinline__ bool collides_exact64(int model, double sigma_hs, double vhs_c, double two_kt, double m_r,
                                              double vhs_exp, double conc, double dt, double vm, double u1)
{
    CollideK<double> ck{sigma_hs, vhs_c, two_kt, m_r, vhs_exp, 0.0, 0.0, 0.0};
    return u1 < collision_probability<double>(model, ck, conc, dt, vm);
}

template <typename R>
__device__ inline bool collides(const StepParams& S, const SpeciesK& k, double conc, R vx, R vy, R vz, double u1)
{
    R vm = m_sqrt<R>(vx * vx + vy * vy + vz * vz);
    if (S.model == 1) {  // HS: no transcendental, evaluate directly
        CollideK<R> ck{(R)k.sigma_hs, (R)0, (R)0, (R)0, (R)0, (R)0, (R)0, (R)0};
        return (R)u1 < collision_probability<R>(1, ck, (R)conc, (R)S.dt, vm);
    }
    float vmf = (float)vm;
    CollideK<float> cf{(float)k.sigma_hs, (float)k.vhs_c, (float)k.two_kt, (float)k.m_r, (float)k.vhs_exp,
                       0.0f, 0.0f, 0.0f};
    float p32 = collision_probability<float>(S.model, cf, (float)conc, (float)S.dt, vmf);
    if (!std::is_same<R, double>::value) return (float)u1 < p32;
    double p = (double)p32;
    if (!(p > 1e-30 && p < 1e30))  // outside the range the FP32 bound covers
        return collides_exact64(S.model, k.sigma_hs, k.vhs_c, k.two_kt, k.m_r, k.vhs_exp, conc, S.dt, (double)vm, u1);
    if (u1 < p * (1.0 - 1e-4)) return true;
    if (u1 >= p * (1.0 + 1e-4)) return false;
    return collides_exact64(S.model, k.sigma_hs, k.vhs_c, k.two_kt, k.m_r, k.vhs_exp, conc, S.dt, (double)vm, u1);
}

// Post-collision velocity of a particle whose collision test passed (rare:
// about once per 1e4 particle-steps), out of line to keep the streaming kernel
// lean.  The variates u2, u3 are drawn here, not by the caller, so that few of
// the caller's registers are live at the call; arguments and result by value
// so nothing of the caller escapes to memory.  FP64 in both precisions:
// evaluated in FP32, sin(chi) = sqrt((1 - cos)(1 + cos)) of a forward scatter
// (cos near 1 from an FP64 variate rounded to FP32) loses most of its digits.
struct V3 {
    double x, y, z;
};
__device__ __noinline__ V3 collide_velocity(int model, const SpeciesK* k, uint64_t seed, uint32_t key, uint32_t vid,
                                            uint32_t step, double vx, double vy, double vz)
{
    double u2, u3;
    variates(seed, key, vid, step, 1u, u2, u3);
    scatter_velocity(model, k->mass_cp, k->mass_ct, k->inv_alpha, u2, u3, vx, vy, vz);
    return V3{vx, vy, vz};
}

// Barycentric classification with a certified threshold: strictly inside,
// strictly outside exactly one face (strictly inside the other three), or
// uncertain.  tau bounds the rounding of lambda at points within 2 diam of v0
// (diam: largest vertex coordinate difference).  Since x - v0 = sum_j
// lambda_j (v_j - v0), |x - v0|_inf <= diam * sum |lambda_j|, so the
// classification is only trusted when sum |lambda| <= 2: a long step that
// ends far outside the tet is left to the exact walk.
enum { kInside = 0, kOneFace = 1, kUncertain = 2 };
__device__ inline int classify(const double lam[4], double tau, int& fneg)
{
    int npos = 0, nneg = 0;
    fneg = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        npos += lam[q] > tau;
        if (lam[q] < -tau) {
            ++nneg;
            fneg = q;
        }
    }
    if (fabs(lam[0]) + fabs(lam[1]) + fabs(lam[2]) + fabs(lam[3]) > 2.0) return kUncertain;
    if (npos == 4) return kInside;
    if (npos == 3 && nneg == 1) return kOneFace;
    return kUncertain;
}

// Certified walk of the segment p0 -> p1 (absolute FP64, p0 strictly inside
// t) through as many tets as it crosses, with the FP64 barycentric maps, for
// steps longer than a tet.  In each tet the exit face is the face f with
// lambda_f(p1) < 0 whose plane the segment meets first: s_f = l0_f / (l0_f -
// l1_f) is smallest, i.e. D_fg = l0_f l1_g - l0_g l1_f > 0 for every other
// such g.  Signs and orderings are trusted only beyond their rounding bound:
// lambda at a point with sum|lambda| = S errs by at most tau * max(1, S/2)
// (tau covers S <= 2 and the bound is linear in |x - v0| <= diam * S).  The
// entry point lies inside a face (all other lambdas > 0), so the exit point
// found this way lies inside the exit face.  Anything closer than the bounds
// (edges, vertices, end points near a face) is left to the exact walk.
enum { kWalkInside = 0, kWalkSurface = 1, kWalkOpen = 2, kWalkUnsure = 3 };
constexpr int kWalkHops = 16;

__device__ inline double walk_tol(double tau, const double l[4])
{
    const double s = (fabs(l[0]) + fabs(l[1])) + (fabs(l[2]) + fabs(l[3]));
    return tau * fmax(1.0, 0.5 * s) * (1.0 + 1e-6);
}

__device__ inline int walk64(const DevMesh& M, int32_t t, const double* p0, const double* p1, int32_t& out)
{
    constexpr double kE = 1.1102230246251565e-16;  // 2^-53
    const double tau = M.tau;
    int32_t cur = t;
    for (int hop = 0; hop < kWalkHops; ++hop) {
        double l1[4];
        bary_at(M, cur, p1[0], p1[1], p1[2], l1);
        const double t1 = walk_tol(tau, l1);
        int nneg = 0, npos = 0, f1 = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            npos += l1[q] > t1;
            if (l1[q] < -t1) {
                ++nneg;
                f1 = q;
            }
        }
        if (npos + nneg < 4) return kWalkUnsure;  // p1 too close to a face plane of cur
        if (nneg == 0) {
            out = cur;
            return kWalkInside;
        }
        int exit_f = f1;
        if (nneg > 1) {
            double l0[4];
            bary_at(M, cur, p0[0], p0[1], p0[2], l0);
            const double t0 = walk_tol(tau, l0);
            exit_f = -1;
            for (int f = 0; f < 4 && exit_f < 0; ++f) {
                if (!(l1[f] < -t1)) continue;
                bool first = true;
                for (int g = 0; g < 4; ++g) {
                    if (g == f || !(l1[g] < -t1)) continue;
                    const double a = l0[f] * l1[g], b = l0[g] * l1[f];
                    const double err = t0 * (fabs(l1[f]) + fabs(l1[g]) + 2.0 * t1) +
                                       t1 * (fabs(l0[f]) + fabs(l0[g]) + 2.0 * t0) + 4.0 * kE * (fabs(a) + fabs(b));
                    if (!(a - b > err)) {
                        first = false;
                        break;
                    }
                }
                if (first) exit_f = f;
            }
            if (exit_f < 0) return kWalkUnsure;
        }
        const int32_t code = nth(M.tnbr[cur], exit_f);
        if (code <= -2) {
            out = -2 - code;
            return kWalkSurface;
        }
        if (code == -1) return kWalkOpen;
        cur = code;
    }
    return kWalkUnsure;
}

enum StepOutcome { kDone = 0, kSettledNow = 1, kDeferNow = 2, kFrozenNow = 3 };

// settle record for the adjoint (rare; out of line, scalar arguments)
__device__ __noinline__ void record_settle(SettleRec* rec, uint32_t pidx, uint32_t step, double x0, double y0,
                                           double z0)
{
    SettleRec r;
    r.x0 = x0;
    r.y0 = y0;
    r.z0 = z0;
    r.step = (int32_t)step;
    r.pad = 0;
    rec[pidx] = r;
}

// Where the segment xa -> xa + D meets the plane of triangle `tri`:
// xa + s D with s = n.(A - xa) / n.D, n = (B - A) x (C - A).
__device__ inline void hit_plane(const DevMesh& M, int32_t tri, const double* xa, const double* D, double* n,
                                 double& s)
{
    double A[3], B[3], C[3];
    load_node(M, M.tris[3 * (int64_t)tri], A);
    load_node(M, M.tris[3 * (int64_t)tri + 1], B);
    load_node(M, M.tris[3 * (int64_t)tri + 2], C);
    double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    double w[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    n[0] = u[1] * w[2] - u[2] * w[1];
    n[1] = u[2] * w[0] - u[0] * w[2];
    n[2] = u[0] * w[1] - u[1] * w[0];
    double nD = (n[0] * D[0] + n[1] * D[1]) + n[2] * D[2];
    double nA = (n[0] * (A[0] - xa[0]) + n[1] * (A[1] - xa[1])) + n[2] * (A[2] - xa[2]);
    s = nA / nD;
}

// Movement records of the tracked particles (ParticleMovementTracker.cpp:4-16):
// one absolute point appended to the slot of load index li, if it has one.
// Points beyond the slot's capacity are counted, not stored.  One thread owns
// a slot at a time (the particle's lane in the step kernels), so no atomics.
// The callers test S.trk_slot (kernel-uniform) and look the slot up here, so
// that with recording off no register is held across the step loop.
__device__ __noinline__ void track_point(const int32_t* slot_of, int32_t* cnt, double* xyz, int32_t max_pts,
                                         uint32_t li, double x, double y, double z)
{
    const int32_t slot = slot_of[li];
    if (slot < 0) return;
    const int32_t c = cnt[slot];
    if (c < max_pts) {
        double* q = xyz + 3 * ((int64_t)slot * max_pts + c);
        q[0] = x;
        q[1] = y;
        q[2] = z;
    }
    cnt[slot] = c + 1;
}

// the position at the start of a step (recorded after the push, before the
// move: ParticleDynamicsProcessor.cpp:182-184)
template <typename R>
__device__ __forceinline__ void track_start(const DevMesh& M, const StepParams& S, uint32_t li, int32_t t, R x, R y,
                                            R z)
{
    double a[3];
    abs_pos(M, t, x, y, z, a);
    track_point(S.trk_slot, S.trk_count, S.trk_xyz, S.trk_max, li, a[0], a[1], a[2]);
}

// the point where the settling segment xa -> (x1 in the frame of t) meets its
// triangle (ParticleSurfaceCollisionHandler.cpp:91-99, :134-138)
template <typename R>
__device__ __forceinline__ void track_hit(const DevMesh& M, const StepParams& S, uint32_t li, const double* xa,
                                          int32_t t, R x1, R y1, R z1, int32_t tri)
{
    double b[3], D[3], n[3], s;
    abs_pos(M, t, x1, y1, z1, b);
    for (int k = 0; k < 3; ++k) D[k] = b[k] - xa[k];
    hit_plane(M, tri, xa, D, n, s);
    track_point(S.trk_slot, S.trk_count, S.trk_xyz, S.trk_max, li, xa[0] + s * D[0], xa[1] + s * D[1],
                xa[2] + s * D[2]);
}

// position in R (FP32 in the owner's frame in mixed mode), velocity FP64:
// accumulated in FP32, a per-tet increment much smaller than |v| rounds the
// same way step after step, and the velocity drifts linearly (about 6e-8 of
// |v| per step)
template <typename R>
struct PState {
    R x, y, z;
    double vx, vy, vz;
    int32_t t, f;
};

// One step of one particle (ParticleDynamicsProcessor.cpp:166-199 stage order,
// then the relocation and charge of the next gather).  Relocations that the
// barycentric maps cannot certify return kDeferNow with the state at x1 in the
// frame of the unchanged owner; the caller resolves them exactly.
// The owning tet's data is re-read each step (L1/L2 hits while the particle
// stays in it) to keep the register footprint of the streaming kernel small.
// The same index, through an empty asm statement (no instruction): the
// compiler cannot reuse the per-lane addresses of the state loads at the start
// of a step for the stores at its end, which kept eight 64-bit addresses live
// across the FP64 step and spilled them to scratch; the stores recompute
// theirs from the index and the kernel-argument bases instead.  (In the mixed
// kernels the allocation without it measured faster.)
__device__ __forceinline__ int64_t fresh_index(int64_t i)
{
    asm volatile("" : "+v"(i));
    return i;
}

struct StepResult {
    int outcome;
    int32_t hit_tri;
};

template <typename R, bool kBatch, bool kWalk, typename Sink>
__device__ __forceinline__ StepResult step_one(const DevMesh& M, const StepParams& S, const SpeciesK& k,
                                               const RunK* __restrict__ rk, uint32_t vid, uint32_t step,
                                               int remaining, PState<R>& s, Sink& sink)
{
    const double tau = tau_of<R>(M);
    // 2. |v| < 10 m/s and 3. outside the expanded bbox: frozen this step and
    //    every later one (nothing changes its state); it still carries charge
    const R vxr = (R)s.vx, vyr = (R)s.vy, vzr = (R)s.vz;
    R vmod = m_sqrt<R>(vxr * vxr + vyr * vyr + vzr * vzr);
    bool frozen = vmod < R(10) || !in_bbox_at(M, S, s.t, s.x, s.y, s.z);
    double lam[4];
    if (frozen) {
        if (S.charge && s.t >= 0 && k.charge != 0.0) {
            bary_at(M, s.t, s.x, s.y, s.z, lam);
            sink.add(M, s.t, k.charge * (double)remaining, lam);
        }
        return StepResult{kFrozenNow, -1};
    }
    R dt = (R)S.dt;
    // 4. push with the owning tet's field (ParticleDynamicsProcessor.cpp:179-180)
    if (S.em_push && s.t >= 0) {
        if (!kBatch && k.push_slot >= 0) {
            // h = (((E*q)/m)*dt)/2 of this tet and species, evaluated per field
            // and time step with the same operations (-0 where the push is
            // skipped, which leaves v unchanged bit for bit)
            R hx, hy, hz;
            load_push(M, k.push_slot, s.t, hx, hy, hz);
            s.vx = (s.vx + (double)hx) + (double)hx;
            s.vy = (s.vy + (double)hy) + (double)hy;
            s.vz = (s.vz + (double)hz) + (double)hz;
        } else if (!kBatch && k.push_slot == -2) {
            // charge 0: with m > 0 and dt > 0 the increment is the signed zero E*q
            R ex, ey, ez;
            load_E(M, s.t, ex, ey, ez);
            if (dt != R(0) && !(ex == R(0) && ey == R(0) && ez == R(0))) {
                const R q = (R)k.charge;
                const double hx = (double)(ex * q), hy = (double)(ey * q), hz = (double)(ez * q);
                s.vx = (s.vx + hx) + hx;
                s.vy = (s.vy + hy) + hy;
                s.vz = (s.vz + hz) + hz;
            }
        } else {
            R ex, ey, ez;
            run_field<R, kBatch>(M, rk, s.t, ex, ey, ez);
            push<R>(ex, ey, ez, (R)k.charge, (R)k.mass, dt, s.vx, s.vy, s.vz);
        }
    }
    // 6. move (Particle.cpp:157-165).  Mixed: the end point in the owner's
    //    frame is rounded to FP32 stochastically, with 30 hashed bits of
    //    (particle, step, seed).
    R x1, y1, z1;
    if constexpr (std::is_same<R, double>::value) {
        x1 = s.x + s.vx * S.dt;
        y1 = s.y + s.vy * S.dt;
        z1 = s.z + s.vz * S.dt;
    } else {
        const uint32_t h = hash3(vid, step, (uint32_t)S.seed ^ rk->key);
        x1 = round_stochastic((double)s.x + s.vx * S.dt, h & 0x3ffu);
        y1 = round_stochastic((double)s.y + s.vy * S.dt, (h >> 10) & 0x3ffu);
        z1 = round_stochastic((double)s.z + s.vz * S.dt, h >> 22);
    }
    if (x1 == s.x && y1 == s.y && z1 == s.z) {  // degenerate segment (:188-190)
        if (S.charge && s.t >= 0 && k.charge != 0.0) {
            bary_at(M, s.t, s.x, s.y, s.z, lam);
            sink.add(M, s.t, k.charge, lam);
        }
        return StepResult{kDone, -1};
    }
    // 7. gas collision (ParticlePhysicsUpdater.cpp:20-24)
    if (S.model != 0) {
        double u1, u1b;
        variates(S.seed, rk->key, vid, step, 0u, u1, u1b);
        if (collides<R>(S, k, rk->conc, (R)s.vx, (R)s.vy, (R)s.vz, u1)) {
            V3 nv = collide_velocity(S.model, &k, S.seed, rk->key, vid, step, s.vx, s.vy, s.vz);
            s.vx = nv.x;
            s.vy = nv.y;
            s.vz = nv.z;
        }
    }
    // 8. relocation and surface hit on the segment x -> x1.  Certified cases:
    //    x1 strictly inside the owner, or strictly outside exactly one face f
    //    and strictly inside the other three (the segment leaves through the
    //    open face f).  The rest needs the exact walk.
    bool certified = false;
    if (s.t >= 0 && (s.f & kStrict)) {
        bary_at(M, s.t, x1, y1, z1, lam);
        int fneg = 0;
        int cls = classify(lam, tau, fneg);
        if (!std::is_same<R, double>::value && cls == kUncertain) {
            // FP32 too close to a face: certify with the FP64 map at the FP64
            // absolute position of the stored FP32 state
            double p1[3];
            abs_pos(M, s.t, x1, y1, z1, p1);
            bary_at(M, s.t, p1[0], p1[1], p1[2], lam);
            cls = classify(lam, M.tau, fneg);
        }
        if (cls == kInside) {
            certified = true;
        } else if (cls == kOneFace) {
            int32_t code = nth(M.tnbr[s.t], fneg);
            if (code <= -2) {
                s.x = x1;  // crosses this surface triangle first
                s.y = y1;
                s.z = z1;
                return StepResult{kSettledNow, -2 - code};
            }
            // to the neighbour (or out through a face with no triangle)
            double p1[3];
            abs_pos(M, s.t, x1, y1, z1, p1);
            R nx, ny, nz;
            to_frame(M, code, p1, nx, ny, nz);
            if (code == -1) {
                s.t = -1;
                s.f &= ~kStrict;
                s.x = nx;
                s.y = ny;
                s.z = nz;
                return StepResult{kDone, -1};
            }
            bary_at(M, code, nx, ny, nz, lam);
            int f2;
            int c2 = classify(lam, tau, f2);
            if (!std::is_same<R, double>::value && c2 == kUncertain) {
                double q[3];
                abs_pos(M, code, nx, ny, nz, q);
                bary_at(M, code, q[0], q[1], q[2], lam);
                c2 = classify(lam, M.tau, f2);
            }
            if (c2 == kInside) {
                s.t = code;
                x1 = nx;
                y1 = ny;
                z1 = nz;
                certified = true;
            }
        }
    }
    if (s.t < 0) {
        // outside the mesh: a segment wholly beyond one face of the nodes'
        // bounding box meets no tet and no surface triangle (exact compares)
        double p0[3], p1[3];
        abs_pos(M, -1, s.x, s.y, s.z, p0);
        abs_pos(M, -1, x1, y1, z1, p1);
        bool away = false;
#pragma unroll
        for (int q = 0; q < 3; ++q)
            away = away || (p0[q] < M.node_lo[q] && p1[q] < M.node_lo[q]) ||
                   (p0[q] > M.node_hi[q] && p1[q] > M.node_hi[q]);
        if (away) {
            s.x = x1;
            s.y = y1;
            s.z = z1;
            return StepResult{kDone, -1};
        }
    }
    if (kWalk && !certified && s.t >= 0 && (s.f & kStrict)) {
        // a segment through several tets: the certified FP64 walk
        double p0[3], p1[3];
        abs_pos(M, s.t, s.x, s.y, s.z, p0);
        abs_pos(M, s.t, x1, y1, z1, p1);
        int32_t where = -1;
        int w = walk64(M, s.t, p0, p1, where);
        if (!std::is_same<R, double>::value && w == kWalkUnsure) {
            // FP32 end points lie on a lattice (~1e-9 m apart in a 1 cm tet),
            // and about 1e-7 of them land exactly on a face plane (one through
            // v0, e.g. local x == y).  Such a point is moved by 1, 2 and 3 FP32
            // ulps along the step in x, y and z (unequal, so that it leaves
            // planes like x == y, which an equal move would keep it on) -- within
            // the FP32 state's own rounding -- and certified again; a point that
            // stays uncertain goes to the exact walk as before.
            const R ux = s.vx < R(0) ? R(-INFINITY) : R(INFINITY);
            const R uy = s.vy < R(0) ? R(-INFINITY) : R(INFINITY);
            const R uz = s.vz < R(0) ? R(-INFINITY) : R(INFINITY);
            x1 = m_nextafter(x1, ux);
            y1 = m_nextafter(m_nextafter(y1, uy), uy);
            z1 = m_nextafter(m_nextafter(m_nextafter(z1, uz), uz), uz);
            abs_pos(M, s.t, x1, y1, z1, p1);
            w = walk64(M, s.t, p0, p1, where);
        }
        if (w == kWalkSurface) {
            s.x = x1;  // crosses this surface triangle first
            s.y = y1;
            s.z = z1;
            return StepResult{kSettledNow, where};
        }
        if (w != kWalkUnsure) {
            R nx, ny, nz;
            to_frame(M, w == kWalkOpen ? -1 : where, p1, nx, ny, nz);
            if (w == kWalkOpen) {
                s.t = -1;
                s.f &= ~kStrict;
                s.x = nx;
                s.y = ny;
                s.z = nz;
                return StepResult{kDone, -1};
            }
            // the position as stored (FP32 in its owner's frame in mixed mode)
            // must itself be strictly inside the owner the walk found
            bary_at(M, where, nx, ny, nz, lam);
            int f2;
            int c2 = classify(lam, tau, f2);
            if (!std::is_same<R, double>::value && c2 == kUncertain) {
                double q[3];
                abs_pos(M, where, nx, ny, nz, q);
                bary_at(M, where, q[0], q[1], q[2], lam);
                c2 = classify(lam, M.tau, f2);
            }
            if (!std::is_same<R, double>::value) {
                // Rounding the end point to FP32 in the new frame can put the
                // stored point (not the segment's end) just beyond a face of
                // `where`; it is then owned by the neighbour across that face,
                // where it is stored rounded in that neighbour's frame.
                for (int tries = 0; tries < 2 && c2 == kOneFace; ++tries) {
                    const int32_t code = nth(M.tnbr[where], f2);
                    if (code < 0) break;  // a surface or open face: the exact walk decides
                    double q[3];
                    abs_pos(M, where, nx, ny, nz, q);
                    where = code;
                    to_frame(M, where, q, nx, ny, nz);
                    bary_at(M, where, nx, ny, nz, lam);
                    c2 = classify(lam, tau, f2);
                    if (c2 == kUncertain) {
                        abs_pos(M, where, nx, ny, nz, q);
                        bary_at(M, where, q[0], q[1], q[2], lam);
                        c2 = classify(lam, M.tau, f2);
                    }
                }
            }
            if (c2 == kInside) {
                s.t = where;
                x1 = nx;
                y1 = ny;
                z1 = nz;
                certified = true;
            }
        }
    }
    if (!certified) {
        s.x = x1;
        s.y = y1;
        s.z = z1;
        return StepResult{kDeferNow, -1};
    }
    s.x = x1;
    s.y = y1;
    s.z = z1;
    if (S.charge && s.t >= 0 && k.charge != 0.0) sink.add(M, s.t, k.charge, lam);
    return StepResult{kDone, -1};
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void prologue_kernel(double* __restrict__ a, double* __restrict__ b, int64_t n_nodes)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = i; j < n_nodes; j += stride) {
        a[j] = 0.0;
        if (b) b[j] = 0.0;
    }
}

// One tile of the streaming kernel: BS consecutive particles, one per thread,
// S.nsteps steps with the state and the owning tet's data in registers.  A
// particle whose relocation cannot be certified stops and goes to the work
// list with the step it stopped at.  Called with a block-uniform tile.
// kTrack: movement records on (S.trk_slot set), a kernel variant of its own so
// that the streaming kernels without them carry none of their code.
template <typename R, int BS, int HashBits, int Win, bool kBatch, bool kWalk, bool kTrack>
__device__ __forceinline__ void step_tile(const DevMesh& M, const DevParticles<R>& P, const StepParams& S,
                                          const SpeciesK* __restrict__ sp, double* __restrict__ node_charge,
                                          WorkItem* __restrict__ wl, unsigned long long* __restrict__ cnt,
                                          unsigned long long* __restrict__ tri_count, int64_t tile, int64_t n_active,
                                          uint32_t step0, unsigned int* blk_processed, int32_t* hk, double* hv,
                                          double* wv, int32_t* s_wbase)
{
    // The node window is used by single-step launches, whose charge all goes
    // to LDS in the flush at the end of the tile (flush_wave_window); fused
    // launches add to LDS while they step and use the hash alone.  (Laying a
    // hash of twice the slots over the unused window for them was no faster.)
    constexpr int kSlots = 1 << HashBits;
    const bool win = Win > 0 && S.nsteps == 1;
    if (threadIdx.x == 0) {
        *blk_processed = 0u;
        if (Win > 0) *s_wbase = INT32_MAX;
    }
    if (S.charge) {
        for (int h = threadIdx.x; h < kSlots; h += BS) {
            hk[h] = -1;
            hv[h] = 0.0;
        }
        if (win)
            for (int h = threadIdx.x; h < Win; h += BS) wv[h] = 0.0;
    }
    __syncthreads();
    int64_t i = tile * BS + threadIdx.x;
    RegSink<LdsSink<HashBits, Win>> sink{LdsSink<HashBits, Win>{hk, hv, wv, INT32_MAX, node_charge}, -1, 0, 0.0,
                                         0.0, 0.0, 0.0};
    int processed = 0;
    bool deferred = false;
    int defer_k = 0;
    double x0abs[3] = {0.0, 0.0, 0.0};
    int32_t t_defer = -1;
    if (i < n_active) {
        PState<R> s;
        s.f = P.flags[i];
        if (!(s.f & kSettled)) {
            const SpeciesK& k = sp[(s.f >> kSpeciesShift) & 0xff];
            // one simulation: run 0 for every lane (scalar loads); a batch: the lane's run
            const RunK* rk = kBatch ? S.runs + ((uint32_t)s.f >> kRunShift) : S.runs;
            sink.base = kBatch ? (int32_t)rk->node_base : 0;
            s.x = P.x[i];
            s.y = P.y[i];
            s.z = P.z[i];
            s.vx = P.vx[i];
            s.vy = P.vy[i];
            s.vz = P.vz[i];
            s.t = P.tet[i];
            const int32_t t_in = s.t, f_in = s.f;
            // variate counter: collisions need it; the mixed path's stochastic
            // rounding keys on it too
            const uint32_t vid = (S.model != 0 || !std::is_same<R, double>::value) ? P.pidx[i] - rk->first : 0u;
            bool moved = false;
            for (int ks = 0; ks < S.nsteps; ++ks) {
                const R px = s.x, py = s.y, pz = s.z;
                const int32_t t_before = s.t;
                StepResult r = step_one<R, kBatch, kWalk>(M, S, k, rk, vid, step0 + (uint32_t)ks, S.nsteps - ks, s, sink);
                if (r.outcome == kFrozenNow) break;
                ++processed;
                moved = true;
                if (kTrack) track_start<R>(M, S, P.pidx[i], t_before, px, py, pz);
                if (r.outcome == kSettledNow) {
                    s.f = (s.f | kSettled) & ~kStrict;
                    P.hit[i] = r.hit_tri;
                    atomicAdd(&tri_count[rk->tri_base + r.hit_tri], 1ull);
                    S.settle_log[atomicAdd(&cnt[kCntSettled], 1ull)] = P.pidx[i];
                    if (S.settle || kTrack) {
                        double a[3];
                        abs_pos(M, t_before, px, py, pz, a);
                        if (S.settle) record_settle(S.settle, P.pidx[i], step0 + (uint32_t)ks, a[0], a[1], a[2]);
                        if (kTrack) track_hit<R>(M, S, P.pidx[i], a, s.t, s.x, s.y, s.z, r.hit_tri);
                    }
                    break;
                }
                if (r.outcome == kDeferNow) {
                    deferred = true;
                    defer_k = ks;
                    t_defer = t_before;
                    abs_pos(M, t_before, px, py, pz, x0abs);
                    break;
                }
            }
            if (moved) {
                const int64_t j = std::is_same<R, double>::value ? fresh_index(i) : i;
                P.x[j] = s.x;
                P.y[j] = s.y;
                P.z[j] = s.z;
                P.vx[j] = s.vx;
                P.vy[j] = s.vy;
                P.vz[j] = s.vz;
                if (s.t != t_in) P.tet[j] = s.t;
                if (s.f != f_in) P.flags[j] = s.f;
            }
        }
    }
    if (S.charge) {  // every lane is active here
        if (Win > 0 && S.nsteps == 1) sink.flush_wave_window(M, s_wbase);
        else sink.flush_wave(M);
    }
    // work list: one atomic per wavefront, lanes take consecutive slots
    unsigned long long dm = __ballot(deferred);
    if (dm) {
        int lane = threadIdx.x & 63;
        int leader = __ffsll((long long)dm) - 1;
        unsigned int base = 0u;
        if (lane == leader)
            base = (unsigned int)atomicAdd(&cnt[kCntWork + 2 * S.work_slot], (unsigned long long)__popcll(dm));
        base = __shfl(base, leader);
        if (deferred) {
            WorkItem it;
            it.x0 = x0abs[0];
            it.y0 = x0abs[1];
            it.z0 = x0abs[2];
            it.i = (int32_t)i;
            it.t0 = t_defer;
            it.k = defer_k;
            it.pad = 0;
            wl[base + (unsigned)__popcll(dm & ((1ull << lane) - 1ull))] = it;
        }
    }
    // processed count: wavefront sum, LDS per block, one atomic per tile spread over slots
    __syncthreads();
    int wsum;
    if (Win > 0) {  // single-step launches: 0 or 1 per lane
        wsum = (int)__popcll(__ballot(processed != 0));
    } else {
        wsum = processed;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) wsum += __shfl_xor(wsum, off);
    }
    if ((threadIdx.x & 63) == 0 && wsum) atomicAdd(blk_processed, (unsigned int)wsum);
    __syncthreads();
    if (threadIdx.x == 0 && *blk_processed)
        atomicAdd(&cnt[kCntProcSlot0 + (tile & (kProcSlots - 1))], (unsigned long long)*blk_processed);
    if (S.charge) {
        for (int h = threadIdx.x; h < kSlots; h += BS) {
            int32_t node = hk[h];
            if (node >= 0) global_add_f64(&node_charge[node], hv[h]);
        }
        // window slots that received charge (untouched ones hold 0.0)
        if (win) {
            const int32_t wbase = sink.inner.wbase;
            for (int h = threadIdx.x; h < Win; h += BS) {
                const double v = wv[h];
                if (v != 0.0) global_add_f64(&node_charge[wbase + h], v);
            }
        }
    }
    __syncthreads();  // the LDS table and counter are reused by the next tile
}

// Streaming kernel.  With persistent == 0 workgroup b takes tile b; otherwise
// a grid sized to the chip pulls tiles from a queue counter until the live
// prefix is covered (one launch for every run of a batch, whatever their
// sizes).  The charge of a tile is summed in LDS (node window and hash) and
// flushed with one FP64 atomic per distinct node.
template <typename R, int BS, bool kBatch, bool kPersistent, bool kWalk, bool kTrack = false, bool kWindow = false>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) step_main(DevMesh M, DevParticles<R> P, StepParams S,
                                                       const SpeciesK* __restrict__ sp,
                                                       double* __restrict__ node_charge, WorkItem* __restrict__ wl,
                                                       unsigned long long* __restrict__ cnt,
                                                       unsigned long long* __restrict__ tri_count)
{
    // kWindow (single-step launches of one simulation, 256 threads): a
    // 3,072-node window (24 KB) and a 1,024-slot hash (12 KB), 36 KB in all,
    // so that four workgroups (the 16 waves the 128-VGPR budget allows) share
    // a CU's 160 KB.  Otherwise the hash alone, 8 slots per thread: fused
    // launches add to LDS while they step, and batches interleave their runs
    // in the tet order.  (The window variant run fused was 5% slower on the
    // 3M-tet chamber, whose node layers lie 6,600 ids apart.)
    static_assert(!kWindow || (BS == 256 && !kBatch), "the window is for 256-thread single-run launches");
    constexpr int kWin = kWindow ? 3072 : 0;
    constexpr int kBits = kWin > 0 ? 10 : (BS == 256 ? 11 : 10);
    __shared__ unsigned int blk_processed;
    __shared__ int64_t s_tile;
    __shared__ int32_t s_wbase;
    __shared__ int32_t hk[1 << kBits];
    __shared__ double hv[1 << kBits];
    __shared__ double wv[kWin > 0 ? kWin : 1];
    const uint32_t step0 = (uint32_t)cnt[kCntStep + S.work_slot];
    {
        // clear the next launch's charge buffers and counters, and give it its first step
        int64_t g = blockIdx.x * (int64_t)BS + threadIdx.x;
        int64_t gs = (int64_t)gridDim.x * BS;
        if (S.zero_charge)
            for (int64_t j = g; j < S.charge_len; j += gs) S.zero_charge[j] = 0.0;
        if (g == 0) {
            cnt[kCntWork + 2 * (1 - S.work_slot)] = 0ull;
            cnt[kCntStep + (1 - S.work_slot)] = (unsigned long long)step0 + (unsigned long long)S.nsteps;
            cnt[kCntTile + (1 - S.work_slot)] = 0ull;
        }
    }
    const int64_t n_active = *S.n_active;
    if (!kPersistent) {
        if ((int64_t)blockIdx.x * BS < n_active)  // else the whole block is past the live particles
            step_tile<R, BS, kBits, kWin, kBatch, kWalk, kTrack>(M, P, S, sp, node_charge, wl, cnt, tri_count,
                                                                 blockIdx.x, n_active, step0, &blk_processed, hk, hv,
                                                                 wv, &s_wbase);
        return;
    }
    for (;;) {
        if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(&cnt[kCntTile + S.work_slot], 1ull);
        __syncthreads();
        const int64_t tile = s_tile;
        if (tile * BS >= n_active) break;  // queue drained: every workgroup exits here
        step_tile<R, BS, kBits, kWin, kBatch, kWalk, kTrack>(M, P, S, sp, node_charge, wl, cnt, tri_count, tile,
                                                             n_active, step0, &blk_processed, hk, hv, wv, &s_wbase);
    }
}

// Exact relocation of a deferred step: the state holds x1 in the frame of the
// owner at the step start; x0abs is the absolute start.  Returns true when the
// particle settled.
template <typename R, typename Sink>
__device__ inline bool finish_exact(const DevMesh& M, const StepParams& S, const SpeciesK& k, const double* x0abs,
                                    PState<R>& s, Sink& sink, int32_t& hit_tri, int& exh, int& err)
{
    double x1[3];
    abs_pos(M, s.t, s.x, s.y, s.z, x1);
    Resolved r = resolve(M, x0abs, x1, s.t, (s.f & kStrict) != 0);
    exh += r.exhaustive;
    err |= r.err;
    if (r.hit >= 0) {
        hit_tri = r.hit;
        return true;
    }
    to_frame(M, r.tet, x1, s.x, s.y, s.z);
    int strict = r.strict;
    if (!std::is_same<R, double>::value && r.tet >= 0) {
        // the stored (rounded) position decides the fast path next step
        double q[3];
        abs_pos(M, r.tet, s.x, s.y, s.z, q);
        strict = tet_contains(M, r.tet, q) == 2;
    }
    s.t = r.tet;
    s.f = strict ? (s.f | kStrict) : (s.f & ~kStrict);
    if (S.charge && s.t >= 0 && k.charge != 0.0) {
        double lam[4];
        bary_at(M, s.t, s.x, s.y, s.z, lam);
        sink.add(M, s.t, k.charge, lam);
    }
    return false;
}

// Straggler kernel: finishes each deferred relocation exactly, then runs the
// particle's remaining steps of the launch with the certified fast path,
// resolving exactly again whenever a step defers.  (Holding it to 128 VGPRs
// for 4 waves per SIMD spills ~300 registers and was slower than 1 wave.)
template <typename R, bool kWalk>
__global__ void __launch_bounds__(64) step_resolve(DevMesh M, DevParticles<R> P, StepParams S,
                                                   const SpeciesK* __restrict__ sp, double* __restrict__ node_charge,
                                                   const WorkItem* __restrict__ wl,
                                                   unsigned long long* __restrict__ cnt,
                                                   unsigned long long* __restrict__ tri_count)
{
    const uint64_t nw = cnt[kCntWork + 2 * S.work_slot];
    const uint32_t step0 = (uint32_t)cnt[kCntStep + S.work_slot];
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nw;
         w += (uint64_t)gridDim.x * blockDim.x) {
        RegSink<GlobalSink> sink{GlobalSink{node_charge}, -1, 0, 0.0, 0.0, 0.0, 0.0};
        WorkItem it = wl[w];
        int64_t i = it.i;
        PState<R> s;
        s.f = P.flags[i];
        s.t = it.t0;
        s.x = P.x[i];
        s.y = P.y[i];
        s.z = P.z[i];
        s.vx = P.vx[i];
        s.vy = P.vy[i];
        s.vz = P.vz[i];
        const SpeciesK& k = sp[(s.f >> kSpeciesShift) & 0xff];
        const RunK* rk = S.runs + ((uint32_t)s.f >> kRunShift);
        sink.base = (int32_t)rk->node_base;
        int exh = 0, err = 0, nexact = 0;
        int32_t hit_tri = -1;
        double x0[3] = {it.x0, it.y0, it.z0};
        // (the step's start was recorded by the streaming kernel)
        bool settled = finish_exact(M, S, k, x0, s, sink, hit_tri, exh, err);
        if (settled && S.trk_slot) track_hit<R>(M, S, P.pidx[i], x0, s.t, s.x, s.y, s.z, hit_tri);
        uint32_t settle_step = step0 + (uint32_t)it.k;  // step and start of the settling segment
        double sx = x0[0], sy = x0[1], sz = x0[2];
        if (!settled) {
            const uint32_t vid = (S.model != 0 || !std::is_same<R, double>::value) ? P.pidx[i] - rk->first : 0u;
            unsigned long long extra = 0;
            for (int ks = it.k + 1; ks < S.nsteps; ++ks) {
                const R px = s.x, py = s.y, pz = s.z;
                const int32_t t_before = s.t;
                StepResult rr = step_one<R, true, kWalk>(M, S, k, rk, vid, step0 + (uint32_t)ks, S.nsteps - ks, s, sink);
                if (rr.outcome == kFrozenNow) break;
                ++extra;
                if (S.trk_slot) track_start<R>(M, S, P.pidx[i], t_before, px, py, pz);
                if (rr.outcome == kDeferNow || rr.outcome == kSettledNow) {
                    double x0s[3];
                    abs_pos(M, t_before, px, py, pz, x0s);
                    if (rr.outcome == kDeferNow) {
                        ++nexact;
                        settled = finish_exact(M, S, k, x0s, s, sink, hit_tri, exh, err);
                    } else {
                        hit_tri = rr.hit_tri;
                        settled = true;
                    }
                    if (settled && S.trk_slot) track_hit<R>(M, S, P.pidx[i], x0s, s.t, s.x, s.y, s.z, hit_tri);
                    settle_step = step0 + (uint32_t)ks;
                    sx = x0s[0];
                    sy = x0s[1];
                    sz = x0s[2];
                }
                if (settled) break;
            }
            if (extra) atomicAdd(&cnt[kCntProcSlot0 + (w & (kProcSlots - 1))], extra);
        }
        if (settled) {
            s.f = (s.f | kSettled) & ~kStrict;
            P.hit[i] = hit_tri;
            atomicAdd(&tri_count[rk->tri_base + hit_tri], 1ull);
            S.settle_log[atomicAdd(&cnt[kCntSettled], 1ull)] = P.pidx[i];
            if (S.settle) record_settle(S.settle, P.pidx[i], settle_step, sx, sy, sz);
        }
        sink.flush(M);
        P.x[i] = s.x;
        P.y[i] = s.y;
        P.z[i] = s.z;
        P.vx[i] = s.vx;
        P.vy[i] = s.vy;
        P.vz[i] = s.vz;
        P.tet[i] = s.t;
        P.flags[i] = s.f;
        if (err) atomicOr(&cnt[kCntError], 1ull);
        if (exh) atomicAdd(&cnt[kCntExhaustive], (unsigned long long)exh);
        if (nexact) atomicAdd(&cnt[kCntDeferred], (unsigned long long)nexact);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&cnt[kCntDeferred], (unsigned long long)nw);
}

// Owners of absolute positions (min-id closed containment); with `lx` the
// positions are also written in the frame of their owner (mixed layout).
__global__ void locate_kernel(DevMesh M, const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ z, int64_t n, int32_t* __restrict__ tet,
                              int32_t* __restrict__ flags, float* __restrict__ lx, float* __restrict__ ly,
                              float* __restrict__ lz, unsigned long long* __restrict__ cnt)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p[3] = {x[i], y[i], z[i]};
    int strict = 0, err = 0;
    int32_t t = locate_grid(M, p, &strict, &err);
    if (lx) {
        float a, b, c;
        to_frame(M, t, p, a, b, c);
        lx[i] = a;
        ly[i] = b;
        lz[i] = c;
        if (t >= 0) {
            double q[3];
            abs_pos(M, t, a, b, c, q);
            int r = tet_contains(M, t, q);
            if (r < 0) err = 1;
            strict = r == 2;
        }
    }
    if (err && cnt) atomicOr(&cnt[kCntError], 1ull);
    tet[i] = t;
    if (flags) flags[i] = strict ? (flags[i] | kStrict) : (flags[i] & ~kStrict);
}

template <typename R>
__global__ void gather_kernel(DevMesh M, DevParticles<R> P, const SpeciesK* __restrict__ sp,
                              const RunK* __restrict__ runs, double* __restrict__ node_charge)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    int32_t f = P.flags[i];
    int32_t t = P.tet[i];
    if ((f & kSettled) || t < 0) return;
    double q = sp[(f >> kSpeciesShift) & 0xff].charge;
    if (q == 0.0) return;
    double lam[4];
    bary_at(M, t, P.x[i], P.y[i], P.z[i], lam);
    scatter4(node_charge + runs[(uint32_t)f >> kRunShift].node_base, M.tnode[t], q, lam);
}

// Occupancy and summed charge of each tet (one owner per unsettled particle).
template <typename R>
__global__ void tet_occupancy_kernel(DevParticles<R> P, const SpeciesK* __restrict__ sp, double* __restrict__ tet_q,
                                     int32_t* __restrict__ tet_n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    int32_t f = P.flags[i], t = P.tet[i];
    if ((f & kSettled) || t < 0) return;
    atomicAdd(&tet_n[t], 1);
    double q = sp[(f >> kSpeciesShift) & 0xff].charge;
    if (q != 0.0) global_add_f64(&tet_q[t], q);
}

// rho_n = sum(rho_t V_t) / sum V_t over the occupied tets around n, ascending tet id
// (NodeChargeDensityProcessor.cpp:69-93).
__global__ void node_density_kernel(const int32_t* __restrict__ adj_start, const int32_t* __restrict__ adj_list,
                                    const double* __restrict__ tet_vol, const double* __restrict__ tet_q,
                                    const int32_t* __restrict__ tet_n, int64_t n_nodes, double* __restrict__ rho,
                                    int32_t* __restrict__ written)
{
    int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    double tc = 0.0, tv = 0.0;
    for (int32_t j = adj_start[n]; j < adj_start[n + 1]; ++j) {
        int32_t t = adj_list[j];
        if (tet_n[t] == 0) continue;
        double v = tet_vol[t];
        double rt = tet_q[t] / v;
        tc += rt * v;
        tv += v;
    }
    rho[n] = tv > 0.0 ? tc / tv : 0.0;
    if (written) written[n] = tv > 0.0;
}

__global__ void node_density_bary_kernel(const double* __restrict__ q, const double* __restrict__ node_vol,
                                         int64_t n_nodes, double* __restrict__ rho, int32_t* __restrict__ written)
{
    int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    rho[n] = q[n] / node_vol[n];
    if (written) written[n] = 1;
}

// E = -grad(phi) = -sum_j (phi_j - phi_0) A_j  (A_j: gradient of lambda_j)
__global__ void field_kernel(DevMesh M, const double* __restrict__ phi, double* __restrict__ E, float* __restrict__ E32)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= M.n_tets) return;
    int4 nv = M.tnode[t];
    const double* g = M.geo + 12 * t;
    double p0 = phi[nv.x], d1 = phi[nv.y] - p0, d2 = phi[nv.z] - p0, d3 = phi[nv.w] - p0;
    double ex = -((d1 * g[3] + d2 * g[6]) + d3 * g[9]);
    double ey = -((d1 * g[4] + d2 * g[7]) + d3 * g[10]);
    double ez = -((d1 * g[5] + d2 * g[8]) + d3 * g[11]);
    E[3 * t] = ex;
    E[3 * t + 1] = ey;
    E[3 * t + 2] = ez;
    if (E32) {
        E32[4 * t] = (float)ex;
        E32[4 * t + 1] = (float)ey;
        E32[4 * t + 2] = (float)ez;
        E32[4 * t + 3] = 0.0f;
    }
}

// Push increments h = (((E*q)/m)*dt)/2 of field basis 0 for each charged
// species slot (push<R>'s arithmetic, FP64 from E and FP32 from the FP32 field),
// -0 where push<R> returns early (dt == 0 or E == 0 exactly).
struct PushSlots {
    double q[kMaxPushSlots], m[kMaxPushSlots];
    int32_t n;
};
template <typename R>
__device__ inline R push_increment(R e, R q, R m, R dt)
{
    return (((e * q) / m) * dt) / R(2);
}
__global__ void push_table_kernel(const double* __restrict__ E, const float* __restrict__ E32, int64_t nt,
                                  PushSlots ps, double dt, double* __restrict__ H, float* __restrict__ H32)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nt) return;
    for (int k = 0; k < ps.n; ++k) {
        if (H) {
            const double ex = E[3 * t], ey = E[3 * t + 1], ez = E[3 * t + 2];
            const bool skip = dt == 0.0 || (ex == 0.0 && ey == 0.0 && ez == 0.0);
            double* h = H + 3 * ((int64_t)k * nt + t);
            h[0] = skip ? -0.0 : push_increment<double>(ex, ps.q[k], ps.m[k], dt);
            h[1] = skip ? -0.0 : push_increment<double>(ey, ps.q[k], ps.m[k], dt);
            h[2] = skip ? -0.0 : push_increment<double>(ez, ps.q[k], ps.m[k], dt);
        }
        if (H32) {
            const float ex = E32[4 * t], ey = E32[4 * t + 1], ez = E32[4 * t + 2];
            const float q = (float)ps.q[k], m = (float)ps.m[k], d = (float)dt;
            const bool skip = d == 0.0f || (ex == 0.0f && ey == 0.0f && ez == 0.0f);
            float4 h;
            h.x = skip ? -0.0f : push_increment<float>(ex, q, m, d);
            h.y = skip ? -0.0f : push_increment<float>(ey, q, m, d);
            h.z = skip ? -0.0f : push_increment<float>(ez, q, m, d);
            h.w = 0.0f;
            reinterpret_cast<float4*>(H32)[(int64_t)k * nt + t] = h;
        }
    }
}

void launch_push_table(hipStream_t s, const double* E, const float* E32, int64_t nt, const double* q,
                       const double* m, int32_t n_slots, double dt, double* H, float* H32)
{
    PushSlots ps{};
    ps.n = n_slots;
    for (int k = 0; k < n_slots && k < kMaxPushSlots; ++k) {
        ps.q[k] = q[k];
        ps.m[k] = m[k];
    }
    if (nt > 0 && n_slots > 0)
        push_table_kernel<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(E, E32, nt, ps, dt, H, H32);
}

__global__ void e32_kernel(const double* __restrict__ E, float* __restrict__ E32, int64_t nt)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nt) return;
    E32[4 * t] = (float)E[3 * t];
    E32[4 * t + 1] = (float)E[3 * t + 1];
    E32[4 * t + 2] = (float)E[3 * t + 2];
    E32[4 * t + 3] = 0.0f;
}

__global__ void orient_batch_kernel(const double* __restrict__ pts, int64_t n, int32_t* __restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* p = pts + 12 * i;
    out[i] = orient3d(p, p + 3, p + 6, p + 9);
}

// the shared elementary functions (dm_*) over n arguments, for checks:
// which = 0 pow(x, y), 1 sin x and cos x, 2 acos x, 3 atan2(y, x), 4 log x, 5 exp x
__global__ void math_batch_kernel(int which, const double* __restrict__ x, const double* __restrict__ y, int64_t n,
                                  double* __restrict__ out, double* __restrict__ out2)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = 0.0, r2 = 0.0;
    if (which == 0) r = dm_pow(x[i], y[i]);
    else if (which == 1) dm_sincos(x[i], &r, &r2);
    else if (which == 2) r = dm_acos(x[i]);
    else if (which == 3) r = dm_atan2(y[i], x[i]);
    else if (which == 4) r = dm_log(x[i]);
    else r = dm_exp(x[i]);
    out[i] = r;
    out2[i] = r2;
}

// absolute FP64 positions of the mixed layout, for read-back
__global__ void abs_positions_kernel(DevMesh M, const float* __restrict__ lx, const float* __restrict__ ly,
                                     const float* __restrict__ lz, const int32_t* __restrict__ tet, int64_t n,
                                     double* __restrict__ x, double* __restrict__ y, double* __restrict__ z)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p[3];
    abs_pos(M, tet[i], lx[i], ly[i], lz[i], p);
    x[i] = p[0];
    y[i] = p[1];
    z[i] = p[2];
}

// ---------------------------------------------------------------------------
// Adjoint of the collisionless push and the surface-hit map.
//
// Forward, step k with owner t_k of x_k:
//   v_{k+1} = v_k + a(t_k) dt,  a = (q/m) (s0 E_0 + s1 E_1),  x_{k+1} = x_k + v_{k+1} dt.
// The particle settles in step k* on triangle T (plane n.x = n.A); its hit
// point is x_h = x_a + s D with x_a = x_{k*}, D = x_{k*+1} - x_a and
// s = n.(A - x_a) / n.D.  Linearised: dx_h = P (dx_a + s dt dv_{k*+1}) with
// P = I - D n^T / (n.D).  For a loss L with lambda = dL/dx_h the reverse sweep is
//   lx = P^T lambda                    adjoint of every x_k (E is constant per tet)
//   mu_k* = s dt P^T lambda            adjoint of v_{k*+1}
//   mu_{k-1} = mu_k + dt lx
//   dL/ds_j += mu_k . (q/m) E_j(t_k) dt,   k = k*, ..., k0.
// The owners t_k are recomputed backwards -- x_{k-1} = x_k - v_k dt with v_k
// the push of t_k undone -- instead of storing the trajectory.
// ---------------------------------------------------------------------------

// Which particles the next step records a movement for, by load index: not
// settled, |v| >= 10 m/s and inside the expanded bounding box (the filters of
// ParticleDynamicsProcessor.cpp:166-176, as step_one applies them).
template <typename R>
__global__ void track_eligible_kernel(DevMesh M, DevParticles<R> P, StepParams S, uint8_t* __restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int32_t f = P.flags[i];
    const R vx = (R)P.vx[i], vy = (R)P.vy[i], vz = (R)P.vz[i];
    const R vmod = m_sqrt<R>(vx * vx + vy * vy + vz * vz);
    const bool live = !(f & kSettled) && !(vmod < R(10)) && in_bbox_at(M, S, P.tet[i], P.x[i], P.y[i], P.z[i]);
    out[P.pidx[i]] = live ? 1 : 0;
}

// hit points, triangles and settle steps of the settled particles, in load order
template <typename R>
__global__ void hit_points_kernel(DevMesh M, DevParticles<R> P, const SettleRec* __restrict__ rec,
                                  double* __restrict__ xh, int32_t* __restrict__ tri, int32_t* __restrict__ step)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    uint32_t p = P.pidx[i];
    SettleRec r = rec[p];
    if (!(P.flags[i] & kSettled) || r.step < 0) {
        xh[3 * (int64_t)p] = xh[3 * (int64_t)p + 1] = xh[3 * (int64_t)p + 2] = 0.0;
        tri[p] = -1;
        step[p] = -1;
        return;
    }
    double xb[3], xa[3] = {r.x0, r.y0, r.z0}, D[3], n[3], s;
    abs_pos(M, P.tet[i], P.x[i], P.y[i], P.z[i], xb);
    for (int k = 0; k < 3; ++k) D[k] = xb[k] - xa[k];
    hit_plane(M, P.hit[i], xa, D, n, s);
    for (int k = 0; k < 3; ++k) xh[3 * (int64_t)p + k] = xa[k] + s * D[k];
    tri[p] = P.hit[i];
    step[p] = r.step;
}

template <typename R>
__global__ void __launch_bounds__(256) adjoint_push_kernel(DevMesh M, DevParticles<R> P,
                                                           const SpeciesK* __restrict__ sp,
                                                           const RunK* __restrict__ runs, double dt, int32_t k0,
                                                           const SettleRec* __restrict__ rec,
                                                           const double* __restrict__ lambda,
                                                           double* __restrict__ grad,
                                                           unsigned long long* __restrict__ cnt)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double g0 = 0.0, g1 = 0.0;
    int err = 0;
    if (i < P.n) {
        const int32_t f = P.flags[i];
        const uint32_t p = P.pidx[i];
        const double lam[3] = {lambda[3 * (int64_t)p], lambda[3 * (int64_t)p + 1], lambda[3 * (int64_t)p + 2]};
        const SettleRec r = rec[p];
        if ((f & kSettled) && r.step >= k0 && (lam[0] != 0.0 || lam[1] != 0.0 || lam[2] != 0.0)) {
            const SpeciesK& k = sp[(f >> kSpeciesShift) & 0xff];
            const RunK& rk = runs[(uint32_t)f >> kRunShift];
            const double qm = k.charge / k.mass;
            int32_t t = P.tet[i];
            double xb[3], x[3] = {r.x0, r.y0, r.z0}, D[3], n[3], s;
            abs_pos(M, t, P.x[i], P.y[i], P.z[i], xb);
            for (int c = 0; c < 3; ++c) D[c] = xb[c] - x[c];
            hit_plane(M, P.hit[i], x, D, n, s);
            const double nD = (n[0] * D[0] + n[1] * D[1]) + n[2] * D[2];
            const double Dl = (D[0] * lam[0] + D[1] * lam[1]) + D[2] * lam[2];
            double pl[3], mu[3];
            for (int c = 0; c < 3; ++c) {
                pl[c] = lam[c] - n[c] * (Dl / nD);
                mu[c] = (s * dt) * pl[c];
            }
            double v[3] = {(double)P.vx[i], (double)P.vy[i], (double)P.vz[i]};
            for (int32_t kk = r.step;; --kk) {
                double e0[3] = {0.0, 0.0, 0.0}, e1[3] = {0.0, 0.0, 0.0};
                if (t >= 0) {
                    for (int c = 0; c < 3; ++c) e0[c] = M.E[3 * (int64_t)t + c];
                    if (M.E1)
                        for (int c = 0; c < 3; ++c) e1[c] = M.E1[3 * (int64_t)t + c];
                }
                g0 += (((mu[0] * e0[0] + mu[1] * e0[1]) + mu[2] * e0[2]) * qm) * dt;
                g1 += (((mu[0] * e1[0] + mu[1] * e1[1]) + mu[2] * e1[2]) * qm) * dt;
                if (kk <= k0) break;
                if (t >= 0) {  // undo the push of step kk: forward (v + h) + h
                    double E[3];
                    for (int c = 0; c < 3; ++c) E[c] = rk.s0 * e0[c] + rk.s1 * e1[c];
                    if (!(E[0] == 0.0 && E[1] == 0.0 && E[2] == 0.0)) {
                        for (int c = 0; c < 3; ++c) {
                            double h = (((E[c] * k.charge) / k.mass) * dt) / 2.0;
                            v[c] = (v[c] - h) - h;
                        }
                    }
                }
                for (int c = 0; c < 3; ++c) x[c] = x[c] - v[c] * dt;
                int strict = 0;
                t = t >= 0 ? locate_near(M, t, x, &strict, &err) : locate_grid(M, x, &strict, &err);
                for (int c = 0; c < 3; ++c) mu[c] = mu[c] + dt * pl[c];
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        g0 += __shfl_xor(g0, off);
        g1 += __shfl_xor(g1, off);
    }
    if ((threadIdx.x & 63) == 0 && (g0 != 0.0 || g1 != 0.0)) {
        global_add_f64(&grad[0], g0);
        global_add_f64(&grad[1], g1);
    }
    if (err) atomicOr(&cnt[kCntError], 1ull);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_zero_charge(hipStream_t s, double* a, double* b, int64_t n_nodes)
{
    unsigned nb = std::max(1u, std::min(blocks_for(n_nodes, 256), 2048u));
    prologue_kernel<<<nb, 256, 0, s>>>(a, b, n_nodes);
}

template <typename R>
void launch_step(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const StepParams& S,
                 const SpeciesK* sp, double* node_charge, WorkItem* wl, unsigned long long* cnt,
                 unsigned long long* tri_count, const StepLaunch& L, hipEvent_t ev_mid)
{
    if (P.n > 0 && S.trk_slot) {
        // movement records: the tracking variants (256-thread workgroups, one per tile)
        const unsigned grid = blocks_for(P.n, 256);
        switch ((L.batch ? 2 : 0) + (L.walk ? 1 : 0)) {
#define SG_STEP_TRACK(B_, W_) \
    step_main<R, 256, B_, false, W_, true><<<grid, 256, 0, s>>>(M, P, S, sp, node_charge, wl, cnt, tri_count)
        case 0: SG_STEP_TRACK(false, false); break;
        case 1: SG_STEP_TRACK(false, true); break;
        case 2: SG_STEP_TRACK(true, false); break;
        default: SG_STEP_TRACK(true, true); break;
#undef SG_STEP_TRACK
        }
    } else if (P.n > 0 && S.nsteps == 1 && !L.persistent && !L.batch && (L.walk || L.block != 128)) {
        // single-step launches of one simulation: the node-window variants
        const unsigned grid = blocks_for(P.n, 256);
        if (L.walk)
            step_main<R, 256, false, false, true, false, true>
                <<<grid, 256, 0, s>>>(M, P, S, sp, node_charge, wl, cnt, tri_count);
        else
            step_main<R, 256, false, false, false, false, true>
                <<<grid, 256, 0, s>>>(M, P, S, sp, node_charge, wl, cnt, tri_count);
    } else if (P.n > 0) {
        const unsigned grid = L.persistent ? (unsigned)L.grid : blocks_for(P.n, L.block == 128 ? 128 : 256);
        // the multi-tet walk is instantiated for 256-thread workgroups only
        // (the host sets block = 256 with it)
        const int v = L.walk ? 8 + (L.batch ? 2 : 0) + (L.persistent ? 1 : 0)
                             : (L.block == 128 ? 4 : 0) + (L.batch ? 2 : 0) + (L.persistent ? 1 : 0);
        switch (v) {
#define SG_STEP_MAIN(BS_, B_, P_, W_) \
    step_main<R, BS_, B_, P_, W_><<<grid, BS_, 0, s>>>(M, P, S, sp, node_charge, wl, cnt, tri_count)
        case 0: SG_STEP_MAIN(256, false, false, false); break;
        case 1: SG_STEP_MAIN(256, false, true, false); break;
        case 2: SG_STEP_MAIN(256, true, false, false); break;
        case 3: SG_STEP_MAIN(256, true, true, false); break;
        case 4: SG_STEP_MAIN(128, false, false, false); break;
        case 5: SG_STEP_MAIN(128, false, true, false); break;
        case 6: SG_STEP_MAIN(128, true, false, false); break;
        case 7: SG_STEP_MAIN(128, true, true, false); break;
        case 8: SG_STEP_MAIN(256, false, false, true); break;
        case 9: SG_STEP_MAIN(256, false, true, true); break;
        case 10: SG_STEP_MAIN(256, true, false, true); break;
        default: SG_STEP_MAIN(256, true, true, true); break;
#undef SG_STEP_MAIN
        }
    }
    if (ev_mid) (void)hipEventRecord(ev_mid, s);
    // the work list is only known on the device: a grid of up to 1,024
    // single-wave workgroups -- one per SIMD, all that can be resident at the
    // kernel's 300+ VGPRs -- and the rest of the list by grid stride; 64 for
    // a single step, whose list is short (one exact relocation per deferred
    // particle) and mostly empty.  (4,096 cost 10 us per launch with an empty
    // list, 13% of a single step; 64 saves 0.3 us per step against 256.)
    const unsigned cap = S.nsteps == 1 ? 64u : 1024u;
    unsigned nb = std::max(1u, std::min(blocks_for(P.n, 64), cap));
    if (L.walk)
        step_resolve<R, true><<<nb, 64, 0, s>>>(M, P, S, sp, node_charge, wl, cnt, tri_count);
    else
        step_resolve<R, false><<<nb, 64, 0, s>>>(M, P, S, sp, node_charge, wl, cnt, tri_count);
}

void launch_locate(hipStream_t s, const DevMesh& M, const double* x, const double* y, const double* z, int64_t n,
                   int32_t* tet, int32_t* flags, float* lx, float* ly, float* lz, unsigned long long* cnt)
{
    if (n > 0) locate_kernel<<<blocks_for(n, 128), 128, 0, s>>>(M, x, y, z, n, tet, flags, lx, ly, lz, cnt);
}

template <typename R>
void launch_gather(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const SpeciesK* sp, const RunK* runs,
                   double* node_charge)
{
    if (P.n > 0) gather_kernel<R><<<blocks_for(P.n, 256), 256, 0, s>>>(M, P, sp, runs, node_charge);
}

template <typename R>
void launch_hit_points(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const SettleRec* rec, double* xh,
                       int32_t* tri, int32_t* step)
{
    if (P.n > 0) hit_points_kernel<R><<<blocks_for(P.n, 256), 256, 0, s>>>(M, P, rec, xh, tri, step);
}

template <typename R>
void launch_track_eligible(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const StepParams& S,
                           uint8_t* out)
{
    if (P.n > 0) track_eligible_kernel<R><<<blocks_for(P.n, 256), 256, 0, s>>>(M, P, S, out);
}
template void launch_track_eligible<double>(hipStream_t, const DevMesh&, const DevParticles<double>&,
                                            const StepParams&, uint8_t*);
template void launch_track_eligible<float>(hipStream_t, const DevMesh&, const DevParticles<float>&,
                                           const StepParams&, uint8_t*);

template <typename R>
void launch_adjoint_push(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const SpeciesK* sp,
                         const RunK* runs, double dt, int32_t k0, const SettleRec* rec, const double* lambda,
                         double* grad, unsigned long long* cnt)
{
    if (P.n > 0)
        adjoint_push_kernel<R><<<blocks_for(P.n, 256), 256, 0, s>>>(M, P, sp, runs, dt, k0, rec, lambda, grad, cnt);
}
template void launch_hit_points<double>(hipStream_t, const DevMesh&, const DevParticles<double>&, const SettleRec*,
                                        double*, int32_t*, int32_t*);
template void launch_hit_points<float>(hipStream_t, const DevMesh&, const DevParticles<float>&, const SettleRec*,
                                       double*, int32_t*, int32_t*);
template void launch_adjoint_push<double>(hipStream_t, const DevMesh&, const DevParticles<double>&, const SpeciesK*,
                                          const RunK*, double, int32_t, const SettleRec*, const double*, double*,
                                          unsigned long long*);
template void launch_adjoint_push<float>(hipStream_t, const DevMesh&, const DevParticles<float>&, const SpeciesK*,
                                         const RunK*, double, int32_t, const SettleRec*, const double*, double*,
                                         unsigned long long*);

template <typename R>
void launch_node_density(hipStream_t s, const DevParticles<R>& P, const SpeciesK* sp, int32_t mode,
                         const double* node_q, const double* node_vol, const int32_t* adj_start,
                         const int32_t* adj_list, const double* tet_vol, double* tet_q, int32_t* tet_n,
                         int64_t n_nodes, double* rho, int32_t* written)
{
    if (mode == 0) {
        node_density_bary_kernel<<<blocks_for(n_nodes, 256), 256, 0, s>>>(node_q, node_vol, n_nodes, rho, written);
        return;
    }
    if (P.n > 0) tet_occupancy_kernel<R><<<blocks_for(P.n, 256), 256, 0, s>>>(P, sp, tet_q, tet_n);
    node_density_kernel<<<blocks_for(n_nodes, 256), 256, 0, s>>>(adj_start, adj_list, tet_vol, tet_q, tet_n, n_nodes,
                                                                  rho, written);
}
template void launch_node_density<double>(hipStream_t, const DevParticles<double>&, const SpeciesK*, int32_t,
                                          const double*, const double*, const int32_t*, const int32_t*,
                                          const double*, double*, int32_t*, int64_t, double*, int32_t*);
template void launch_node_density<float>(hipStream_t, const DevParticles<float>&, const SpeciesK*, int32_t,
                                         const double*, const double*, const int32_t*, const int32_t*,
                                         const double*, double*, int32_t*, int64_t, double*, int32_t*);

void launch_field(hipStream_t s, const DevMesh& M, const double* phi, double* E, float* E32)
{
    field_kernel<<<blocks_for(M.n_tets, 256), 256, 0, s>>>(M, phi, E, E32);
}

void launch_e32(hipStream_t s, const double* E, float* E32, int64_t nt)
{
    e32_kernel<<<blocks_for(nt, 256), 256, 0, s>>>(E, E32, nt);
}

void launch_orient_batch(hipStream_t s, const double* pts, int64_t n, int32_t* out)
{
    if (n > 0) orient_batch_kernel<<<blocks_for(n, 64), 64, 0, s>>>(pts, n, out);
}

void launch_math_batch(hipStream_t s, int which, const double* x, const double* y, int64_t n, double* out,
                       double* out2)
{
    if (n > 0) math_batch_kernel<<<blocks_for(n, 256), 256, 0, s>>>(which, x, y, n, out, out2);
}

void launch_abs_positions(hipStream_t s, const DevMesh& M, const float* lx, const float* ly, const float* lz,
                          const int32_t* tet, int64_t n, double* x, double* y, double* z)
{
    if (n > 0) abs_positions_kernel<<<blocks_for(n, 256), 256, 0, s>>>(M, lx, ly, lz, tet, n, x, y, z);
}

template void launch_step<double>(hipStream_t, const DevMesh&, const DevParticles<double>&, const StepParams&,
                                  const SpeciesK*, double*, WorkItem*, unsigned long long*, unsigned long long*,
                                  const StepLaunch&, hipEvent_t);
template void launch_step<float>(hipStream_t, const DevMesh&, const DevParticles<float>&, const StepParams&,
                                 const SpeciesK*, double*, WorkItem*, unsigned long long*, unsigned long long*,
                                 const StepLaunch&, hipEvent_t);
template void launch_gather<double>(hipStream_t, const DevMesh&, const DevParticles<double>&, const SpeciesK*,
                                    const RunK*, double*);
template void launch_gather<float>(hipStream_t, const DevMesh&, const DevParticles<float>&, const SpeciesK*,
                                   const RunK*, double*);

}  // namespace sg
