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

__device__ Resolved resolve_exhaustive(const DevMesh& M, const double* x0, const double* x1)
{
    Resolved r{-1, -1, 0, 1, 0};
    r.hit = first_hit_grid(M, x0, x1, &r.err);
    if (r.hit < 0) {
        int strict = 0;
        r.tet = locate_grid(M, x1, &strict, &r.err);
        r.strict = strict;
    }
    return r;
}

// Straight walk from t0 (x0 strictly inside) along x0->x1 through exit faces.
// Generic crossings are decided with exact signs; any zero sign (the segment
// meets an edge, a vertex or ends on a face) falls back to the grid search,
// which applies the reference rules directly.
__device__ Resolved resolve(const DevMesh& M, const double* x0, const double* x1, int32_t t0, bool strict0)
{
    if (t0 < 0 || !strict0) return resolve_exhaustive(M, x0, x1);
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
            if (s[f] == 2) return resolve_exhaustive(M, x0, x1);
            neg += s[f] < 0;
            zero += s[f] == 0;
        }
        if (neg == 0) {
            if (zero == 0) return Resolved{-1, t, 1, 0, 0};
            return resolve_exhaustive(M, x0, x1);  // x1 on a face/edge/vertex of t
        }
        int exit_f = -1, nexit = 0;
        for (int f = 0; f < 4; ++f) {
            if (f == entry || s[f] >= 0) continue;
            const double* a = P[kFace[f][0]];
            const double* b = P[kFace[f][1]];
            const double* c = P[kFace[f][2]];
            int o1 = orient3d(x0, x1, a, b), o2 = orient3d(x0, x1, b, c), o3 = orient3d(x0, x1, c, a);
            if (o1 == 0 || o2 == 0 || o3 == 0 || o1 == 2 || o2 == 2 || o3 == 2)
                return resolve_exhaustive(M, x0, x1);
            if (o1 == o2 && o2 == o3) {
                exit_f = f;
                ++nexit;
            }
        }
        if (nexit != 1) return resolve_exhaustive(M, x0, x1);
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
    return resolve_exhaustive(M, x0, x1);
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

__device__ inline void scatter4(double* __restrict__ node_charge, const int4& nv, double q, const double lam[4])
{
    unsafeAtomicAdd(&node_charge[nv.x], q * lam[0]);
    unsafeAtomicAdd(&node_charge[nv.y], q * lam[1]);
    unsafeAtomicAdd(&node_charge[nv.z], q * lam[2]);
    unsafeAtomicAdd(&node_charge[nv.w], q * lam[3]);
}

// Block-local reduction of the charge scatter: contributions to the same node
// from the block's particles are summed in an LDS hash table (open addressing,
// linear probing) and flushed with one FP64 atomic per distinct node.  With
// particles sorted by tet a 256-particle block touches ~100-150 nodes instead
// of issuing 1,024 global atomics.
constexpr int kHashBits = 11;
constexpr int kHashSlots = 1 << kHashBits;

__device__ inline void lds_charge_add(int32_t* hk, double* hv, double* __restrict__ node_charge, int32_t node,
                                      double val)
{
    uint32_t h = ((uint32_t)node * 2654435761u) >> (32 - kHashBits);
    for (int probe = 0; probe < 32; ++probe) {
        int32_t prev = atomicCAS(&hk[h], -1, node);
        if (prev == -1 || prev == node) {
            atomicAdd(&hv[h], val);
            return;
        }
        h = (h + 1) & (kHashSlots - 1);
    }
    unsafeAtomicAdd(&node_charge[node], val);  // crowded table: straight to memory
}

__device__ inline void lds_scatter4(int32_t* hk, double* hv, double* __restrict__ node_charge, const int4& nv,
                                    double q, const double lam[4])
{
    lds_charge_add(hk, hv, node_charge, nv.x, q * lam[0]);
    lds_charge_add(hk, hv, node_charge, nv.y, q * lam[1]);
    lds_charge_add(hk, hv, node_charge, nv.z, q * lam[2]);
    lds_charge_add(hk, hv, node_charge, nv.w, q * lam[3]);
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

template <typename R>
__global__ void __launch_bounds__(256) step_main(DevMesh M, DevParticles<R> P, StepParams S,
                                                 const SpeciesK* __restrict__ sp, double* __restrict__ node_charge,
                                                 WorkItem* __restrict__ wl, unsigned long long* __restrict__ cnt,
                                                 unsigned long long* __restrict__ tri_count)
{
    __shared__ unsigned int blk_processed;
    __shared__ int32_t hk[kHashSlots];
    __shared__ double hv[kHashSlots];
    {
        // clear the next step's charge buffer and work-list counter
        int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
        int64_t gs = (int64_t)gridDim.x * blockDim.x;
        if (S.zero_charge)
            for (int64_t j = g; j < M.n_nodes; j += gs) S.zero_charge[j] = 0.0;
        if (g == 0) cnt[kCntWork + 2 * (1 - S.work_slot)] = 0ull;
    }
    const int64_t n_active = *S.n_active;
    if ((int64_t)blockIdx.x * blockDim.x >= n_active) return;  // whole block past the live particles
    if (threadIdx.x == 0) blk_processed = 0u;
    if (S.charge) {
        for (int h = threadIdx.x; h < kHashSlots; h += blockDim.x) {
            hk[h] = -1;
            hv[h] = 0.0;
        }
        __syncthreads();
    }
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool proc = false;
    bool deferred = false;
    double xd = 0.0, yd = 0.0, zd = 0.0;
    int32_t t0 = -1;
    if (i < n_active) {
        int32_t f = P.flags[i];
        if (!(f & kSettled)) {
            const SpeciesK& k = sp[(f >> kSpeciesShift) & 0xff];
            R x = P.x[i], y = P.y[i], z = P.z[i];
            R vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
            t0 = P.tet[i];
            int32_t t1 = t0;  // owner after the step
            // 2. |v| < 10 m/s and 3. outside the expanded bbox freeze the particle
            //    for this step (ParticleDynamicsProcessor.cpp:170-176)
            R vmod = m_sqrt<R>(vx * vx + vy * vy + vz * vz);
            bool frozen = vmod < R(10);
            xd = (double)x;
            yd = (double)y;
            zd = (double)z;
            frozen = frozen || !(xd >= S.lo[0] && xd <= S.hi[0] && yd >= S.lo[1] && yd <= S.hi[1] &&
                                 zd >= S.lo[2] && zd <= S.hi[2]);
            bool have_lam = false;
            bool settled_now = false;
            double lam[4];
            if (!frozen) {
                proc = true;
                R dt = (R)S.dt;
                // 4. push with the owning tet's field (ParticleDynamicsProcessor.cpp:179-180)
                if (S.em_push && t0 >= 0) {
                    const double* E = M.E + 3 * (int64_t)t0;
                    push<R>((R)E[0], (R)E[1], (R)E[2], (R)k.charge, (R)k.mass, dt, vx, vy, vz);
                }
                // 6. move (Particle.cpp:157-165)
                R x1 = x + vx * dt, y1 = y + vy * dt, z1 = z + vz * dt;
                if (!(x1 == x && y1 == y && z1 == z)) {
                    // 7. gas collision (ParticlePhysicsUpdater.cpp:20-24)
                    if (S.model != 0) {
                        double u1, u1b;
                        variates(S.seed, S.run, P.pidx[i], S.step, 0u, u1, u1b);
                        CollideK<R> ck{(R)k.sigma_hs, (R)k.vhs_c, (R)k.two_kt, (R)k.m_r, (R)k.vhs_exp,
                                       (R)k.mass_cp, (R)k.mass_ct, (R)k.inv_alpha};
                        R vm = m_sqrt<R>(vx * vx + vy * vy + vz * vz);
                        R pc = collision_probability<R>(S.model, ck, (R)S.conc, dt, vm);
                        if ((R)u1 < pc) {
                            double u2, u3;
                            variates(S.seed, S.run, P.pidx[i], S.step, 1u, u2, u3);
                            scatter_velocity<R>(S.model, ck, (R)u2, (R)u3, vx, vy, vz);
                        }
                    }
                    // 8. relocation and surface hit on the segment x -> x1.
                    //    Certified cases: x1 strictly inside the owner (no face met), or
                    //    x1 strictly outside exactly one face f and strictly inside the
                    //    other three, so the segment leaves through the open face f.
                    //    Everything else goes to the exact walk of step_resolve.
                    if (t0 >= 0 && (f & kStrict)) {
                        const double tau = M.tau;
                        bary64(M.geo + 12 * (int64_t)t0, (double)x1, (double)y1, (double)z1, lam);
                        int npos = 0, nneg = 0, fneg = 0;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            npos += lam[q] > tau;
                            if (lam[q] < -tau) {
                                ++nneg;
                                fneg = q;
                            }
                        }
                        if (npos == 4) {
                            have_lam = true;
                        } else if (npos == 3 && nneg == 1) {
                            int32_t code = nth(M.tnbr[t0], fneg);
                            if (code <= -2) {
                                settled_now = true;  // crosses the surface triangle -2-code
                                int32_t tri = -2 - code;
                                P.flags[i] = (f | kSettled) & ~kStrict;
                                P.hit[i] = tri;
                                atomicAdd(&tri_count[tri], 1ull);
                                atomicAdd(&cnt[kCntSettled], 1ull);
                            } else if (code == -1) {
                                t1 = -1;  // left through a face with no triangle
                                P.tet[i] = -1;
                                P.flags[i] = f & ~kStrict;
                            } else {
                                bary64(M.geo + 12 * (int64_t)code, (double)x1, (double)y1, (double)z1, lam);
                                double mn = fmin(fmin(lam[0], lam[1]), fmin(lam[2], lam[3]));
                                if (mn > tau) {
                                    t1 = code;
                                    have_lam = true;
                                    P.tet[i] = code;
                                } else {
                                    deferred = true;
                                }
                            }
                        } else {
                            deferred = true;
                        }
                    } else {
                        deferred = true;
                    }
                    P.x[i] = x1;
                    P.y[i] = y1;
                    P.z[i] = z1;
                    x = x1;
                    y = y1;
                    z = z1;
                }
                P.vx[i] = vx;
                P.vy[i] = vy;
                P.vz[i] = vz;
            }
            // node charge of the end-of-step position (the next gather)
            if (S.charge && !deferred && !settled_now && t1 >= 0 && k.charge != 0.0) {
                if (!have_lam) bary64(M.geo + 12 * (int64_t)t1, (double)x, (double)y, (double)z, lam);
                lds_scatter4(hk, hv, node_charge, M.tnode[t1], k.charge, lam);
            }
        }
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
            it.x0 = xd;
            it.y0 = yd;
            it.z0 = zd;
            it.i = (int32_t)i;
            it.t0 = t0;
            wl[base + (unsigned)__popcll(dm & ((1ull << lane) - 1ull))] = it;
        }
    }
    // processed count: LDS per block, one atomic per block spread over slots
    __syncthreads();
    unsigned long long pm = __ballot(proc);
    if ((threadIdx.x & 63) == 0 && pm) atomicAdd(&blk_processed, (unsigned int)__popcll(pm));
    __syncthreads();
    if (threadIdx.x == 0 && blk_processed)
        atomicAdd(&cnt[kCntProcSlot0 + (blockIdx.x & (kProcSlots - 1))], (unsigned long long)blk_processed);
    if (S.charge) {
        for (int h = threadIdx.x; h < kHashSlots; h += blockDim.x) {
            int32_t node = hk[h];
            if (node >= 0) unsafeAtomicAdd(&node_charge[node], hv[h]);
        }
    }
}

template <typename R>
__global__ void __launch_bounds__(256) step_resolve(DevMesh M, DevParticles<R> P, StepParams S,
                                                    const SpeciesK* __restrict__ sp,
                                                    double* __restrict__ node_charge, const WorkItem* __restrict__ wl,
                                                    unsigned long long* __restrict__ cnt,
                                                    unsigned long long* __restrict__ tri_count)
{
    const uint64_t nw = cnt[kCntWork + 2 * S.work_slot];
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nw;
         w += (uint64_t)gridDim.x * blockDim.x) {
        WorkItem it = wl[w];
        int64_t i = it.i;
        double x0[3] = {it.x0, it.y0, it.z0};
        double x1[3] = {(double)P.x[i], (double)P.y[i], (double)P.z[i]};
        int32_t f = P.flags[i];
        Resolved r = resolve(M, x0, x1, it.t0, (f & kStrict) != 0);
        if (r.err) atomicOr(&cnt[kCntError], 1ull);
        if (r.exhaustive) atomicAdd(&cnt[kCntExhaustive], 1ull);
        if (r.hit >= 0) {
            // ParticleSettler::settle (ParticleSettler.cpp:7-26)
            P.flags[i] = (f | kSettled) & ~kStrict;
            P.hit[i] = r.hit;
            atomicAdd(&tri_count[r.hit], 1ull);
            atomicAdd(&cnt[kCntSettled], 1ull);
        } else {
            P.tet[i] = r.tet;
            P.flags[i] = r.strict ? (f | kStrict) : (f & ~kStrict);
            const SpeciesK& k = sp[(f >> kSpeciesShift) & 0xff];
            if (S.charge && r.tet >= 0 && k.charge != 0.0) {
                double lam[4];
                bary64(M.geo + 12 * (int64_t)r.tet, x1[0], x1[1], x1[2], lam);
                scatter4(node_charge, M.tnode[r.tet], k.charge, lam);
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&cnt[kCntDeferred], (unsigned long long)nw);
}

template <typename R>
__global__ void locate_kernel(DevMesh M, const R* __restrict__ x, const R* __restrict__ y, const R* __restrict__ z,
                              int64_t n, int32_t* __restrict__ tet, int32_t* __restrict__ flags,
                              unsigned long long* __restrict__ cnt)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p[3] = {(double)x[i], (double)y[i], (double)z[i]};
    int strict = 0, err = 0;
    int32_t t = locate_grid(M, p, &strict, &err);
    if (err && cnt) atomicOr(&cnt[kCntError], 1ull);
    tet[i] = t;
    if (flags) flags[i] = strict ? (flags[i] | kStrict) : (flags[i] & ~kStrict);
}

template <typename R>
__global__ void gather_kernel(DevMesh M, DevParticles<R> P, const SpeciesK* __restrict__ sp,
                              double* __restrict__ node_charge)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    int32_t f = P.flags[i];
    int32_t t = P.tet[i];
    if ((f & kSettled) || t < 0) return;
    double q = sp[(f >> kSpeciesShift) & 0xff].charge;
    if (q == 0.0) return;
    double lam[4];
    bary64(M.geo + 12 * (int64_t)t, (double)P.x[i], (double)P.y[i], (double)P.z[i], lam);
    scatter4(node_charge, M.tnode[t], q, lam);
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
                 unsigned long long* tri_count, hipEvent_t ev_mid)
{
    if (P.n > 0) step_main<R><<<blocks_for(P.n, 256), 256, 0, s>>>(M, P, S, sp, node_charge, wl, cnt, tri_count);
    if (ev_mid) (void)hipEventRecord(ev_mid, s);
    unsigned nb = std::max(1u, std::min(blocks_for(P.n, 256), 64u));
    step_resolve<R><<<nb, 256, 0, s>>>(M, P, S, sp, node_charge, wl, cnt, tri_count);
}

template <typename R>
void launch_locate(hipStream_t s, const DevMesh& M, const R* x, const R* y, const R* z, int64_t n, int32_t* tet,
                   int32_t* flags, unsigned long long* cnt)
{
    if (n > 0) locate_kernel<R><<<blocks_for(n, 128), 128, 0, s>>>(M, x, y, z, n, tet, flags, cnt);
}

template <typename R>
void launch_gather(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const SpeciesK* sp, double* node_charge)
{
    if (P.n > 0) gather_kernel<R><<<blocks_for(P.n, 256), 256, 0, s>>>(M, P, sp, node_charge);
}

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

template void launch_step<double>(hipStream_t, const DevMesh&, const DevParticles<double>&, const StepParams&,
                                  const SpeciesK*, double*, WorkItem*, unsigned long long*, unsigned long long*,
                                  hipEvent_t);
template void launch_locate<double>(hipStream_t, const DevMesh&, const double*, const double*, const double*, int64_t,
                                    int32_t*, int32_t*, unsigned long long*);
template void launch_gather<double>(hipStream_t, const DevMesh&, const DevParticles<double>&, const SpeciesK*,
                                    double*);

}  // namespace sg
