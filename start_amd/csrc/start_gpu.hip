// C ABI of libstart_gpu (declared in include/start_gpu.h): device state of one
// simulation context (or one batch of independent runs on a shared mesh) and
// the host side of the step loop.
#include "start_gpu.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "kernels.h"
#include "mesh_prep.h"
#include "sources.h"

namespace sg {
void launch_zero_charge(hipStream_t s, double* a, double* b, int64_t n);
template <typename R>
void launch_step(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const StepParams& S, const SpeciesK* sp,
                 double* node_charge, WorkItem* wl, unsigned long long* cnt, unsigned long long* tri_count,
                 const StepLaunch& L, hipEvent_t ev_mid);
void launch_locate(hipStream_t s, const DevMesh& M, const double* x, const double* y, const double* z, int64_t n,
                   int32_t* tet, int32_t* flags, float* lx, float* ly, float* lz, unsigned long long* cnt);
template <typename R>
void launch_gather(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const SpeciesK* sp, const RunK* runs,
                   double* node_charge);
template <typename R>
void launch_node_density(hipStream_t s, const DevParticles<R>& P, const SpeciesK* sp, int32_t mode,
                         const double* node_q, const double* node_vol, const int32_t* adj_start,
                         const int32_t* adj_list, const double* tet_vol, double* tet_q, int32_t* tet_n,
                         int64_t n_nodes, double* rho, int32_t* written);
template <typename R>
void launch_hit_points(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const SettleRec* rec, double* xh,
                       int32_t* tri, int32_t* step);
template <typename R>
void launch_adjoint_push(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const SpeciesK* sp,
                         const RunK* runs, double dt, int32_t k0, const SettleRec* rec, const double* lambda,
                         double* grad, unsigned long long* cnt);
template <typename R>
void launch_track_eligible(hipStream_t s, const DevMesh& M, const DevParticles<R>& P, const StepParams& S,
                           uint8_t* out);
void launch_field(hipStream_t s, const DevMesh& M, const double* phi, double* E, float* E32);
void launch_e32(hipStream_t s, const double* E, float* E32, int64_t nt);
void launch_push_table(hipStream_t s, const double* E, const float* E32, int64_t nt, const double* q,
                       const double* m, int32_t n_slots, double dt, double* H, float* H32);
void launch_orient_batch(hipStream_t s, const double* pts, int64_t n, int32_t* out);
void launch_math_batch(hipStream_t s, int which, const double* x, const double* y, int64_t n, double* out,
                       double* out2);
void launch_abs_positions(hipStream_t s, const DevMesh& M, const float* lx, const float* ly, const float* lz,
                          const int32_t* tet, int64_t n, double* x, double* y, double* z);
size_t sort_temp_bytes(int64_t n, int nbits);
template <typename R>
hipError_t sort_particles(hipStream_t s, const DevParticles<R>& P, R* x, R* y, R* z, double* vx, double* vy, double* vz,
                          int32_t* tet, int32_t* flags, int32_t* hit, uint32_t* pidx, int32_t* keys,
                          int32_t* keys_out, int32_t* idx, int32_t* perm, void* temp, size_t temp_bytes,
                          int32_t n_tets, int nbits, int64_t* n_active);
}  // namespace sg

using namespace sg;

// Staging of start_locate_noalloc: the x, y and z columns (FP64) and the
// owners (int32) of at most START_NOALLOC_LOCATE_MAX points, in one file-scope
// device array.  Every GPU the library runs on holds its own copy until the
// process ends; every call on a GPU shares that GPU's copy.
constexpr size_t kNoAllocBytes = (size_t)START_NOALLOC_LOCATE_MAX * (3 * sizeof(double) + sizeof(int32_t));
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::string g_create_error;

struct TimedStep {
    hipEvent_t a0, a1, b1;
};

// one set of particle arrays; positions are double (FP64) or float in the
// owner's frame (mixed), velocities double in both
struct PBuf {
    void *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
    int32_t *tet = nullptr, *flags = nullptr, *hit = nullptr;
    uint32_t* pidx = nullptr;
};

// A captured chunk of step launches: valid for the live buffer set, the
// charge/work-list parities and the launch configuration it was captured with.
struct GraphEntry {
    const void* ax;
    int cb, wslot;
    int32_t chunk;
    start_launch_cfg cfg;
    hipGraphExec_t exec;
    int cb_after, wslot_after, latest_after;
};

bool same_cfg(const start_launch_cfg& a, const start_launch_cfg& b)
{
    return a.steps_per_launch == b.steps_per_launch && a.block == b.block && a.persistent == b.persistent &&
           a.sort_interval == b.sort_interval && a.walk == b.walk;
}

// Device address of `scratch` per device.  In host code the array's name is a
// host-side handle, not a device address, so kernels and copies take these.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The address on the current device, resolved once per device (called by
// start_create with the context's device current).
hipError_t scratch_address(void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

struct start_ctx {
    std::string err;
    hipStream_t stream = nullptr;
    HostMesh hm;
    start_params prm{};
    start_launch_cfg cfg{};
    int n_cu = 256;
    std::string arch = "unknown";  // the device's gcnArchName, part of the autotuner's cache key
    void* scratch = nullptr;  // device address of the no-allocation staging array on this context's device
    std::vector<start_species_desc> species_desc;
    bool mixed = false;
    size_t es = sizeof(double);  // bytes of one position component
    std::vector<SpeciesK> species;
    // device mesh
    double* d_nodes = nullptr;
    double* d_geo = nullptr;
    float* d_geo32 = nullptr;
    double* d_E = nullptr;
    float* d_E32 = nullptr;
    double* d_E1 = nullptr;  // field basis 1 (allocated on first use)
    float* d_E1_32 = nullptr;
    // node potential hand-off: device copy, pinned staging buffer, and the
    // event after which the staging buffer may be rewritten
    double* d_phi = nullptr;
    double* h_phi = nullptr;
    hipEvent_t phi_copied = nullptr;
    bool phi_pending = false;
    // push increments per charged species slot and tet (single simulation),
    // refreshed before the next step whenever the field or the time step changed
    double* d_H = nullptr;
    float* d_H32 = nullptr;
    int32_t n_push = 0;
    bool push_dirty = true;
    std::vector<uint8_t> present;  // species of the loaded population
    int4* d_tnode = nullptr;
    int4* d_tnbr = nullptr;
    int32_t* d_tris = nullptr;
    uint8_t* d_tri_deg = nullptr;
    int32_t *d_tgs = nullptr, *d_tgl = nullptr, *d_sgs = nullptr, *d_sgl = nullptr;
    SpeciesK* d_species = nullptr;
    float tau32 = 0.0f;
    // runs: one for a single simulation, n_runs for a batch
    std::vector<RunK> runs;
    RunK* d_runs = nullptr;
    unsigned long long* d_tri_count = nullptr;  // n_runs * n_tris
    double* d_node_charge = nullptr;            // two buffers of n_runs * n_nodes
    int cb = 0;                                 // buffer the next launch scatters into
    int latest = 0;                             // buffer holding the most recent node charge
    int wslot = 0;                              // work-list / step / tile counter slot of the next launch
    // node charge density tables (built on first use)
    double *d_tet_vol = nullptr, *d_node_vol = nullptr, *d_tet_q = nullptr;
    int32_t *d_adj_start = nullptr, *d_adj_list = nullptr, *d_tet_n = nullptr;
    // node density output: device buffers and pinned host staging, kept
    // between calls (the step loop asks for it every step)
    double* d_rho = nullptr;
    int32_t* d_written = nullptr;
    double* h_rho = nullptr;
    int32_t* h_written = nullptr;
    unsigned long long* h_cnt = nullptr;  // page-locked copy of the counters (start_read_stats)
    uint32_t* h_settle = nullptr;         // page-locked staging of settled load indices
    int64_t h_settle_cap = 0;
    // particles: n live rows in buffers of cap rows (appends grow cap geometrically)
    int64_t n = 0;
    int64_t cap = 0;
    PBuf A, B;  // live arrays and the sort target
    std::vector<uint64_t> ids;
    WorkItem* d_wl = nullptr;
    int32_t *d_keys = nullptr, *d_keys_out = nullptr, *d_idx = nullptr, *d_perm = nullptr;
    void* d_sort_temp = nullptr;
    size_t sort_temp = 0;
    int sort_bits = 1;
    int64_t* d_nactive = nullptr;
    unsigned long long* d_cnt = nullptr;
    uint64_t step = 0;
    uint64_t last_sort = 0;
    // load indices in settle order (cap entries: a particle settles once)
    uint32_t* d_settle_log = nullptr;
    // settle records for the adjoint (by load index), when recording
    SettleRec* d_settle = nullptr;
    bool record = false;
    // movement records: slot of each load index, points per slot, the points
    // (trk_max per slot), and the load index of each slot
    int32_t* d_trk_slot = nullptr;
    int32_t* d_trk_count = nullptr;
    double* d_trk_xyz = nullptr;
    int32_t trk_max = 0;
    int64_t trk_cap = 0;  // particles the caller asked to track (appended ones may join)
    std::vector<int64_t> trk_rows;
    // captured step chunks
    std::vector<GraphEntry> graphs;
    // timing
    bool timing = false;
    std::vector<TimedStep> events;
    size_t ev_used = 0;

    int64_t n_runs() const { return (int64_t)runs.size(); }
    int64_t charge_len() const { return n_runs() * hm.n_nodes; }

    DevMesh mesh_view() const
    {
        DevMesh M{};
        M.nodes = d_nodes;
        M.geo = d_geo;
        M.E = d_E;
        M.E32 = d_E32;
        M.E1 = d_E1;
        M.E1_32 = d_E1_32;
        M.geo32 = d_geo32;
        M.H = d_H;
        M.H32 = d_H32;
        M.tnode = d_tnode;
        M.tnbr = d_tnbr;
        M.tris = d_tris;
        M.tri_degenerate = d_tri_deg;
        M.tgrid_start = d_tgs;
        M.tgrid_list = d_tgl;
        M.sgrid_start = d_sgs;
        M.sgrid_list = d_sgl;
        M.grid = hm.grid;
        for (int k = 0; k < 3; ++k) {
            M.node_lo[k] = hm.node_lo[k];
            M.node_hi[k] = hm.node_hi[k];
        }
        M.n_nodes = hm.n_nodes;
        M.n_tets = hm.n_tets;
        M.n_tris = hm.n_tris;
        M.tau = hm.tau;
        M.tau32 = tau32;
        return M;
    }
    template <typename R>
    DevParticles<R> view() const
    {
        return DevParticles<R>{(R*)A.x, (R*)A.y, (R*)A.z, (double*)A.vx, (double*)A.vy, (double*)A.vz,
                               A.tet,   A.flags, A.hit,  A.pidx,  n};
    }
    StepLaunch launch_shape() const
    {
        StepLaunch L{};
        L.walk = cfg.walk ? 1 : 0;
        L.block = cfg.block == 128 && !L.walk ? 128 : 256;
        L.persistent = cfg.persistent ? 1 : 0;
        L.grid = n_cu * (1024 / L.block);  // 4 waves per SIMD x 4 SIMDs per CU
        // per-lane run parameters unless the field is E_0 as given
        L.batch = (runs.size() > 1 || d_E1 || runs[0].s0 != 1.0 || runs[0].s1 != 0.0) ? 1 : 0;
        return L;
    }
};

namespace {

int fail(start_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPCHECK(ctx, call)                                                                       \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail((ctx), START_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device allocation that leaves no error behind when it fails: HIP keeps a
// failed hipMalloc's error as the thread's last error, which the next
// hipGetLastError() after a kernel launch would report as the launch's.
hipError_t dev_malloc(void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
    }
    return e;
}

hipError_t host_pinned_malloc(void** p, size_t bytes)
{
    hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
    }
    return e;
}

template <typename T>
hipError_t dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    if (count > SIZE_MAX / sizeof(T)) return hipErrorOutOfMemory;
    return dev_malloc((void**)p, sizeof(T) * count);
}

hipError_t alloc_pbuf(PBuf& b, int64_t n, size_t es)
{
    hipError_t e = hipSuccess;
    void** f[6] = {&b.x, &b.y, &b.z, &b.vx, &b.vy, &b.vz};
    for (int k = 0; k < 6; ++k)
        if (e == hipSuccess) e = dev_malloc(f[k], (k < 3 ? es : sizeof(double)) * (size_t)std::max<int64_t>(n, 1));
    if (e == hipSuccess) e = dalloc(&b.tet, n);
    if (e == hipSuccess) e = dalloc(&b.flags, n);
    if (e == hipSuccess) e = dalloc(&b.hit, n);
    if (e == hipSuccess) e = dalloc(&b.pidx, n);
    return e;
}

void free_pbuf(PBuf& b)
{
    void* ptrs[] = {b.x, b.y, b.z, b.vx, b.vy, b.vz, b.tet, b.flags, b.hit, b.pidx};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    b = PBuf{};
}

void drop_graphs(start_ctx* c)
{
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

void free_tracking(start_ctx* c)
{
    void* ptrs[] = {c->d_trk_slot, c->d_trk_count, c->d_trk_xyz};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    c->d_trk_slot = c->d_trk_count = nullptr;
    c->d_trk_xyz = nullptr;
    c->trk_max = 0;
    c->trk_cap = 0;
    c->trk_rows.clear();
}

// The sort target, work list and sort scratch (contents never outlive a
// launch or a sort, so a growth reallocates them without copying).
void free_work_buffers(start_ctx* c)
{
    free_pbuf(c->B);
    void* ptrs[] = {c->d_wl, c->d_keys, c->d_keys_out, c->d_idx, c->d_perm, c->d_sort_temp};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    c->d_wl = nullptr;
    c->d_keys = c->d_keys_out = c->d_idx = c->d_perm = nullptr;
    c->d_sort_temp = nullptr;
    c->sort_temp = 0;
}

hipError_t alloc_work_buffers(start_ctx* c, int64_t cap)
{
    hipError_t he = dalloc(&c->d_wl, cap);
    if (he == hipSuccess) he = alloc_pbuf(c->B, cap, c->es);
    if (he == hipSuccess) he = dalloc(&c->d_keys, cap);
    if (he == hipSuccess) he = dalloc(&c->d_keys_out, cap);
    if (he == hipSuccess) he = dalloc(&c->d_idx, cap);
    if (he == hipSuccess) he = dalloc(&c->d_perm, cap);
    if (he == hipSuccess) {
        int bits = 1;
        while (bits < 31 && ((int64_t)1 << bits) <= c->hm.n_tets + 1) ++bits;
        c->sort_bits = bits;
        c->sort_temp = sort_temp_bytes(cap, bits);
        he = dev_malloc(&c->d_sort_temp, std::max<size_t>(c->sort_temp, 16));
    }
    return he;
}

void free_particles(start_ctx* c)
{
    drop_graphs(c);
    free_tracking(c);
    free_pbuf(c->A);
    free_work_buffers(c);
    if (c->d_settle) (void)hipFree(c->d_settle);
    if (c->d_settle_log) (void)hipFree(c->d_settle_log);
    c->d_settle = nullptr;
    c->d_settle_log = nullptr;
    c->n = 0;
    c->cap = 0;
}

// radix sort by owning tet, settled particles last, then swap the buffer sets
template <typename R>
hipError_t sort_typed(start_ctx* c)
{
    PBuf& B = c->B;
    return sort_particles<R>(c->stream, c->view<R>(), (R*)B.x, (R*)B.y, (R*)B.z, (double*)B.vx, (double*)B.vy, (double*)B.vz, B.tet,
                             B.flags, B.hit, B.pidx, c->d_keys, c->d_keys_out, c->d_idx, c->d_perm, c->d_sort_temp,
                             c->sort_temp, (int32_t)c->hm.n_tets, c->sort_bits, c->d_nactive);
}

int sort_now(start_ctx* c)
{
    if (c->n == 0 || !c->B.x) return START_OK;
    hipError_t e = c->mixed ? sort_typed<float>(c) : sort_typed<double>(c);
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("particle sort: ") + hipGetErrorString(e));
    std::swap(c->A, c->B);
    c->last_sort = c->step;
    return START_OK;
}

// Species constants with the reference's expressions (HSModel.cpp:8-12,
// VHSModel.cpp:12-22, VSSModel.cpp:32), evaluated once.
SpeciesK species_constants(const start_species_desc& s, const start_gas_desc& g)
{
    const double KT_reference = 297.0 * 1.380658e-23;  // Constants.hpp:12
    SpeciesK k{};
    double p_mass = s.mass, t_mass = g.mass, d = s.diameter;
    k.mass = p_mass;
    k.charge = s.charge;
    k.sigma_hs = kPi * pow(d, 2);
    k.vhs_c = pow(d, 2) / tgamma(5. / 2. - g.omega);
    k.two_kt = 2 * KT_reference;
    k.m_r = p_mass * t_mass / (p_mass + t_mass);
    k.vhs_exp = g.omega - 1. / 2.;
    k.mass_cp = p_mass / (t_mass + p_mass);
    k.mass_ct = t_mass / (t_mass + p_mass);
    k.inv_alpha = g.alpha != 0.0 ? 1. / g.alpha : 0.0;
    k.push_slot = s.charge == 0.0 ? -2 : -1;
    k.pad = 0;
    return k;
}

// FP32 copy of the barycentric maps (A row-major, then v0) and the FP32
// certification threshold.  For a stored local position l (|l|_inf <= D):
// lambda_j = a_j . l with a_j rounded to FP32 and a 3-term FP32 dot product
// errs by <= (eps32 + gamma_3) |A|_inf D ~ 4 eps32 R (R = |A|_inf D);
// lambda_0 = 1 - ((l1 + l2) + l3) adds the three and gamma_3 (1 + sum|l_j|),
// so |lambda32 - lambda| <= (12 R + 7) eps32; tau32 takes 1.25x that at
// D = 2 diam (the step kernels classify only points with sum |lambda| <= 2,
// which lie within 2 diam of v0), plus 8 rho for the FP64 map's own error
// (rho bounds the row sums of |A E - I|).  A tet whose FP32 bound reaches
// 0.05 makes FP32 certify nothing: every classification then falls back to
// the FP64 map (correct, slower).
float build_geo32(const HostMesh& hm, std::vector<float>& g32)
{
    const double eps32 = 5.9604644775390625e-08;  // 2^-24
    g32.assign((size_t)(12 * hm.n_tets), 0.0f);
    double tau = 0.0;
    for (int64_t t = 0; t < hm.n_tets; ++t) {
        const double* g = &hm.geo[12 * (size_t)t];
        float* o = &g32[12 * (size_t)t];
        double rowsum = 0.0;
        for (int r = 0; r < 3; ++r) {
            double s = 0.0;
            for (int q = 0; q < 3; ++q) {
                o[3 * r + q] = (float)g[3 + 3 * r + q];
                s += fabs(g[3 + 3 * r + q]);
            }
            rowsum = std::max(rowsum, s);
        }
        o[9] = (float)g[0];
        o[10] = (float)g[1];
        o[11] = (float)g[2];
        double diam = 0.0;
        const int32_t* v = &hm.tets[4 * (size_t)t];
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b)
                for (int k = 0; k < 3; ++k)
                    diam = std::max(diam, fabs(hm.nodes[3 * (size_t)v[a] + k] - hm.nodes[3 * (size_t)v[b] + k]));
        tau = std::max(tau, 1.25 * eps32 * (12.0 * rowsum * 2.0 * diam + 7.0) + 8.0 * hm.rho[(size_t)t]);
    }
    return tau < 0.05 ? (float)tau : INFINITY;
}

// Tet volumes ((b-a) . ((c-a) x (d-a)))/6 of the positively oriented tets,
// lumped node volumes (sum over the tets around a node, ascending tet id, / 4)
// and the node -> tets adjacency in ascending tet id.
int ensure_density_tables(start_ctx* c)
{
    if (c->d_tet_vol) return START_OK;
    const HostMesh& hm = c->hm;
    const int64_t T = hm.n_tets, M = hm.n_nodes;
    std::vector<double> tv((size_t)T), nv((size_t)M, 0.0);
    std::vector<int32_t> start((size_t)M + 1, 0), list((size_t)(4 * T));
    for (int64_t t = 0; t < T; ++t) {
        const int32_t* v = &hm.tets[4 * (size_t)t];
        const double *a = &hm.nodes[3 * (size_t)v[0]], *b = &hm.nodes[3 * (size_t)v[1]],
                     *cc = &hm.nodes[3 * (size_t)v[2]], *d = &hm.nodes[3 * (size_t)v[3]];
        double u[3], w[3], s[3];
        for (int k = 0; k < 3; ++k) {
            u[k] = b[k] - a[k];
            w[k] = cc[k] - a[k];
            s[k] = d[k] - a[k];
        }
        double cx = w[1] * s[2] - w[2] * s[1];
        double cy = w[2] * s[0] - w[0] * s[2];
        double cz = w[0] * s[1] - w[1] * s[0];
        double vol = ((u[0] * cx + u[1] * cy) + u[2] * cz) / 6.0;
        tv[(size_t)t] = vol;
        for (int j = 0; j < 4; ++j) {
            nv[(size_t)v[j]] += vol;
            start[(size_t)v[j] + 1] += 1;
        }
    }
    for (int64_t n = 0; n < M; ++n) {
        nv[(size_t)n] = nv[(size_t)n] / 4.0;
        start[(size_t)n + 1] += start[(size_t)n];
    }
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int64_t t = 0; t < T; ++t)
        for (int j = 0; j < 4; ++j) list[(size_t)fill[(size_t)hm.tets[4 * (size_t)t + j]]++] = (int32_t)t;
    hipError_t e = dalloc(&c->d_tet_vol, T);
    if (e == hipSuccess) e = dalloc(&c->d_node_vol, M);
    if (e == hipSuccess) e = dalloc(&c->d_tet_q, T);
    if (e == hipSuccess) e = dalloc(&c->d_tet_n, T);
    if (e == hipSuccess) e = dalloc(&c->d_adj_start, M + 1);
    if (e == hipSuccess) e = dalloc(&c->d_adj_list, 4 * T);
    if (e == hipSuccess) e = hipMemcpy(c->d_tet_vol, tv.data(), sizeof(double) * T, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_node_vol, nv.data(), sizeof(double) * M, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_adj_start, start.data(), sizeof(int32_t) * (M + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_adj_list, list.data(), sizeof(int32_t) * 4 * T, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        void* ptrs[] = {c->d_tet_vol, c->d_node_vol, c->d_tet_q, c->d_tet_n, c->d_adj_start, c->d_adj_list};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        c->d_tet_vol = c->d_node_vol = c->d_tet_q = nullptr;
        c->d_tet_n = c->d_adj_start = c->d_adj_list = nullptr;
        return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                    std::string("density tables: ") + hipGetErrorString(e));
    }
    return START_OK;
}

int check_kernel_errors(start_ctx* c)
{
    unsigned long long h[kCntTotal];
    HIPCHECK(c, hipMemcpyAsync(h, c->d_cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (h[kCntError] & 1ull) return fail(c, START_ERR_KERNEL, "exact predicate expansion capacity exceeded");
    return START_OK;
}

int upload_runs(start_ctx* c)
{
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (c->d_runs) (void)hipFree(c->d_runs);
    c->d_runs = nullptr;
    HIPCHECK(c, dalloc(&c->d_runs, c->runs.size()));
    HIPCHECK(c, hipMemcpy(c->d_runs, c->runs.data(), sizeof(RunK) * c->runs.size(), hipMemcpyHostToDevice));
    drop_graphs(c);  // captured launches hold the old table pointer
    return START_OK;
}

// Table slots for the charged species of the population (up to
// kMaxPushSlots; others push in the kernel), and the tables they need.
int assign_push_slots(start_ctx* c)
{
    std::vector<int32_t> slot(c->species.size(), -1);
    int32_t n = 0;
    for (size_t s = 0; s < c->species.size(); ++s) {
        if (c->species[s].charge == 0.0) slot[s] = -2;
        else if (s < c->present.size() && c->present[s] && n < kMaxPushSlots) slot[s] = n++;
    }
    bool same = n == c->n_push;
    for (size_t s = 0; s < c->species.size(); ++s) same = same && c->species[s].push_slot == slot[s];
    if (same) return START_OK;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (size_t s = 0; s < c->species.size(); ++s) c->species[s].push_slot = slot[s];
    if (c->d_H) (void)hipFree(c->d_H);
    if (c->d_H32) (void)hipFree(c->d_H32);
    c->d_H = nullptr;
    c->d_H32 = nullptr;
    c->n_push = n;
    const int64_t T = c->hm.n_tets;
    hipError_t e = hipSuccess;
    if (n > 0) e = c->mixed ? dalloc(&c->d_H32, 4 * (int64_t)n * T) : dalloc(&c->d_H, 3 * (int64_t)n * T);
    if (e == hipSuccess)
        e = hipMemcpy(c->d_species, c->species.data(), sizeof(SpeciesK) * c->species.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        for (auto& k : c->species) k.push_slot = k.charge == 0.0 ? -2 : -1;
        c->n_push = 0;
        (void)hipMemcpy(c->d_species, c->species.data(), sizeof(SpeciesK) * c->species.size(), hipMemcpyHostToDevice);
        return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                    std::string("push tables: ") + hipGetErrorString(e));
    }
    c->push_dirty = true;
    drop_graphs(c);  // captured launches hold the table pointers
    return START_OK;
}

// Species of n new particles join the population's set.
int mark_present(start_ctx* c, const int32_t* species, int64_t n)
{
    c->present.resize(c->species.size(), 0);
    for (int64_t i = 0; i < n; ++i) c->present[(size_t)species[i]] = 1;
    return assign_push_slots(c);
}

// The tables of the current field and time step, queued on the library's
// stream before the next step.
int refresh_push_tables(start_ctx* c)
{
    if (!c->push_dirty) return START_OK;
    if (c->n_push > 0) {
        std::vector<double> q, m;
        for (const auto& k : c->species)
            if (k.push_slot >= 0) {
                q.resize((size_t)std::max<int32_t>((int32_t)q.size(), k.push_slot + 1));
                m.resize(q.size());
                q[(size_t)k.push_slot] = k.charge;
                m[(size_t)k.push_slot] = k.mass;
            }
        launch_push_table(c->stream, c->d_E, c->d_E32, c->hm.n_tets, q.data(), m.data(), c->n_push, c->prm.dt, c->d_H,
                          c->d_H32);
        HIPCHECK(c, hipGetLastError());
    }
    c->push_dirty = false;
    return START_OK;
}

// Per-run output buffers: node charge (two sets) and triangle counters.
int alloc_run_outputs(start_ctx* c)
{
    if (c->d_node_charge) (void)hipFree(c->d_node_charge);
    if (c->d_tri_count) (void)hipFree(c->d_tri_count);
    c->d_node_charge = nullptr;
    c->d_tri_count = nullptr;
    const int64_t R = c->n_runs();
    hipError_t e = dalloc(&c->d_node_charge, 2 * R * c->hm.n_nodes);
    if (e == hipSuccess) e = dalloc(&c->d_tri_count, R * std::max<int64_t>(c->hm.n_tris, 1));
    if (e != hipSuccess)
        return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                    std::string("run buffers: ") + hipGetErrorString(e));
    return START_OK;
}

// Fills the step arguments of the current configuration.
StepParams step_params(const start_ctx* c)
{
    StepParams S{};
    S.dt = c->prm.dt;
    for (int k = 0; k < 3; ++k) {
        S.lo[k] = c->hm.bbox_lo[k];
        S.hi[k] = c->hm.bbox_hi[k];
    }
    S.seed = c->prm.seed;
    S.runs = c->d_runs;
    S.model = c->prm.gas.model;
    S.em_push = c->prm.em_push;
    S.charge = c->prm.charge_every_step;
    S.nspecies = (int32_t)c->species.size();
    S.n_active = c->d_nactive;
    S.charge_len = c->charge_len();
    S.settle = c->record ? c->d_settle : nullptr;
    S.settle_log = c->d_settle_log;
    S.trk_slot = c->d_trk_slot;
    S.trk_count = c->d_trk_count;
    S.trk_xyz = c->d_trk_xyz;
    S.trk_max = c->trk_max;
    return S;
}

// Enqueues `chunk` steps as launches of at most steps_per_launch steps and
// advances the host-side parities.  No sort happens inside a chunk.
int enqueue_chunk(start_ctx* c, int32_t chunk)
{
    DevMesh M = c->mesh_view();
    StepParams S = step_params(c);
    const StepLaunch L = c->launch_shape();
    const int64_t CL = c->charge_len();
    const int32_t K = std::max(1, c->cfg.steps_per_launch);
    int32_t done = 0;
    while (done < chunk) {
        int32_t kk = std::min(K, chunk - done);
        double* nc = c->prm.charge_every_step ? c->d_node_charge + (size_t)c->cb * CL : nullptr;
        S.zero_charge = c->prm.charge_every_step ? c->d_node_charge + (size_t)(1 - c->cb) * CL : nullptr;
        S.work_slot = c->wslot;
        S.nsteps = kk;
        TimedStep* ev = nullptr;
        if (c->timing) {
            if (c->ev_used == c->events.size()) {
                TimedStep t{};
                HIPCHECK(c, hipEventCreate(&t.a0));
                HIPCHECK(c, hipEventCreate(&t.a1));
                HIPCHECK(c, hipEventCreate(&t.b1));
                c->events.push_back(t);
            }
            ev = &c->events[c->ev_used++];
        }
        if (ev) HIPCHECK(c, hipEventRecord(ev->a0, c->stream));
        if (c->mixed)
            launch_step<float>(c->stream, M, c->view<float>(), S, c->d_species, nc, c->d_wl, c->d_cnt, c->d_tri_count,
                               L, ev ? ev->a1 : nullptr);
        else
            launch_step<double>(c->stream, M, c->view<double>(), S, c->d_species, nc, c->d_wl, c->d_cnt,
                                c->d_tri_count, L, ev ? ev->a1 : nullptr);
        if (ev) HIPCHECK(c, hipEventRecord(ev->b1, c->stream));
        HIPCHECK(c, hipGetLastError());
        done += kk;
        c->wslot ^= 1;
        if (nc) {
            c->latest = c->cb;
            c->cb ^= 1;
        }
    }
    return START_OK;
}

// A chunk through a captured graph: replays the one captured for this state
// and configuration, capturing it first if there is none.
int run_chunk_graph(start_ctx* c, int32_t chunk)
{
    for (auto& g : c->graphs) {
        if (g.ax == c->A.x && g.cb == c->cb && g.wslot == c->wslot && g.chunk == chunk && same_cfg(g.cfg, c->cfg)) {
            HIPCHECK(c, hipGraphLaunch(g.exec, c->stream));
            c->cb = g.cb_after;
            c->wslot = g.wslot_after;
            c->latest = g.latest_after;
            return START_OK;
        }
    }
    if (c->graphs.size() >= 16) drop_graphs(c);
    GraphEntry g{};
    g.ax = c->A.x;
    g.cb = c->cb;
    g.wslot = c->wslot;
    g.chunk = chunk;
    g.cfg = c->cfg;
    HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_chunk(c, chunk);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc != START_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    g.cb_after = c->cb;
    g.wslot_after = c->wslot;
    g.latest_after = c->latest;
    c->graphs.push_back(g);
    HIPCHECK(c, hipGraphLaunch(g.exec, c->stream));  // the capture only recorded the launches
    return START_OK;
}

// Copies of the mutable device state, for the autotuner's trials.
struct Snapshot {
    PBuf P;
    unsigned long long* cnt = nullptr;
    unsigned long long* tri = nullptr;
    double* charge = nullptr;
    int64_t* nactive = nullptr;
    int cb = 0, latest = 0, wslot = 0;
    uint64_t step = 0, last_sort = 0;
};

hipError_t copy_pbuf(const PBuf& dst, const PBuf& src, int64_t n, size_t es, hipStream_t s)
{
    const void* from[10] = {src.x, src.y, src.z, src.vx, src.vy, src.vz, src.tet, src.flags, src.hit, src.pidx};
    void* to[10] = {dst.x, dst.y, dst.z, dst.vx, dst.vy, dst.vz, dst.tet, dst.flags, dst.hit, dst.pidx};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 10 && e == hipSuccess; ++k)
        e = hipMemcpyAsync(to[k], from[k], (k < 3 ? es : k < 6 ? sizeof(double) : 4) * (size_t)n,
                           hipMemcpyDeviceToDevice, s);
    return e;
}

int snapshot_take(start_ctx* c, Snapshot& s)
{
    const int64_t R = c->n_runs();
    hipError_t e = alloc_pbuf(s.P, c->n, c->es);
    if (e == hipSuccess) e = dalloc(&s.cnt, kCntTotal);
    if (e == hipSuccess) e = dalloc(&s.tri, R * std::max<int64_t>(c->hm.n_tris, 1));
    if (e == hipSuccess) e = dalloc(&s.charge, 2 * R * c->hm.n_nodes);
    if (e == hipSuccess) e = dalloc(&s.nactive, 1);
    if (e == hipSuccess) e = copy_pbuf(s.P, c->A, c->n, c->es, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.cnt, c->d_cnt, sizeof(unsigned long long) * kCntTotal, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.tri, c->d_tri_count, sizeof(unsigned long long) * R * std::max<int64_t>(c->hm.n_tris, 1),
                           hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.charge, c->d_node_charge, sizeof(double) * 2 * R * c->hm.n_nodes,
                           hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.nactive, c->d_nactive, sizeof(int64_t), hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    s.cb = c->cb;
    s.latest = c->latest;
    s.wslot = c->wslot;
    s.step = c->step;
    s.last_sort = c->last_sort;
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("state snapshot: ") + hipGetErrorString(e));
    return START_OK;
}

int snapshot_restore(start_ctx* c, const Snapshot& s)
{
    const int64_t R = c->n_runs();
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    hipError_t e = copy_pbuf(c->A, s.P, c->n, c->es, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_cnt, s.cnt, sizeof(unsigned long long) * kCntTotal, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_tri_count, s.tri, sizeof(unsigned long long) * R * std::max<int64_t>(c->hm.n_tris, 1),
                           hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_node_charge, s.charge, sizeof(double) * 2 * R * c->hm.n_nodes,
                           hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_nactive, s.nactive, sizeof(int64_t), hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    c->cb = s.cb;
    c->latest = s.latest;
    c->wslot = s.wslot;
    c->step = s.step;
    c->last_sort = s.last_sort;
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("state restore: ") + hipGetErrorString(e));
    return START_OK;
}

void snapshot_free(Snapshot& s)
{
    free_pbuf(s.P);
    void* ptrs[] = {s.cnt, s.tri, s.charge, s.nactive};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    s = Snapshot{};
}

int check_cfg(start_ctx* c, const start_launch_cfg& g)
{
    if (g.steps_per_launch < 1) return fail(c, START_ERR_ARG, "steps_per_launch must be >= 1");
    if (g.block != 128 && g.block != 256) return fail(c, START_ERR_ARG, "block must be 128 or 256");
    if (g.sort_interval < 0) return fail(c, START_ERR_ARG, "sort_interval must be >= 0");
    if (g.walk != 0 && g.walk != 1) return fail(c, START_ERR_ARG, "walk must be 0 or 1");
    return START_OK;
}

// Allocates a population of n particles in runs rk (both buffer sets, work
// list, sort scratch), clears the counters and outputs and resets the step
// state; the arrays are filled by the caller.
int begin_population(start_ctx* c, const std::vector<RunK>& rk, int64_t n, const uint64_t* ids)
{
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    free_particles(c);
    c->runs = rk;
    int rc = upload_runs(c);
    if (rc != START_OK) return rc;
    rc = alloc_run_outputs(c);
    if (rc != START_OK) return rc;
    hipError_t he = alloc_pbuf(c->A, n, c->es);
    if (he == hipSuccess) he = alloc_work_buffers(c, n);
    if (he == hipSuccess) he = dalloc(&c->d_settle_log, n);
    if (he != hipSuccess) {
        free_particles(c);
        return fail(c, he == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                    std::string("particle allocation failed: ") + hipGetErrorString(he));
    }
    c->n = n;
    c->cap = n;
    c->ids.assign(ids ? ids : (const uint64_t*)nullptr, ids ? ids + n : (const uint64_t*)nullptr);
    if (!ids) {
        c->ids.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) c->ids[(size_t)i] = (uint64_t)i;
    }
    const int64_t CL = c->charge_len();
    HIPCHECK(c, hipMemsetAsync(c->d_cnt, 0, sizeof(unsigned long long) * kCntTotal, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_tri_count, 0,
                               sizeof(unsigned long long) * c->n_runs() * std::max<int64_t>(c->hm.n_tris, 1),
                               c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_node_charge, 0, sizeof(double) * 2 * CL, c->stream));
    if (c->record) {  // no particle has settled yet: step = -1 in every record
        HIPCHECK(c, dalloc(&c->d_settle, n));
        HIPCHECK(c, hipMemsetAsync(c->d_settle, 0xff, sizeof(SettleRec) * std::max<int64_t>(n, 1), c->stream));
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->cb = 0;
    c->latest = 0;
    c->wslot = 0;
    c->step = 0;
    c->last_sort = 0;
    return START_OK;
}

// Room for `need` particles: a larger buffer set (1.5x the old capacity, at
// least need) with the live rows copied over, and the records kept by load
// index (settle records, movement slots) extended with empty entries.  No
// particle state, counter, node charge or step changes.
int grow_particles(start_ctx* c, int64_t need)
{
    if (need <= c->cap) return START_OK;
    int64_t cap = std::max<int64_t>({need, c->cap + c->cap / 2, 1024});
    cap = std::min<int64_t>(cap, std::max<int64_t>(need, 0x7fffffff));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);  // captured launches hold the old buffers
    PBuf A2;
    SettleRec* s2 = nullptr;
    int32_t* slot2 = nullptr;
    uint32_t* log2 = nullptr;
    hipError_t he = alloc_pbuf(A2, cap, c->es);
    if (he == hipSuccess && c->n > 0) he = copy_pbuf(A2, c->A, c->n, c->es, c->stream);
    if (he == hipSuccess) he = dalloc(&log2, cap);
    if (he == hipSuccess && c->n > 0 && c->d_settle_log)  // entries past the settled count are never read
        he = hipMemcpyAsync(log2, c->d_settle_log, sizeof(uint32_t) * (size_t)c->n, hipMemcpyDeviceToDevice,
                            c->stream);
    if (he == hipSuccess && c->d_settle) {  // step = -1: not settled
        he = dalloc(&s2, cap);
        if (he == hipSuccess) he = hipMemsetAsync(s2, 0xff, sizeof(SettleRec) * (size_t)cap, c->stream);
        if (he == hipSuccess && c->n > 0)
            he = hipMemcpyAsync(s2, c->d_settle, sizeof(SettleRec) * (size_t)c->n, hipMemcpyDeviceToDevice,
                                c->stream);
    }
    if (he == hipSuccess && c->d_trk_slot) {  // -1: not tracked
        he = dalloc(&slot2, cap);
        if (he == hipSuccess) he = hipMemsetAsync(slot2, 0xff, sizeof(int32_t) * (size_t)cap, c->stream);
        if (he == hipSuccess && c->n > 0)
            he = hipMemcpyAsync(slot2, c->d_trk_slot, sizeof(int32_t) * (size_t)c->n, hipMemcpyDeviceToDevice,
                                c->stream);
    }
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) {
        free_pbuf(A2);
        if (s2) (void)hipFree(s2);
        if (slot2) (void)hipFree(slot2);
        if (log2) (void)hipFree(log2);
        return fail(c, he == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                    std::string("particle buffers (growing): ") + hipGetErrorString(he));
    }
    free_pbuf(c->A);
    c->A = A2;
    if (c->d_settle_log) (void)hipFree(c->d_settle_log);
    c->d_settle_log = log2;
    if (s2) {
        (void)hipFree(c->d_settle);
        c->d_settle = s2;
    }
    if (slot2) {
        (void)hipFree(c->d_trk_slot);
        c->d_trk_slot = slot2;
    }
    free_work_buffers(c);
    he = alloc_work_buffers(c, cap);
    if (he != hipSuccess) {
        free_work_buffers(c);  // the state is kept; start_step refuses until buffers exist
        return fail(c, he == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                    std::string("work buffers (growing): ") + hipGetErrorString(he));
    }
    c->cap = cap;
    return START_OK;
}

// Rows [first, first + m) of the live buffers from host arrays: the state
// (positions in the frame of their owner in mixed mode, velocities FP64),
// the flags given, hit = -1, load indices first.., and owners by the locate
// kernel (min-id closed containment).
int upload_rows(start_ctx* c, int64_t first, int64_t m, const double* x, const double* y, const double* z,
                const double* vx, const double* vy, const double* vz, const int32_t* flags)
{
    if (m == 0) return START_OK;
    std::vector<uint32_t> pidx((size_t)m);
    for (int64_t i = 0; i < m; ++i) pidx[(size_t)i] = (uint32_t)(first + i);
    std::vector<int32_t> hit((size_t)m, -1);
    HIPCHECK(c, hipMemcpy(c->A.hit + first, hit.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->A.pidx + first, pidx.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->A.flags + first, flags, sizeof(int32_t) * m, hipMemcpyHostToDevice));
    const size_t b8 = sizeof(double) * (size_t)m;
    DevMesh M = c->mesh_view();
    if (!c->mixed) {
        const double* src[6] = {x, y, z, vx, vy, vz};
        void* dst[6] = {c->A.x, c->A.y, c->A.z, c->A.vx, c->A.vy, c->A.vz};
        for (int k = 0; k < 6; ++k) HIPCHECK(c, hipMemcpy((double*)dst[k] + first, src[k], b8, hipMemcpyHostToDevice));
        launch_locate(c->stream, M, (double*)c->A.x + first, (double*)c->A.y + first, (double*)c->A.z + first, m,
                      c->A.tet + first, c->A.flags + first, nullptr, nullptr, nullptr, c->d_cnt);
    } else {
        // positions go to the frame of their owner
        double *dx = nullptr, *dy = nullptr, *dz = nullptr;
        hipError_t e = dalloc(&dx, m);
        if (e == hipSuccess) e = dalloc(&dy, m);
        if (e == hipSuccess) e = dalloc(&dz, m);
        if (e == hipSuccess) e = hipMemcpy(dx, x, b8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dy, y, b8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dz, z, b8, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            launch_locate(c->stream, M, dx, dy, dz, m, c->A.tet + first, c->A.flags + first, (float*)c->A.x + first,
                          (float*)c->A.y + first, (float*)c->A.z + first, c->d_cnt);
            e = hipStreamSynchronize(c->stream);
        }
        if (dx) (void)hipFree(dx);
        if (dy) (void)hipFree(dy);
        if (dz) (void)hipFree(dz);
        if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("particle upload: ") + hipGetErrorString(e));
        const double* vs[3] = {vx, vy, vz};
        void* vd[3] = {c->A.vx, c->A.vy, c->A.vz};
        for (int k = 0; k < 3; ++k)
            HIPCHECK(c, hipMemcpy((double*)vd[k] + first, vs[k], b8, hipMemcpyHostToDevice));
    }
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return START_OK;
}

// Movement slots, in load order, for the particles with load index >= first
// that the next step moves (not settled, |v| >= 10 m/s, inside the expanded
// bounding box), while fewer than trk_cap are tracked -- the reference's map
// takes a particle while it holds fewer than maxParticles
// (ParticleMovementTracker.cpp:11-12).  first = 0: a whole population (any row
// order); first > 0: rows first..n-1 just appended, which hold load indices
// first..n-1.
int assign_track_slots(start_ctx* c, int64_t first)
{
    const int64_t n = c->n;
    const int64_t m_old = (int64_t)c->trk_rows.size();
    if (c->trk_cap <= m_old || first >= n) return START_OK;
    uint8_t* d_el = nullptr;
    HIPCHECK(c, dalloc(&d_el, n));
    StepParams S = step_params(c);
    if (c->mixed) {
        DevParticles<float> P = c->view<float>();
        if (first > 0)
            P = DevParticles<float>{P.x + first, P.y + first, P.z + first, P.vx + first, P.vy + first, P.vz + first,
                                    P.tet + first, P.flags + first, P.hit + first, P.pidx + first, n - first};
        launch_track_eligible<float>(c->stream, c->mesh_view(), P, S, d_el);
    } else {
        DevParticles<double> P = c->view<double>();
        if (first > 0)
            P = DevParticles<double>{P.x + first, P.y + first, P.z + first, P.vx + first, P.vy + first,
                                     P.vz + first, P.tet + first, P.flags + first, P.hit + first, P.pidx + first,
                                     n - first};
        launch_track_eligible<double>(c->stream, c->mesh_view(), P, S, d_el);
    }
    std::vector<uint8_t> el((size_t)(n - first));
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(el.data(), d_el + first, (size_t)(n - first), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_el);
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("movement tracking: ") + hipGetErrorString(e));
    std::vector<int32_t> slot((size_t)(n - first), -1);
    for (int64_t li = first; li < n && (int64_t)c->trk_rows.size() < c->trk_cap; ++li)
        if (el[(size_t)(li - first)]) {
            slot[(size_t)(li - first)] = (int32_t)c->trk_rows.size();
            c->trk_rows.push_back(li);
        }
    const int64_t m_new = (int64_t)c->trk_rows.size();
    if (m_new == m_old) return START_OK;
    // Every buffer first, the slots last: a step launched after a failure
    // here must not find a slot without its record space.
    if (!c->d_trk_slot) {
        e = dalloc(&c->d_trk_slot, std::max(c->cap, n));
        if (e == hipSuccess) e = hipMemset(c->d_trk_slot, 0xff, sizeof(int32_t) * (size_t)std::max(c->cap, n));
    }
    int32_t* cnt2 = nullptr;
    double* xyz2 = nullptr;
    const size_t per = 3 * (size_t)c->trk_max;
    if (e == hipSuccess) e = dalloc(&cnt2, m_new);
    if (e == hipSuccess) e = dalloc(&xyz2, (size_t)m_new * per);
    if (e == hipSuccess) e = hipMemset(cnt2, 0, sizeof(int32_t) * (size_t)m_new);
    if (e == hipSuccess && m_old > 0) {
        e = hipMemcpy(cnt2, c->d_trk_count, sizeof(int32_t) * (size_t)m_old, hipMemcpyDeviceToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(xyz2, c->d_trk_xyz, sizeof(double) * (size_t)m_old * per, hipMemcpyDeviceToDevice);
    }
    if (e == hipSuccess)
        e = hipMemcpy(c->d_trk_slot + first, slot.data(), sizeof(int32_t) * (size_t)(n - first), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (cnt2) (void)hipFree(cnt2);
        if (xyz2) (void)hipFree(xyz2);
        c->trk_rows.resize((size_t)m_old);
        if (m_old == 0) free_tracking(c);  // nothing was tracked: tracking stays off
        else (void)hipMemset(c->d_trk_slot + first, 0xff, sizeof(int32_t) * (size_t)(n - first));
        return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                    std::string("movement tracking: ") + hipGetErrorString(e));
    }
    if (c->d_trk_count) (void)hipFree(c->d_trk_count);
    if (c->d_trk_xyz) (void)hipFree(c->d_trk_xyz);
    c->d_trk_count = cnt2;
    c->d_trk_xyz = xyz2;
    drop_graphs(c);  // captured launches hold the record pointers
    return START_OK;
}

// hit = -1, load indices 0..n-1 and the live count of a new population.
int init_population_tail(start_ctx* c, int64_t n)
{
    std::vector<uint32_t> pidx((size_t)n);
    for (int64_t i = 0; i < n; ++i) pidx[(size_t)i] = (uint32_t)i;
    std::vector<int32_t> hit((size_t)n, -1);
    HIPCHECK(c, hipMemcpy(c->A.hit, hit.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->A.pidx, pidx.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    int64_t nact = n;
    HIPCHECK(c, hipMemcpy(c->d_nactive, &nact, sizeof(int64_t), hipMemcpyHostToDevice));
    return START_OK;
}

}  // namespace

extern "C" {

double start_gas_concentration(double pressure_pa, double temperature_k)
{
    const double R = 8.314, N_av = 6.02214076e23;  // Constants.hpp:10,13
    return (pressure_pa / (R * temperature_k)) * N_av;
}

const char* start_last_error(const start_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int start_create(start_ctx** out, const start_mesh_desc* mesh, const start_params* params)
{
    if (!out || !mesh || !params) return fail(nullptr, START_ERR_ARG, "null argument");
    *out = nullptr;
    if (params->n_species <= 0 || params->n_species > kMaxSpecies || !params->species)
        return fail(nullptr, START_ERR_ARG, "n_species must be in [1, 32] with a species table");
    if (params->precision != START_PREC_FP64 && params->precision != START_PREC_MIXED)
        return fail(nullptr, START_ERR_ARG, "unknown precision");
    if (params->gas.model < START_MODEL_NONE || params->gas.model > START_MODEL_VSS)
        return fail(nullptr, START_ERR_ARG, "unknown collision model");
    if (params->steps_per_launch < 0 || params->sort_interval < 0)
        return fail(nullptr, START_ERR_ARG, "steps_per_launch and sort_interval must be >= 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, START_ERR_HIP, "no HIP device available");
    start_ctx* c = new (std::nothrow) start_ctx();
    if (!c) return fail(nullptr, START_ERR_NOMEM, "host allocation failed");
    c->prm = *params;
    c->prm.species = nullptr;
    c->cfg.steps_per_launch = std::max(1, params->steps_per_launch);
    c->cfg.block = 256;
    c->cfg.persistent = 0;
    c->cfg.sort_interval = params->sort_interval;
    c->cfg.graph = 0;
    c->cfg.walk = 1;
    {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
            c->n_cu = ncu;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) c->arch = prop.gcnArchName;
    }
    c->mixed = params->precision == START_PREC_MIXED;
    c->es = c->mixed ? sizeof(float) : sizeof(double);
    double expansion = params->bbox_expansion > 0.0 ? params->bbox_expansion : 1.01;
    std::string e = prepare_mesh(c->hm, mesh->n_nodes, mesh->nodes, mesh->n_tets, mesh->tets, mesh->n_tris,
                                 mesh->tris, expansion);
    if (!e.empty()) {
        delete c;
        return fail(nullptr, START_ERR_MESH, e);
    }
    for (int s = 0; s < params->n_species; ++s) {
        c->species_desc.push_back(params->species[s]);
        c->species.push_back(species_constants(params->species[s], params->gas));
    }
    RunK r0{};
    r0.conc = params->gas.concentration;
    r0.s0 = 1.0;
    r0.s1 = 0.0;
    r0.key = params->run;
    c->runs.assign(1, r0);
    std::vector<float> g32;
    if (c->mixed) {
        c->tau32 = build_geo32(c->hm, g32);
    }

    auto up = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    const HostMesh& hm = c->hm;
    int64_t T = hm.n_tets, M = hm.n_nodes, S = hm.n_tris;
    std::vector<int4> tn((size_t)T), tb((size_t)T);
    for (int64_t t = 0; t < T; ++t) {
        tn[(size_t)t] = make_int4(hm.tets[4 * t], hm.tets[4 * t + 1], hm.tets[4 * t + 2], hm.tets[4 * t + 3]);
        tb[(size_t)t] = make_int4(hm.nbr[4 * t], hm.nbr[4 * t + 1], hm.nbr[4 * t + 2], hm.nbr[4 * t + 3]);
    }
    hipError_t he = hipSuccess;
    auto chk = [&](hipError_t x) {
        if (he == hipSuccess && x != hipSuccess) he = x;
    };
    chk(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    chk(scratch_address(&c->scratch));
    chk(dalloc(&c->d_nodes, 3 * M));
    chk(dalloc(&c->d_geo, 12 * T));
    chk(dalloc(&c->d_E, 3 * T));
    if (c->mixed) {
        chk(dalloc(&c->d_geo32, 12 * T));
        chk(dalloc(&c->d_E32, 4 * T));
    }
    chk(dalloc(&c->d_tnode, T));
    chk(dalloc(&c->d_tnbr, T));
    chk(dalloc(&c->d_tris, 3 * S));
    chk(dalloc(&c->d_tri_deg, S));
    chk(dalloc(&c->d_tgs, hm.tgrid_start.size()));
    chk(dalloc(&c->d_tgl, hm.tgrid_list.size()));
    chk(dalloc(&c->d_sgs, hm.sgrid_start.size()));
    chk(dalloc(&c->d_sgl, hm.sgrid_list.size()));
    chk(dalloc(&c->d_species, c->species.size()));
    chk(dalloc(&c->d_cnt, kCntTotal));
    chk(dalloc(&c->d_nactive, 1));
    chk(dalloc(&c->d_runs, 1));
    chk(dalloc(&c->d_node_charge, 2 * M));
    chk(dalloc(&c->d_tri_count, std::max<int64_t>(S, 1)));
    if (he == hipSuccess) {
        chk(up(c->d_nodes, hm.nodes.data(), sizeof(double) * 3 * M));
        chk(up(c->d_geo, hm.geo.data(), sizeof(double) * 12 * T));
        chk(hipMemset(c->d_E, 0, sizeof(double) * 3 * T));
        if (c->mixed) {
            chk(up(c->d_geo32, g32.data(), sizeof(float) * 12 * T));
            chk(hipMemset(c->d_E32, 0, sizeof(float) * 4 * T));
        }
        chk(up(c->d_tnode, tn.data(), sizeof(int4) * T));
        chk(up(c->d_tnbr, tb.data(), sizeof(int4) * T));
        if (S) chk(up(c->d_tris, hm.tris.data(), sizeof(int32_t) * 3 * S));
        if (S) chk(up(c->d_tri_deg, hm.tri_degenerate.data(), S));
        chk(up(c->d_tgs, hm.tgrid_start.data(), sizeof(int32_t) * hm.tgrid_start.size()));
        chk(up(c->d_tgl, hm.tgrid_list.data(), sizeof(int32_t) * hm.tgrid_list.size()));
        chk(up(c->d_sgs, hm.sgrid_start.data(), sizeof(int32_t) * hm.sgrid_start.size()));
        chk(up(c->d_sgl, hm.sgrid_list.data(), sizeof(int32_t) * hm.sgrid_list.size()));
        chk(hipMemset(c->d_tri_count, 0, sizeof(unsigned long long) * std::max<int64_t>(S, 1)));
        chk(hipMemset(c->d_node_charge, 0, sizeof(double) * 2 * M));
        chk(up(c->d_species, c->species.data(), sizeof(SpeciesK) * c->species.size()));
        chk(up(c->d_runs, c->runs.data(), sizeof(RunK)));
        chk(hipMemset(c->d_cnt, 0, sizeof(unsigned long long) * kCntTotal));
    }
    if (he != hipSuccess) {
        std::string msg = std::string("device set-up failed: ") + hipGetErrorString(he);
        start_destroy(c);
        return fail(nullptr, he == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP, msg);
    }
    *out = c;
    return START_OK;
}

void start_destroy(start_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_particles(c);
    void* ptrs[] = {c->d_nodes,     c->d_geo,       c->d_geo32,       c->d_E,         c->d_E32,      c->d_E1,
                    c->d_E1_32,     c->d_tnode,     c->d_tnbr,        c->d_tris,      c->d_tri_deg,  c->d_tgs,
                    c->d_tgl,       c->d_sgs,       c->d_sgl,         c->d_tri_count, c->d_node_charge,
                    c->d_species,   c->d_cnt,       c->d_nactive,     c->d_runs,      c->d_tet_vol,  c->d_node_vol,
                    c->d_tet_q,     c->d_adj_start, c->d_adj_list,    c->d_tet_n,     c->d_phi,
                    c->d_rho,       c->d_written,   c->d_H,           c->d_H32};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->h_phi) (void)hipHostFree(c->h_phi);
    if (c->h_rho) (void)hipHostFree(c->h_rho);
    if (c->h_written) (void)hipHostFree(c->h_written);
    if (c->h_cnt) (void)hipHostFree(c->h_cnt);
    if (c->h_settle) (void)hipHostFree(c->h_settle);
    if (c->phi_copied) (void)hipEventDestroy(c->phi_copied);
    for (auto& e : c->events) {
        (void)hipEventDestroy(e.a0);
        (void)hipEventDestroy(e.a1);
        (void)hipEventDestroy(e.b1);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int start_load_batch(start_ctx* c, int32_t n_runs, const start_run_desc* runs, const double* x, const double* y,
                     const double* z, const double* vx, const double* vy, const double* vz, const int32_t* species,
                     const uint64_t* ids)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    if (n_runs < 1 || n_runs > kMaxRuns || !runs) return fail(c, START_ERR_ARG, "n_runs must be in [1, 32768]");
    int64_t n = 0;
    for (int32_t r = 0; r < n_runs; ++r) {
        if (runs[r].n_particles < 0) return fail(c, START_ERR_ARG, "negative run size");
        if (!(runs[r].concentration >= 0.0)) return fail(c, START_ERR_ARG, "negative gas concentration");
        n += runs[r].n_particles;
    }
    if (n > (int64_t)0x7fffffff) return fail(c, START_ERR_ARG, "particle count out of range");
    if ((int64_t)n_runs * c->hm.n_nodes > (int64_t)0x7fffffff)
        return fail(c, START_ERR_ARG, "n_runs * n_nodes exceeds the node-charge index range");
    if (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !species)) return fail(c, START_ERR_ARG, "null particle array");
    std::vector<int32_t> flags((size_t)n);
    std::vector<RunK> rk((size_t)n_runs);
    {
        int64_t i = 0;
        for (int32_t r = 0; r < n_runs; ++r) {
            rk[(size_t)r].conc = runs[r].concentration;
            rk[(size_t)r].s0 = runs[r].field_scale[0];
            rk[(size_t)r].s1 = runs[r].field_scale[1];
            rk[(size_t)r].key = runs[r].run;
            rk[(size_t)r].first = (uint32_t)i;
            rk[(size_t)r].node_base = (int64_t)r * c->hm.n_nodes;
            rk[(size_t)r].tri_base = (int64_t)r * c->hm.n_tris;
            for (int64_t j = 0; j < runs[r].n_particles; ++j, ++i) {
                if (species[i] < 0 || species[i] >= (int32_t)c->species.size())
                    return fail(c, START_ERR_ARG, "species index out of range at particle " + std::to_string(i));
                flags[(size_t)i] = (species[i] << kSpeciesShift) | (r << kRunShift);
            }
        }
    }
    int rc = begin_population(c, rk, n, ids);
    if (rc != START_OK) return rc;
    c->present.assign(c->species.size(), 0);
    rc = mark_present(c, species, n);
    if (rc != START_OK || n == 0) return rc;
    int64_t nact = n;
    HIPCHECK(c, hipMemcpy(c->d_nactive, &nact, sizeof(int64_t), hipMemcpyHostToDevice));
    rc = upload_rows(c, 0, n, x, y, z, vx, vy, vz, flags.data());
    if (rc != START_OK) return rc;
    HIPCHECK(c, hipGetLastError());
    if (c->cfg.sort_interval > 0) {
        int rc2 = sort_now(c);
        if (rc2 != START_OK) return rc2;
    }
    return check_kernel_errors(c);
}

int start_generate(start_ctx* c, int32_t n_sources, const start_source_desc* src, uint64_t seed)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    if (n_sources < 1 || !src) return fail(c, START_ERR_ARG, "no particle sources");
    std::vector<SourceK> sk((size_t)n_sources);
    int64_t n = 0, n_cells = 0;
    for (int32_t s = 0; s < n_sources; ++s) {
        const start_source_desc& d = src[s];
        const std::string where = "source " + std::to_string(s) + ": ";
        if (d.kind != START_SOURCE_POINT && d.kind != START_SOURCE_SURFACE)
            return fail(c, START_ERR_ARG, where + "unknown source kind");
        if (d.species < 0 || d.species >= (int32_t)c->species.size())
            return fail(c, START_ERR_ARG, where + "species index out of range");
        if (d.count <= 0)  // ParticleGeneratorHost.cpp:55-59
            return fail(c, START_ERR_ARG, where + "there is no need to generate 0 objects");
        if (d.energy < START_ENERGY_MONO || d.energy > START_ENERGY_LOGNORMAL)
            return fail(c, START_ERR_ARG, where + "unknown energy distribution");
        if (d.angle < START_ANGLE_FAN || d.angle > START_ANGLE_CONE)
            return fail(c, START_ERR_ARG, where + "unknown direction distribution");
        if (d.energy == START_ENERGY_MONO && d.energy_eV == 0.0)  // ParticleGeneratorHost.cpp:61-66
            return fail(c, START_ERR_ARG, where + "source with zero energy");
        if (d.energy == START_ENERGY_MAXWELLIAN && !(d.p0 > 0.0))
            return fail(c, START_ERR_ARG, where + "Maxwellian temperature must be > 0");
        if (d.energy == START_ENERGY_THOMPSON && !(d.p0 > 0.0 && d.p1 > 0.0))
            return fail(c, START_ERR_ARG, where + "Thompson U_b and E_max must be > 0");
        if (d.energy == START_ENERGY_LOGNORMAL && !(d.p0 > 0.0 && d.p1 >= 0.0))
            return fail(c, START_ERR_ARG, where + "lognormal mean must be > 0 and sigma >= 0");
        SourceK& k = sk[(size_t)s];
        k.kind = d.kind;
        k.energy = d.energy;
        k.angle = d.angle;
        k.flags = d.species << kSpeciesShift;
        k.first = n;
        k.count = d.count;
        k.energy_eV = d.energy_eV;
        k.p0 = d.p0;
        k.p1 = d.p1;
        k.mass = c->species[(size_t)d.species].mass;
        for (int q = 0; q < 3; ++q) k.pos[q] = d.position[q];
        k.theta = d.theta;
        k.phi = d.phi;
        k.expansion = d.expansion;
        k.displace = d.displace;
        k.cell0 = n_cells;
        k.n_cells = 0;
        if (d.kind == START_SOURCE_SURFACE) {
            if (!d.centres || !d.normals || d.n_cells <= 0)  // ParticleGeneratorHost.cpp:103-107
                return fail(c, START_ERR_ARG, where + "base coordinates for surface source are empty");
            for (int64_t j = 0; j < d.n_cells; ++j) {
                const double* nn = d.normals + 3 * j;
                if (nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2] == 0.0)  // :130-134
                    return fail(c, START_ERR_ARG, where + "normal vector magnitude is zero");
            }
            k.n_cells = d.n_cells;
            n_cells += d.n_cells;
        }
        n += d.count;
    }
    if (n > (int64_t)0x7fffffff) return fail(c, START_ERR_ARG, "particle count out of range");
    // Cells of every surface source in one array, each source's cells in a
    // seeded random order: the first count % n_cells of them take one particle
    // more, as the reference hands the remainder to randomly shuffled cells
    // (ParticleGeneratorHost.cpp:109-118).  Fisher-Yates with Philox variates
    // (counter {k, 2, source, 0}).
    std::vector<double> cen((size_t)std::max<int64_t>(3 * n_cells, 3)), nor(cen.size());
    for (int32_t s = 0; s < n_sources; ++s) {
        const SourceK& k = sk[(size_t)s];
        if (k.n_cells == 0) continue;
        std::vector<int64_t> order((size_t)k.n_cells);
        for (int64_t j = 0; j < k.n_cells; ++j) order[(size_t)j] = j;
        for (int64_t j = k.n_cells - 1; j > 0; --j) {
            uint32_t w[4];
            philox4x32_10((uint32_t)j, 2u, (uint32_t)s, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), w);
            const double u = (double)(((uint64_t)w[0] << 21) | (w[1] >> 11)) * (1.0 / 9007199254740992.0);
            const int64_t r = std::min<int64_t>((int64_t)(u * (double)(j + 1)), j);
            std::swap(order[(size_t)j], order[(size_t)r]);
        }
        for (int64_t j = 0; j < k.n_cells; ++j)
            for (int q = 0; q < 3; ++q) {
                cen[(size_t)(3 * (k.cell0 + j) + q)] = src[s].centres[3 * order[(size_t)j] + q];
                nor[(size_t)(3 * (k.cell0 + j) + q)] = src[s].normals[3 * order[(size_t)j] + q];
            }
    }
    std::vector<RunK> rk(1);
    rk[0].conc = c->prm.gas.concentration;
    rk[0].s0 = 1.0;
    rk[0].s1 = 0.0;
    rk[0].key = c->prm.run;
    rk[0].first = 0;
    rk[0].node_base = 0;
    rk[0].tri_base = 0;
    int rc = begin_population(c, rk, n, nullptr);
    if (rc != START_OK) return rc;
    {
        std::vector<int32_t> sps((size_t)n_sources);
        for (int32_t s = 0; s < n_sources; ++s) sps[(size_t)s] = src[s].species;
        c->present.assign(c->species.size(), 0);
        rc = mark_present(c, sps.data(), n_sources);
        if (rc != START_OK) return rc;
    }
    rc = init_population_tail(c, n);
    if (rc != START_OK) return rc;
    double *d_cen = nullptr, *d_nor = nullptr, *t[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipError_t e = dalloc(&d_cen, cen.size());
    if (e == hipSuccess) e = dalloc(&d_nor, nor.size());
    if (e == hipSuccess) e = hipMemcpy(d_cen, cen.data(), sizeof(double) * cen.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_nor, nor.data(), sizeof(double) * nor.size(), hipMemcpyHostToDevice);
    // FP64 state: drawn in place; mixed: positions drawn into FP64 scratch,
    // then to the owner's frame (by the locate); velocities in place
    double* dst[6] = {(double*)c->A.x, (double*)c->A.y, (double*)c->A.z,
                      (double*)c->A.vx, (double*)c->A.vy, (double*)c->A.vz};
    if (c->mixed)
        for (int q = 0; q < 3 && e == hipSuccess; ++q) {
            e = dalloc(&t[q], n);
            dst[q] = t[q];
        }
    if (e == hipSuccess) {
        for (int32_t s = 0; s < n_sources; ++s)
            launch_generate(c->stream, sk[(size_t)s], (uint32_t)s, seed, d_cen, d_nor, dst[0], dst[1], dst[2],
                            dst[3], dst[4], dst[5], c->A.flags);
        DevMesh M = c->mesh_view();
        if (!c->mixed) {
            launch_locate(c->stream, M, dst[0], dst[1], dst[2], n, c->A.tet, c->A.flags, nullptr, nullptr, nullptr,
                          c->d_cnt);
        } else {
            launch_locate(c->stream, M, dst[0], dst[1], dst[2], n, c->A.tet, c->A.flags, (float*)c->A.x,
                          (float*)c->A.y, (float*)c->A.z, c->d_cnt);
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(d_cen);
    (void)hipFree(d_nor);
    for (double* p : t)
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("particle sources: ") + hipGetErrorString(e));
    if (c->cfg.sort_interval > 0) {
        rc = sort_now(c);
        if (rc != START_OK) return rc;
    }
    return check_kernel_errors(c);
}

int start_append_particles(start_ctx* c, int64_t m, const double* x, const double* y, const double* z,
                           const double* vx, const double* vy, const double* vz, const int32_t* species,
                           const uint64_t* ids)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    if (m < 0 || c->n + m > (int64_t)0x7fffffff) return fail(c, START_ERR_ARG, "particle count out of range");
    if (m == 0) return START_OK;
    if (c->n_runs() != 1) return fail(c, START_ERR_STATE, "particles can be appended to a single simulation only");
    if (!x || !y || !z || !vx || !vy || !vz || !species) return fail(c, START_ERR_ARG, "null particle array");
    std::vector<int32_t> flags((size_t)m);
    for (int64_t i = 0; i < m; ++i) {
        if (species[i] < 0 || species[i] >= (int32_t)c->species.size())
            return fail(c, START_ERR_ARG, "species index out of range at particle " + std::to_string(i));
        flags[(size_t)i] = species[i] << kSpeciesShift;  // run 0
    }
    const int64_t n0 = c->n;
    int rc = grow_particles(c, n0 + m);
    if (rc != START_OK) return rc;
    rc = mark_present(c, species, m);
    if (rc != START_OK) return rc;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    // launches read their first step from the device counter, which steps
    // taken while the context held no particles did not advance
    const unsigned long long step_now = (unsigned long long)c->step;
    HIPCHECK(c, hipMemcpy(c->d_cnt + kCntStep + c->wslot, &step_now, sizeof(step_now), hipMemcpyHostToDevice));
    rc = upload_rows(c, n0, m, x, y, z, vx, vy, vz, flags.data());
    if (rc != START_OK) return rc;
    c->n = n0 + m;
    for (int64_t i = 0; i < m; ++i) c->ids.push_back(ids ? ids[i] : (uint64_t)(n0 + i));
    // after a sort the settled particles close the live prefix; the kernels
    // cover every row until the next sort compacts them again
    int64_t nact = c->n;
    HIPCHECK(c, hipMemcpy(c->d_nactive, &nact, sizeof(int64_t), hipMemcpyHostToDevice));
    drop_graphs(c);  // captured launches hold the old particle count
    rc = assign_track_slots(c, n0);
    if (rc != START_OK) return rc;
    return check_kernel_errors(c);
}

int start_load_particles(start_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                         const double* vx, const double* vy, const double* vz, const int32_t* species,
                         const uint64_t* ids)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    if (n < 0 || n > (int64_t)0x7fffffff) return fail(c, START_ERR_ARG, "particle count out of range");
    start_run_desc r{};
    r.n_particles = n;
    r.field_scale[0] = 1.0;
    r.field_scale[1] = 0.0;
    r.concentration = c->prm.gas.concentration;
    r.run = c->prm.run;
    return start_load_batch(c, 1, &r, x, y, z, vx, vy, vz, species, ids);
}

int start_set_gas(start_ctx* c, const start_gas_desc* g, const double* diameters)
{
    if (!c || !g) return fail(c, START_ERR_ARG, "null argument");
    if (g->model < START_MODEL_NONE || g->model > START_MODEL_VSS) return fail(c, START_ERR_ARG, "unknown collision model");
    if (!(g->concentration >= 0.0)) return fail(c, START_ERR_ARG, "negative gas concentration");
    std::vector<start_species_desc> sd = c->species_desc;
    if (diameters)
        for (size_t s = 0; s < sd.size(); ++s) {
            if (!(diameters[s] >= 0.0)) return fail(c, START_ERR_ARG, "negative collision diameter");
            sd[s].diameter = diameters[s];
        }
    std::vector<SpeciesK> k;
    for (size_t s = 0; s < sd.size(); ++s) {
        k.push_back(species_constants(sd[s], *g));
        k.back().push_slot = c->species[s].push_slot;  // masses and charges are unchanged
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(c->d_species, k.data(), sizeof(SpeciesK) * k.size(), hipMemcpyHostToDevice));
    c->species_desc = sd;
    c->species = k;
    c->prm.gas = *g;
    if (c->n_runs() == 1) {  // a single simulation takes the concentration with the gas
        c->runs[0].conc = g->concentration;
        int rc = upload_runs(c);
        if (rc != START_OK) return rc;
    }
    drop_graphs(c);
    return START_OK;
}

int start_set_time_step(start_ctx* c, double dt)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    if (!(dt >= 0.0)) return fail(c, START_ERR_ARG, "time step must be >= 0");
    if (dt != c->prm.dt) {
        drop_graphs(c);
        c->push_dirty = true;
    }
    c->prm.dt = dt;
    return START_OK;
}

int start_set_tet_field_basis(start_ctx* c, int32_t which, const double* E)
{
    if (!c || !E) return fail(c, START_ERR_ARG, "null argument");
    if (which != 0 && which != 1) return fail(c, START_ERR_ARG, "field basis must be 0 or 1");
    const int64_t T = c->hm.n_tets;
    if (which == 1 && !c->d_E1) {
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        HIPCHECK(c, dalloc(&c->d_E1, 3 * T));
        if (c->mixed) HIPCHECK(c, dalloc(&c->d_E1_32, 4 * T));
        drop_graphs(c);
    }
    double* dE = which ? c->d_E1 : c->d_E;
    float* dE32 = which ? c->d_E1_32 : c->d_E32;
    HIPCHECK(c, hipMemcpyAsync(dE, E, sizeof(double) * 3 * T, hipMemcpyHostToDevice, c->stream));
    if (dE32) launch_e32(c->stream, dE, dE32, T);
    if (which == 0) c->push_dirty = true;
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return START_OK;
}

int start_set_tet_field(start_ctx* c, const double* E) { return start_set_tet_field_basis(c, 0, E); }

int start_read_tet_field(start_ctx* c, int32_t which, double* E)
{
    if (!c || !E) return fail(c, START_ERR_ARG, "null argument");
    if (which != 0 && which != 1) return fail(c, START_ERR_ARG, "field basis must be 0 or 1");
    const double* dE = which ? c->d_E1 : c->d_E;
    if (!dE) return fail(c, START_ERR_STATE, "field basis 1 was never set");
    HIPCHECK(c, hipMemcpyAsync(E, dE, sizeof(double) * 3 * (size_t)c->hm.n_tets, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return START_OK;
}

int start_set_node_potential_basis(start_ctx* c, int32_t which, const double* phi)
{
    if (!c || !phi) return fail(c, START_ERR_ARG, "null argument");
    if (which != 0 && which != 1) return fail(c, START_ERR_ARG, "field basis must be 0 or 1");
    const int64_t T = c->hm.n_tets;
    if (which == 1 && !c->d_E1) {
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        HIPCHECK(c, dalloc(&c->d_E1, 3 * T));
        if (c->mixed) HIPCHECK(c, dalloc(&c->d_E1_32, 4 * T));
        drop_graphs(c);
    }
    // Stream-ordered and asynchronous: launches already queued use the field
    // they were queued with, later ones the new one.  The caller's array is
    // copied into pinned staging before the call returns; the only wait is for
    // the previous call's copy out of that staging buffer.
    const size_t bytes = sizeof(double) * (size_t)c->hm.n_nodes;
    if (!c->d_phi) {
        HIPCHECK(c, dalloc(&c->d_phi, c->hm.n_nodes));
        HIPCHECK(c, host_pinned_malloc((void**)&c->h_phi, std::max<size_t>(bytes, 8)));
        HIPCHECK(c, hipEventCreateWithFlags(&c->phi_copied, hipEventDisableTiming));
    }
    if (c->phi_pending) {
        HIPCHECK(c, hipEventSynchronize(c->phi_copied));
        c->phi_pending = false;
    }
    std::memcpy(c->h_phi, phi, bytes);
    HIPCHECK(c, hipMemcpyAsync(c->d_phi, c->h_phi, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipEventRecord(c->phi_copied, c->stream));
    c->phi_pending = true;
    launch_field(c->stream, c->mesh_view(), c->d_phi, which ? c->d_E1 : c->d_E, which ? c->d_E1_32 : c->d_E32);
    if (which == 0) c->push_dirty = true;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("field from potential: ") + hipGetErrorString(e));
    return START_OK;
}

int start_set_node_potential(start_ctx* c, const double* phi) { return start_set_node_potential_basis(c, 0, phi); }

int start_set_launch(start_ctx* c, const start_launch_cfg* cfg)
{
    if (!c || !cfg) return fail(c, START_ERR_ARG, "null argument");
    int rc = check_cfg(c, *cfg);
    if (rc != START_OK) return rc;
    c->cfg = *cfg;
    c->prm.steps_per_launch = cfg->steps_per_launch;
    c->prm.sort_interval = cfg->sort_interval;
    return START_OK;
}

int start_get_launch(start_ctx* c, start_launch_cfg* cfg)
{
    if (!c || !cfg) return fail(c, START_ERR_ARG, "null argument");
    *cfg = c->cfg;
    return START_OK;
}

int start_step(start_ctx* c, int32_t nsteps)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    if (nsteps < 0) return fail(c, START_ERR_ARG, "negative step count");
    if (c->n == 0) {
        c->step += (uint64_t)nsteps;
        return START_OK;
    }
    if (!c->d_wl || !c->B.x) return fail(c, START_ERR_STATE, "particle work buffers missing (a failed append)");
    {
        int rc = refresh_push_tables(c);
        if (rc != START_OK) return rc;
    }
    // Launches of steps_per_launch steps; before a launch, a sort if
    // sort_interval steps have passed since the last one.  A chunk is the run
    // of launches between two sorts.
    int32_t done = 0;
    const int32_t SI = c->cfg.sort_interval;
    const int32_t K = std::max(1, c->cfg.steps_per_launch);
    while (done < nsteps) {
        if (SI > 0 && c->step - c->last_sort >= (uint64_t)SI) {
            int rc = sort_now(c);
            if (rc != START_OK) return rc;
        }
        int32_t chunk = nsteps - done;
        if (SI > 0) {
            const int64_t to_sort = (int64_t)SI - (int64_t)(c->step - c->last_sort);
            chunk = (int32_t)std::min<int64_t>(chunk, (to_sort + K - 1) / K * K);
        }
        int rc = (c->cfg.graph && !c->timing) ? run_chunk_graph(c, chunk) : enqueue_chunk(c, chunk);
        if (rc != START_OK) return rc;
        c->step += (uint64_t)chunk;
        done += chunk;
    }
    return START_OK;
}

int start_autotune(start_ctx* c, int32_t trial_steps, start_launch_cfg* best, double* ms_per_step)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    if (trial_steps < 1) return fail(c, START_ERR_ARG, "trial_steps must be >= 1");
    if (c->n == 0) return fail(c, START_ERR_STATE, "no particles loaded");
    const start_launch_cfg keep = c->cfg;
    const bool timing = c->timing;
    c->timing = false;
    Snapshot snap;
    int rc = snapshot_take(c, snap);
    if (rc != START_OK) {
        snapshot_free(snap);
        c->timing = timing;
        return rc;
    }
    // the trials record no movements (the restored state does not undo them)
    int32_t* const trk_keep = c->d_trk_slot;
    c->d_trk_slot = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    start_launch_cfg best_cfg = keep;
    double best_ms = -1.0;
    std::vector<std::pair<double, start_launch_cfg>> timed;
    // trial_steps steps of configuration g from the snapshot; keeps the fastest
    auto trial = [&](const start_launch_cfg& g) -> int {
        int r = snapshot_restore(c, snap);
        if (r != START_OK) return r;
        c->cfg = g;
        (void)hipEventRecord(e0, c->stream);
        r = start_step(c, trial_steps);
        (void)hipEventRecord(e1, c->stream);
        if (r != START_OK) return r;
        if (hipEventSynchronize(e1) != hipSuccess) return fail(c, START_ERR_HIP, "autotune timing");
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double per = ms / trial_steps;
        timed.emplace_back(per, g);
        if (best_ms < 0.0 || per < best_ms) {
            best_ms = per;
            best_cfg = g;
        }
        return START_OK;
    };
    const int32_t Ks[] = {10, 25, 50, 100, 200};
    const int32_t blocks[] = {128, 256};
    for (int32_t K : Ks) {
        if (K > trial_steps) continue;
        for (int32_t bs : blocks)
            for (int32_t pers = 0; pers <= 1 && rc == START_OK; ++pers)
                for (int32_t sm = 1; sm <= 4 && rc == START_OK; sm *= 2) {
                    // timed eagerly: a trial under capture would pay its own
                    // graph instantiation; the kept shape gets the caller's graph setting
                    start_launch_cfg g{K, bs, pers, K * sm, 0, keep.walk};
                    if (g.sort_interval > trial_steps) continue;
                    rc = trial(g);
                }
    }
    if (rc == START_OK && best_ms >= 0.0) {  // the multi-tet walk the other way round
        start_launch_cfg g = best_cfg;
        g.walk = 1 - g.walk;
        if (g.walk) g.block = 256;
        rc = trial(g);
    }
    if (rc == START_OK && timed.size() > 1) {
        // one timing of a few ms differs by some % from run to run: the three
        // fastest are timed twice more and ranked by the better of those two
        std::sort(timed.begin(), timed.end(),
                  [](const auto& a, const auto& b) { return a.first < b.first; });
        std::vector<start_launch_cfg> top;
        for (size_t i = 0; i < timed.size() && top.size() < 3; ++i) top.push_back(timed[i].second);
        best_ms = -1.0;
        timed.clear();
        for (int rep = 0; rep < 2 && rc == START_OK; ++rep)
            for (size_t i = 0; i < top.size() && rc == START_OK; ++i) rc = trial(top[i]);
    }
    int rc2 = snapshot_restore(c, snap);
    snapshot_free(snap);
    drop_graphs(c);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->timing = timing;
    c->d_trk_slot = trk_keep;
    best_cfg.graph = keep.graph;
    c->cfg = rc == START_OK ? best_cfg : keep;
    c->prm.steps_per_launch = c->cfg.steps_per_launch;
    c->prm.sort_interval = c->cfg.sort_interval;
    if (rc != START_OK) return rc;
    if (rc2 != START_OK) return rc2;
    if (best) *best = best_cfg;
    if (ms_per_step) *ms_per_step = best_ms;
    return check_kernel_errors(c);
}

namespace {

// The autotuner's cache key: what the launch shape's speed depends on -- the
// device, the precision, the collision model and the work per step (the
// particle count by power of two, the tets, the runs) -- and the trial length.
std::string autotune_key(const start_ctx* c, int32_t trial_steps)
{
    int64_t bucket = 1;
    while (bucket < c->n) bucket <<= 1;
    std::string k = "arch=" + c->arch + ";cu=" + std::to_string(c->n_cu) + ";prec=" + (c->mixed ? "mixed" : "fp64") +
                    ";model=" + std::to_string(c->prm.gas.model) + ";em=" + std::to_string(c->prm.em_push) +
                    ";charge=" + std::to_string(c->prm.charge_every_step) + ";runs=" + std::to_string(c->n_runs()) +
                    ";tets=" + std::to_string(c->hm.n_tets) + ";n<=" + std::to_string(bucket) +
                    ";trial=" + std::to_string(trial_steps);
    for (char& ch : k)
        if (ch == ' ' || ch == '\n' || ch == '|') ch = '_';
    return k;
}

}  // namespace

int start_autotune_cached(start_ctx* c, const char* cache_path, int32_t trial_steps, start_launch_cfg* best,
                          double* ms_per_step, int32_t* from_cache)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    if (!cache_path) return fail(c, START_ERR_ARG, "null cache path");
    if (trial_steps < 1) return fail(c, START_ERR_ARG, "trial_steps must be >= 1");
    if (c->n == 0) return fail(c, START_ERR_STATE, "no particles loaded");
    const std::string key = autotune_key(c, trial_steps);
    // lines "key|K block persistent sort_interval walk|ms"; the last match wins
    bool found = false;
    start_launch_cfg g = c->cfg;
    double ms = 0.0;
    if (FILE* f = fopen(cache_path, "r")) {
        char line[1024];
        while (fgets(line, sizeof line, f)) {
            const char* bar = strchr(line, '|');
            if (!bar || std::string(line, (size_t)(bar - line)) != key) continue;
            start_launch_cfg h = c->cfg;
            double t = 0.0;
            if (sscanf(bar + 1, "%d %d %d %d %d|%lf", &h.steps_per_launch, &h.block, &h.persistent,
                       &h.sort_interval, &h.walk, &t) == 6 &&
                h.steps_per_launch >= 1 && (h.block == 128 || h.block == 256) && (h.persistent == 0 || h.persistent == 1) &&
                h.sort_interval >= 0 && (h.walk == 0 || h.walk == 1)) {
                g = h;
                ms = t;
                found = true;
            }
        }
        fclose(f);
    }
    if (found) {
        g.graph = c->cfg.graph;
        int rc = start_set_launch(c, &g);
        if (rc != START_OK) return rc;
        if (best) *best = g;
        if (ms_per_step) *ms_per_step = ms;
        if (from_cache) *from_cache = 1;
        return START_OK;
    }
    start_launch_cfg b{};
    int rc = start_autotune(c, trial_steps, &b, &ms);
    if (rc != START_OK) return rc;
    // the tuned configuration is set either way; a cache that cannot be
    // written is reported after the outputs are filled in
    if (best) *best = b;
    if (ms_per_step) *ms_per_step = ms;
    if (from_cache) *from_cache = 0;
    FILE* f = fopen(cache_path, "a");
    if (!f) return fail(c, START_ERR_ARG, std::string("cannot append to the autotune cache ") + cache_path);
    const bool ok = fprintf(f, "%s|%d %d %d %d %d|%.6g\n", key.c_str(), b.steps_per_launch, b.block, b.persistent,
                            b.sort_interval, b.walk, ms) > 0;
    if (fclose(f) != 0 || !ok)
        return fail(c, START_ERR_ARG, std::string("cannot append to the autotune cache ") + cache_path);
    return START_OK;
}

int start_track_movements(start_ctx* c, int64_t max_particles, int32_t max_points)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    if (max_particles < 0 || max_points < 0 || (max_particles > 0 && max_points == 0))
        return fail(c, START_ERR_ARG, "max_particles >= 0 and max_points >= 1 are required");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);  // captured launches hold the record pointers
    free_tracking(c);
    if (max_particles == 0) return START_OK;
    // the first max_particles particles, in load order, that the next step
    // records (the reference's map fills in its loop order while it has room);
    // particles appended later join while there is room
    c->trk_cap = max_particles;
    c->trk_max = max_points;
    return assign_track_slots(c, 0);
}

int start_read_movements(start_ctx* c, int64_t* n, uint64_t* ids, int32_t* counts, double* xyz)
{
    if (!c || !n) return fail(c, START_ERR_ARG, "null argument");
    const int64_t m = (int64_t)c->trk_rows.size();
    if (!ids && !counts && !xyz) {
        *n = m;
        return START_OK;
    }
    if (*n < m) return fail(c, START_ERR_ARG, "capacity below the number of tracked particles");
    *n = m;
    if (m == 0) return START_OK;
    if (ids)
        for (int64_t s = 0; s < m; ++s) ids[s] = c->ids[c->trk_rows[s]];
    hipError_t e = hipSuccess;
    if (counts)
        e = hipMemcpyAsync(counts, c->d_trk_count, sizeof(int32_t) * (size_t)m, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && xyz)
        e = hipMemcpyAsync(xyz, c->d_trk_xyz, sizeof(double) * 3 * (size_t)m * (size_t)c->trk_max,
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("movements: ") + hipGetErrorString(e));
    return check_kernel_errors(c);
}

int start_clear_movements(start_ctx* c)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    const int64_t m = (int64_t)c->trk_rows.size();
    if (m == 0) return START_OK;
    HIPCHECK(c, hipMemsetAsync(c->d_trk_count, 0, sizeof(int32_t) * (size_t)m, c->stream));
    return START_OK;
}

int start_record_settle(start_ctx* c, int32_t on)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);  // captured launches hold the record pointer
    c->record = on != 0;
    if (c->record && !c->d_settle && c->n > 0) {  // one record per load index, appends included
        HIPCHECK(c, dalloc(&c->d_settle, c->cap));
        HIPCHECK(c, hipMemset(c->d_settle, 0xff, sizeof(SettleRec) * c->cap));
    }
    return START_OK;
}

int start_read_hits(start_ctx* c, double* hit_xyz, int32_t* hit_tri, int32_t* settle_step)
{
    if (!c || !hit_xyz || !hit_tri || !settle_step) return fail(c, START_ERR_ARG, "null argument");
    if (c->n == 0) return START_OK;
    if (!c->d_settle) return fail(c, START_ERR_STATE, "settle records are off (start_record_settle)");
    const int64_t n = c->n;
    double* dx = nullptr;
    int32_t *dt = nullptr, *ds = nullptr;
    hipError_t e = dalloc(&dx, 3 * n);
    if (e == hipSuccess) e = dalloc(&dt, n);
    if (e == hipSuccess) e = dalloc(&ds, n);
    if (e == hipSuccess) {
        if (c->mixed) launch_hit_points<float>(c->stream, c->mesh_view(), c->view<float>(), c->d_settle, dx, dt, ds);
        else launch_hit_points<double>(c->stream, c->mesh_view(), c->view<double>(), c->d_settle, dx, dt, ds);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hit_xyz, dx, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hit_tri, dt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(settle_step, ds, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (dx) (void)hipFree(dx);
    if (dt) (void)hipFree(dt);
    if (ds) (void)hipFree(ds);
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("hit points: ") + hipGetErrorString(e));
    return START_OK;
}

int start_adjoint_push(start_ctx* c, const double* lambda, int32_t first_step, double* grad)
{
    if (!c || !lambda || !grad) return fail(c, START_ERR_ARG, "null argument");
    if (c->n == 0) {
        grad[0] = grad[1] = 0.0;
        return START_OK;
    }
    if (!c->d_settle) return fail(c, START_ERR_STATE, "settle records are off (start_record_settle)");
    if (c->prm.gas.model != START_MODEL_NONE)
        return fail(c, START_ERR_STATE, "the adjoint covers the collisionless push (gas model NONE)");
    const int64_t n = c->n;
    double *dl = nullptr, *dg = nullptr;
    hipError_t e = dalloc(&dl, 3 * n);
    if (e == hipSuccess) e = dalloc(&dg, 2);
    if (e == hipSuccess) e = hipMemcpyAsync(dl, lambda, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(dg, 0, sizeof(double) * 2, c->stream);
    if (e == hipSuccess) {
        if (c->mixed)
            launch_adjoint_push<float>(c->stream, c->mesh_view(), c->view<float>(), c->d_species, c->d_runs,
                                       c->prm.dt, first_step, c->d_settle, dl, dg, c->d_cnt);
        else
            launch_adjoint_push<double>(c->stream, c->mesh_view(), c->view<double>(), c->d_species, c->d_runs,
                                        c->prm.dt, first_step, c->d_settle, dl, dg, c->d_cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(grad, dg, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (dl) (void)hipFree(dl);
    if (dg) (void)hipFree(dg);
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("adjoint: ") + hipGetErrorString(e));
    return check_kernel_errors(c);
}

int start_gather_charge(start_ctx* c)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    if (c->n == 0) return START_OK;
    const int64_t CL = c->charge_len();
    double* cur = c->d_node_charge + (size_t)c->cb * CL;
    launch_zero_charge(c->stream, cur, c->d_node_charge + (size_t)(1 - c->cb) * CL, CL);
    if (c->mixed) launch_gather<float>(c->stream, c->mesh_view(), c->view<float>(), c->d_species, c->d_runs, cur);
    else launch_gather<double>(c->stream, c->mesh_view(), c->view<double>(), c->d_species, c->d_runs, cur);
    HIPCHECK(c, hipGetLastError());
    c->latest = c->cb;
    c->cb ^= 1;
    return START_OK;
}

int start_read_node_charge_runs(start_ctx* c, double* out)
{
    if (!c || !out) return fail(c, START_ERR_ARG, "null argument");
    HIPCHECK(c, hipMemcpyAsync(out, c->d_node_charge + (size_t)c->latest * c->charge_len(),
                               sizeof(double) * c->charge_len(), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return check_kernel_errors(c);
}

int start_read_node_charge(start_ctx* c, double* out)
{
    if (!c || !out) return fail(c, START_ERR_ARG, "null argument");
    HIPCHECK(c, hipMemcpyAsync(out, c->d_node_charge + (size_t)c->latest * c->charge_len(),
                               sizeof(double) * c->hm.n_nodes, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return check_kernel_errors(c);
}

}  // extern "C"

namespace {

// the node densities of the latest step into the context's pinned host
// buffers h_rho (and h_written when asked), synchronously
int node_density_to_host(start_ctx* c, int32_t mode, bool want_written)
{
    if (mode != START_DENSITY_BARYCENTRIC && mode != START_DENSITY_TET_AVERAGE)
        return fail(c, START_ERR_ARG, "unknown density mode");
    if (c->n_runs() != 1) return fail(c, START_ERR_STATE, "node charge density is defined for a single run");
    int rc = ensure_density_tables(c);
    if (rc != START_OK) return rc;
    const int64_t M = c->hm.n_nodes, T = c->hm.n_tets;
    if (!c->d_rho) {
        hipError_t e = dalloc(&c->d_rho, M);
        if (e == hipSuccess) e = dalloc(&c->d_written, M);
        if (e == hipSuccess) e = host_pinned_malloc((void**)&c->h_rho, sizeof(double) * (size_t)M);
        if (e == hipSuccess)
            e = host_pinned_malloc((void**)&c->h_written, sizeof(int32_t) * (size_t)M);
        if (e != hipSuccess)
            return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                        std::string("node charge density buffers: ") + hipGetErrorString(e));
    }
    if (mode == START_DENSITY_TET_AVERAGE) {
        HIPCHECK(c, hipMemsetAsync(c->d_tet_q, 0, sizeof(double) * T, c->stream));
        HIPCHECK(c, hipMemsetAsync(c->d_tet_n, 0, sizeof(int32_t) * T, c->stream));
    }
    const double* q = c->d_node_charge + (size_t)c->latest * M;
    if (c->mixed)
        launch_node_density<float>(c->stream, c->view<float>(), c->d_species, mode, q, c->d_node_vol,
                                   c->d_adj_start, c->d_adj_list, c->d_tet_vol, c->d_tet_q, c->d_tet_n, M, c->d_rho,
                                   c->d_written);
    else
        launch_node_density<double>(c->stream, c->view<double>(), c->d_species, mode, q, c->d_node_vol,
                                    c->d_adj_start, c->d_adj_list, c->d_tet_vol, c->d_tet_q, c->d_tet_n, M, c->d_rho,
                                    c->d_written);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipMemcpyAsync(c->h_rho, c->d_rho, sizeof(double) * M, hipMemcpyDeviceToHost, c->stream));
    if (want_written)
        HIPCHECK(c, hipMemcpyAsync(c->h_written, c->d_written, sizeof(int32_t) * M, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return START_OK;
}

}  // namespace

extern "C" {

int start_node_charge_density(start_ctx* c, int32_t mode, double* rho, int32_t* written)
{
    if (!c || !rho) return fail(c, START_ERR_ARG, "null argument");
    const int rc = node_density_to_host(c, mode, written != nullptr);
    if (rc != START_OK) return rc;
    const size_t M = (size_t)c->hm.n_nodes;
    std::memcpy(rho, c->h_rho, sizeof(double) * M);
    if (written) std::memcpy(written, c->h_written, sizeof(int32_t) * M);
    return START_OK;
}

int start_node_charge_density_view(start_ctx* c, int32_t mode, const double** rho, const int32_t** written)
{
    if (!c || !rho) return fail(c, START_ERR_ARG, "null argument");
    const int rc = node_density_to_host(c, mode, written != nullptr);
    if (rc != START_OK) return rc;
    *rho = c->h_rho;
    if (written) *written = c->h_written;
    return START_OK;
}

int start_read_particles(start_ctx* c, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                         int32_t* tet, int32_t* status, int32_t* hit_tri)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    int64_t n = c->n;
    if (n == 0) return START_OK;
    hipStream_t s = c->stream;
    // device order differs from load order after a sort: bring back by load index
    std::vector<uint32_t> pidx((size_t)n);
    HIPCHECK(c, hipMemcpyAsync(pidx.data(), c->A.pidx, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
    std::vector<double> tmp((size_t)n);
    std::vector<int32_t> itmp((size_t)n);
    double* pos_src[3] = {nullptr, nullptr, nullptr};
    double* abs_buf = nullptr;
    if (c->mixed && (x || y || z)) {
        HIPCHECK(c, dalloc(&abs_buf, 3 * n));
        launch_abs_positions(s, c->mesh_view(), (float*)c->A.x, (float*)c->A.y, (float*)c->A.z, c->A.tet, n, abs_buf,
                             abs_buf + n, abs_buf + 2 * n);
        pos_src[0] = abs_buf;
        pos_src[1] = abs_buf + n;
        pos_src[2] = abs_buf + 2 * n;
    } else if (!c->mixed) {
        pos_src[0] = (double*)c->A.x;
        pos_src[1] = (double*)c->A.y;
        pos_src[2] = (double*)c->A.z;
    }
    double* ddst[6] = {x, y, z, vx, vy, vz};
    void* vsrc[3] = {c->A.vx, c->A.vy, c->A.vz};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 6 && e == hipSuccess; ++k) {
        if (!ddst[k]) continue;
        const void* src = k < 3 ? (const void*)pos_src[k] : vsrc[k - 3];
        e = hipMemcpyAsync(tmp.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess)
            for (int64_t j = 0; j < n; ++j) ddst[k][pidx[(size_t)j]] = tmp[(size_t)j];
    }
    if (abs_buf) (void)hipFree(abs_buf);
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("read particles: ") + hipGetErrorString(e));
    int32_t* isrc[3] = {c->A.tet, c->A.flags, c->A.hit};
    int32_t* idst[3] = {tet, status, hit_tri};
    for (int k = 0; k < 3; ++k) {
        if (!idst[k]) continue;
        HIPCHECK(c, hipMemcpyAsync(itmp.data(), isrc[k], sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
        HIPCHECK(c, hipStreamSynchronize(s));
        for (int64_t j = 0; j < n; ++j) {
            int32_t v = itmp[(size_t)j];
            if (k == 1) v = (v & kSettled) ? START_SETTLED : START_ACTIVE;
            idst[k][pidx[(size_t)j]] = v;
        }
    }
    HIPCHECK(c, hipStreamSynchronize(s));
    return check_kernel_errors(c);
}

int start_read_tri_counts(start_ctx* c, uint64_t* counts)
{
    if (!c || !counts) return fail(c, START_ERR_ARG, "null argument");
    if (c->hm.n_tris == 0) return START_OK;
    HIPCHECK(c, hipMemcpyAsync(counts, c->d_tri_count, sizeof(uint64_t) * c->hm.n_tris, hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return START_OK;
}

int start_read_tri_counts_runs(start_ctx* c, uint64_t* counts)
{
    if (!c || !counts) return fail(c, START_ERR_ARG, "null argument");
    if (c->hm.n_tris == 0) return START_OK;
    HIPCHECK(c, hipMemcpyAsync(counts, c->d_tri_count, sizeof(uint64_t) * c->n_runs() * c->hm.n_tris,
                               hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return START_OK;
}

int start_read_settled(start_ctx* c, uint64_t* ids, int64_t* n)
{
    if (!c || !n) return fail(c, START_ERR_ARG, "null argument");
    std::vector<int32_t> flags((size_t)c->n);
    std::vector<uint32_t> pidx((size_t)c->n);
    if (c->n) {
        HIPCHECK(c, hipMemcpyAsync(flags.data(), c->A.flags, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipMemcpyAsync(pidx.data(), c->A.pidx, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    std::vector<uint8_t> settled((size_t)c->n, 0);
    for (int64_t j = 0; j < c->n; ++j)
        if (flags[(size_t)j] & kSettled) settled[pidx[(size_t)j]] = 1;
    int64_t cap = *n, k = 0;
    for (int64_t i = 0; i < c->n; ++i)
        if (settled[(size_t)i]) {
            if (ids && k < cap) ids[k] = c->ids[(size_t)i];
            ++k;
        }
    *n = k;
    return START_OK;
}

int start_read_settled_since(start_ctx* c, int64_t first, uint64_t* ids, int64_t* n)
{
    if (!c || !n || first < 0) return fail(c, START_ERR_ARG, "bad argument");
    if (!c->h_cnt) {
        hipError_t e = host_pinned_malloc((void**)&c->h_cnt, sizeof(unsigned long long) * kCntTotal);
        if (e != hipSuccess)
            return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                        std::string("statistics buffer: ") + hipGetErrorString(e));
    }
    HIPCHECK(c, hipMemcpyAsync(c->h_cnt + kCntSettled, c->d_cnt + kCntSettled, sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    const int64_t total = (int64_t)c->h_cnt[kCntSettled], m = std::max<int64_t>(0, total - first);
    if (!ids) {
        *n = m;
        return START_OK;
    }
    if (*n < m) return fail(c, START_ERR_ARG, "capacity below the number of newly settled particles");
    *n = m;
    if (m == 0) return START_OK;
    if (m > c->h_settle_cap) {
        const int64_t cap = std::max<int64_t>(m, 4096);
        uint32_t* p = nullptr;
        hipError_t e = host_pinned_malloc((void**)&p, sizeof(uint32_t) * (size_t)cap);
        if (e != hipSuccess)
            return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                        std::string("settled-id staging: ") + hipGetErrorString(e));
        if (c->h_settle) (void)hipHostFree(c->h_settle);
        c->h_settle = p;
        c->h_settle_cap = cap;
    }
    HIPCHECK(c, hipMemcpyAsync(c->h_settle, c->d_settle_log + first, sizeof(uint32_t) * (size_t)m,
                               hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < m; ++k) ids[k] = c->ids[c->h_settle[k]];
    return START_OK;
}

int start_read_stats(start_ctx* c, start_stats* out)
{
    if (!c || !out) return fail(c, START_ERR_ARG, "null argument");
    if (!c->h_cnt) {  // page-locked, so that the copy is one DMA behind the step
        hipError_t e = host_pinned_malloc((void**)&c->h_cnt, sizeof(unsigned long long) * kCntTotal);
        if (e != hipSuccess)
            return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                        std::string("statistics buffer: ") + hipGetErrorString(e));
    }
    const unsigned long long* h = c->h_cnt;
    HIPCHECK(c, hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(unsigned long long) * kCntTotal, hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    memset(out, 0, sizeof(*out));
    out->step = c->step;
    out->processed = 0;
    for (int k = 0; k < kProcSlots; ++k) out->processed += h[kCntProcSlot0 + k];
    out->settled = h[kCntSettled];
    out->n_particles = (uint64_t)c->n;
    out->deferred = h[kCntDeferred];
    out->exhaustive = h[kCntExhaustive];
    out->stop = (c->n > 0 && h[kCntSettled] >= (unsigned long long)c->n) ? 1 : 0;
    if (h[kCntError] & 1ull) return fail(c, START_ERR_KERNEL, "exact predicate expansion capacity exceeded");
    return START_OK;
}

int start_read_stats_settled(start_ctx* c, start_stats* out, int64_t first, uint64_t* ids, int64_t* n)
{
    if (!c || !out || !n || first < 0 || *n < 0 || (*n > 0 && !ids)) return fail(c, START_ERR_ARG, "bad argument");
    // the counters and a window of the settle log after `first`, in one sync
    const int64_t log_len = c->n;
    const int64_t W = std::max<int64_t>(0, std::min<int64_t>({*n, (int64_t)1024, log_len - first}));
    if (!c->h_cnt) {
        hipError_t e = host_pinned_malloc((void**)&c->h_cnt, sizeof(unsigned long long) * kCntTotal);
        if (e != hipSuccess)
            return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                        std::string("statistics buffer: ") + hipGetErrorString(e));
    }
    if (std::max<int64_t>(W, 1) > c->h_settle_cap) {
        const int64_t cap = std::max<int64_t>(W, 4096);
        uint32_t* p = nullptr;
        hipError_t e = host_pinned_malloc((void**)&p, sizeof(uint32_t) * (size_t)cap);
        if (e != hipSuccess)
            return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                        std::string("settled-id staging: ") + hipGetErrorString(e));
        if (c->h_settle) (void)hipHostFree(c->h_settle);
        c->h_settle = p;
        c->h_settle_cap = cap;
    }
    HIPCHECK(c, hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(unsigned long long) * kCntTotal, hipMemcpyDeviceToHost,
                               c->stream));
    if (W > 0)
        HIPCHECK(c, hipMemcpyAsync(c->h_settle, c->d_settle_log + first, sizeof(uint32_t) * (size_t)W,
                                   hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    const unsigned long long* h = c->h_cnt;
    memset(out, 0, sizeof(*out));
    out->step = c->step;
    for (int k = 0; k < kProcSlots; ++k) out->processed += h[kCntProcSlot0 + k];
    out->settled = h[kCntSettled];
    out->n_particles = (uint64_t)c->n;
    out->deferred = h[kCntDeferred];
    out->exhaustive = h[kCntExhaustive];
    out->stop = (c->n > 0 && h[kCntSettled] >= (unsigned long long)c->n) ? 1 : 0;
    if (h[kCntError] & 1ull) return fail(c, START_ERR_KERNEL, "exact predicate expansion capacity exceeded");
    const int64_t m = std::max<int64_t>(0, (int64_t)h[kCntSettled] - first);
    if (m > *n) return fail(c, START_ERR_ARG, "capacity below the number of newly settled particles");
    *n = m;
    if (m > W) {  // more than the window: the rest in a second copy
        if (m > c->h_settle_cap) {
            uint32_t* p = nullptr;
            hipError_t e = host_pinned_malloc((void**)&p, sizeof(uint32_t) * (size_t)m);
            if (e != hipSuccess)
                return fail(c, e == hipErrorOutOfMemory ? START_ERR_NOMEM : START_ERR_HIP,
                            std::string("settled-id staging: ") + hipGetErrorString(e));
            if (c->h_settle) (void)hipHostFree(c->h_settle);
            c->h_settle = p;
            c->h_settle_cap = m;
            HIPCHECK(c, hipMemcpyAsync(c->h_settle, c->d_settle_log + first, sizeof(uint32_t) * (size_t)m,
                                       hipMemcpyDeviceToHost, c->stream));
        } else {
            HIPCHECK(c, hipMemcpyAsync(c->h_settle + W, c->d_settle_log + first + W, sizeof(uint32_t) * (size_t)(m - W),
                                       hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
    for (int64_t k = 0; k < m; ++k) ids[k] = c->ids[c->h_settle[k]];
    return START_OK;
}

int start_synchronize(start_ctx* c)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return START_OK;
}

int start_locate(start_ctx* c, int64_t n, const double* x, const double* y, const double* z, int32_t* out_tet)
{
    if (!c || n < 0 || (n > 0 && (!x || !y || !z || !out_tet))) return fail(c, START_ERR_ARG, "bad argument");
    if (n == 0) return START_OK;
    double *dx = nullptr, *dy = nullptr, *dz = nullptr;
    int32_t* dt = nullptr;
    hipError_t e = dalloc(&dx, n);
    if (e == hipSuccess) e = dalloc(&dy, n);
    if (e == hipSuccess) e = dalloc(&dz, n);
    if (e == hipSuccess) e = dalloc(&dt, n);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dy, y, sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dz, z, sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        launch_locate(c->stream, c->mesh_view(), dx, dy, dz, n, dt, nullptr, nullptr, nullptr, nullptr, c->d_cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out_tet, dt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dz);
    (void)hipFree(dt);
    if (e != hipSuccess) return fail(c, START_ERR_HIP, std::string("locate: ") + hipGetErrorString(e));
    return check_kernel_errors(c);
}

// start_locate with the points and owners staged in the file-scope `scratch`
// array of this context's device (address resolved by start_create): no device
// allocation.  Every slot the kernel reads (x, y, z of points 0..n-1) is
// written by this call first, so nothing needs clearing between calls.
int start_locate_noalloc(start_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                         int32_t* out_tet)
{
    if (!c || n < 0 || (n > 0 && (!x || !y || !z || !out_tet))) return fail(c, START_ERR_ARG, "bad argument");
    if (n > START_NOALLOC_LOCATE_MAX)
        return fail(c, START_ERR_ARG, "start_locate_noalloc takes at most " +
                                          std::to_string(START_NOALLOC_LOCATE_MAX) + " points per call, got " +
                                          std::to_string(n));
    if (n == 0) return START_OK;
    if (!c->scratch) return fail(c, START_ERR_STATE, "no staging array on this device");
    double* dx = (double*)c->scratch;
    double* dy = dx + START_NOALLOC_LOCATE_MAX;
    double* dz = dy + START_NOALLOC_LOCATE_MAX;
    int32_t* dt = (int32_t*)(dz + START_NOALLOC_LOCATE_MAX);
    const size_t b8 = sizeof(double) * (size_t)n;
    HIPCHECK(c, hipMemcpyAsync(dx, x, b8, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(dy, y, b8, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(dz, z, b8, hipMemcpyHostToDevice, c->stream));
    launch_locate(c->stream, c->mesh_view(), dx, dy, dz, n, dt, nullptr, nullptr, nullptr, nullptr, c->d_cnt);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipMemcpyAsync(out_tet, dt, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return check_kernel_errors(c);
}

int start_mesh_maps(const start_mesh_desc* mesh, int32_t* tets_out, double* geo, double* rho, double* tau,
                    float* geo32, double* tau32)
{
    if (!mesh || !geo || !tau) return fail(nullptr, START_ERR_ARG, "null argument");
    HostMesh hm;
    std::string e = prepare_mesh(hm, mesh->n_nodes, mesh->nodes, mesh->n_tets, mesh->tets, mesh->n_tris, mesh->tris,
                                 1.01);
    if (!e.empty()) return fail(nullptr, START_ERR_MESH, e);
    std::memcpy(geo, hm.geo.data(), sizeof(double) * hm.geo.size());
    if (tets_out) std::memcpy(tets_out, hm.tets.data(), sizeof(int32_t) * hm.tets.size());
    if (rho) std::memcpy(rho, hm.rho.data(), sizeof(double) * hm.rho.size());
    *tau = hm.tau;
    if (geo32 || tau32) {
        std::vector<float> g32;
        const float t32 = build_geo32(hm, g32);
        if (geo32) std::memcpy(geo32, g32.data(), sizeof(float) * g32.size());
        if (tau32) *tau32 = (double)t32;
    }
    return START_OK;
}

int start_orient3d_batch(int64_t n, const double* pts, int32_t* out_sign)
{
    if (n < 0 || (n > 0 && (!pts || !out_sign))) return fail(nullptr, START_ERR_ARG, "bad argument");
    if (n == 0) return START_OK;
    double* dp = nullptr;
    int32_t* ds = nullptr;
    hipError_t e = dalloc(&dp, 12 * n);
    if (e == hipSuccess) e = dalloc(&ds, n);
    if (e == hipSuccess) e = hipMemcpy(dp, pts, sizeof(double) * 12 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_orient_batch(nullptr, dp, n, ds);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out_sign, ds, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
    (void)hipFree(dp);
    (void)hipFree(ds);
    if (e != hipSuccess) return fail(nullptr, START_ERR_HIP, std::string("orient3d batch: ") + hipGetErrorString(e));
    return START_OK;
}

int start_math_batch(int32_t which, int64_t n, const double* x, const double* y, double* out, double* out2)
{
    if (which < 0 || which > 5) return fail(nullptr, START_ERR_ARG, "which must be in [0, 5]");
    if (n < 0 || (n > 0 && (!x || !y || !out || !out2))) return fail(nullptr, START_ERR_ARG, "bad argument");
    if (n == 0) return START_OK;
    double* d = nullptr;
    hipError_t e = dalloc(&d, 4 * n);
    if (e == hipSuccess) e = hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, y, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_math_batch(nullptr, which, d, d + n, n, d + 2 * n, d + 3 * n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out2, d + 3 * n, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(nullptr, START_ERR_HIP, std::string("math batch: ") + hipGetErrorString(e));
    return START_OK;
}

int start_timing_enable(start_ctx* c, int32_t on)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->timing = on != 0;
    c->ev_used = 0;
    return START_OK;
}

int start_timing_read_each(start_ctx* c, double* ms, int64_t* n)
{
    if (!c || !n) return fail(c, START_ERR_ARG, "null argument");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    const int64_t m = (int64_t)c->ev_used;
    if (ms) {
        if (*n < m) return fail(c, START_ERR_ARG, "capacity below the number of timed launches");
        for (int64_t k = 0; k < m; ++k) {
            float t = 0.0f;
            HIPCHECK(c, hipEventElapsedTime(&t, c->events[(size_t)k].a0, c->events[(size_t)k].a1));
            ms[k] = t;
        }
        c->ev_used = 0;
    }
    *n = m;
    return START_OK;
}

int start_timing_read(start_ctx* c, double* step_kernel_ms, uint64_t* launches)
{
    if (!c) return fail(nullptr, START_ERR_ARG, "null context");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    double main_ms = 0.0;
    for (size_t k = 0; k < c->ev_used; ++k) {
        float ms = 0.0f;
        HIPCHECK(c, hipEventElapsedTime(&ms, c->events[k].a0, c->events[k].a1));
        main_ms += ms;
    }
    if (step_kernel_ms) *step_kernel_ms = main_ms;
    if (launches) *launches = c->ev_used;
    c->ev_used = 0;
    return START_OK;
}

}  // extern "C"
