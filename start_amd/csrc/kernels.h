// Device data layout and kernel argument blocks of the particle step.
//
// Particles: structure of arrays in HBM.
//   x y z            FP64 (FP64 mode) or FP32 in the owner's frame (mixed mode)
//   vx vy vz         FP64
//   tet              int32  owning tet at the start of the next step (-1: none)
//   flags            int32  bit0 settled, bit1 position strictly inside `tet`,
//                           bits 8..15 species index, bits 16..31 run index
//   hit              int32  surface triangle the particle settled on (-1: none)
//   pidx             uint32 load index (variate counter)
// Tets: geo (12 doubles: v0, 3x3 barycentric map), E (3 doubles),
//   node ids (int4), face codes (int4: >=0 neighbour, -1 open, <=-2 triangle).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mesh_prep.h"
#include "physics.h"

namespace sg {

constexpr int32_t kSettled = 1;
constexpr int32_t kStrict = 2;
constexpr int kSpeciesShift = 8;
constexpr int kMaxSpecies = 32;
constexpr int kRunShift = 16;
constexpr int kMaxRuns = 1 << 15;

// device counters (unsigned long long)
enum Counter {
    kCntProcessed = 0,
    kCntSettled = 1,
    kCntDeferred = 2,
    kCntExhaustive = 3,
    kCntWork = 4,   // work-list length, slot 0 (slot 1 at kCntWork + 2)
    kCntError = 5,  // bit 0: expansion capacity
    kCntProcSlot0 = 8,  // processed particle-steps, spread over kProcSlots slots
    kCntStep = 72,      // first step of the launch, slot 0 (slot 1 at kCntStep + 1)
    kCntTile = 74,      // tile queue of the persistent launch, slot 0 (slot 1 at kCntTile + 1)
    kCntTotal = 80
};
constexpr int kProcSlots = 64;

// One independent run of a batch: its gas concentration, its field as a
// combination of the two field bases, its variate key and where its particles,
// node charge and triangle counters start.
struct RunK {
    double conc;
    double s0, s1;       // E = s0 * E_0 + s1 * E_1
    uint32_t key;        // folded into the variate key
    uint32_t first;      // load index of the run's first particle
    int64_t node_base;   // run * n_nodes
    int64_t tri_base;    // run * n_tris
};

struct WorkItem {
    double x0, y0, z0;  // absolute start of the deferred step
    int32_t i;          // particle slot
    int32_t t0;         // owner at the start of that step
    int32_t k;          // step index within the launch
    int32_t pad;
};

struct DevMesh {
    const double* nodes;   // 3M
    const double* geo;     // 12T
    const double* E;       // 3T (FP64), field basis 0
    const float* E32;      // 4T (mixed), field basis 0
    const double* E1;      // 3T, field basis 1 (null: one basis)
    const float* E1_32;    // 4T
    const float* geo32;    // 12T (mixed)
    // push increments h = (((E*q)/m)*dt)/2 of field basis 0 per charged
    // species slot and tet: FP64 3 per tet, slot-major; mixed float4 per tet
    const double* H;
    const float* H32;
    const int4* tnode;     // T
    const int4* tnbr;      // T
    const int32_t* tris;   // 3S
    const uint8_t* tri_degenerate;
    const int32_t* tgrid_start;
    const int32_t* tgrid_list;
    const int32_t* sgrid_start;
    const int32_t* sgrid_list;
    GridInfo grid;
    double node_lo[3], node_hi[3];  // bounding box of the nodes (every tet and triangle is inside)
    int64_t n_nodes, n_tets, n_tris;
    double tau;
    float tau32;
};

template <typename R>
struct DevParticles {
    R *x, *y, *z;
    double *vx, *vy, *vz;
    int32_t* tet;
    int32_t* flags;
    int32_t* hit;
    const uint32_t* pidx;
    int64_t n;
};

// Where and when a particle settled, indexed by load index (recorded on
// request, for the adjoint of the push and hit map): the absolute position at
// the start of the step whose segment met the surface, and that step.
struct SettleRec {
    double x0, y0, z0;
    int32_t step;  // -1: not settled
    int32_t pad;
};

// Shape of a step launch: workgroup size and tile scheduling.
struct StepLaunch {
    int block;       // 128 | 256 threads
    int persistent;  // 1: `grid` workgroups pull tiles from a queue
    int grid;
    int batch;       // 1: particles of several runs (per-lane run parameters)
    int walk;        // 1: certified multi-tet walk on the fast path (block 256 only)
};

constexpr int kMaxPushSlots = 4;

struct StepParams {
    double dt;
    double lo[3], hi[3];
    uint64_t seed;
    const RunK* runs;         // run table (one entry for a single simulation)
    int32_t model;
    int32_t em_push;
    int32_t charge;
    int32_t nspecies;
    const int64_t* n_active;  // live particles lead the arrays after a sort
    double* zero_charge;      // node-charge buffers of the next launch, cleared here
    int64_t charge_len;       // n_runs * n_nodes
    int32_t work_slot;        // slot of this launch's work-list, step and tile counters
    int32_t nsteps;           // steps fused in this launch (charge accumulates over them)
    SettleRec* settle;        // null, or one record per load index
    uint32_t* settle_log;     // load index of each particle in the order it settled
                              // (slot = the settled count before it; one per particle)
    // movement records (ParticleMovementTracker.cpp:4-16): null, or the slot
    // of each load index (-1: not tracked), points recorded per slot, and
    // trk_max absolute points per slot
    const int32_t* trk_slot;
    int32_t* trk_count;
    double* trk_xyz;
    int32_t trk_max;
};

}  // namespace sg
