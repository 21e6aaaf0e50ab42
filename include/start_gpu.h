/*
 * start_gpu.h — C ABI of libstart_gpu, the MI355X (gfx950) implementation of
 * Start's per-timestep particle kernel: electromagnetic push in the per-tet
 * field, tetrahedral relocation, HS/VHS/VSS gas collisions, surface hit and
 * settle, and barycentric charge scatter to the mesh nodes.
 *
 * Drop-in boundary.  Start's step loop (ModelingMainDriver.cpp:132-220) calls
 *   NodeChargeDensityProcessor::gather  (NodeChargeDensityProcessor.hpp:39-46)
 *   ParticleDynamicsProcessor::process  (ParticleDynamicsProcessor.hpp:154-170)
 * once per step.  A C++ shim with those two static signatures forwards to the
 * entry points below; the library owns every device buffer between steps.
 *
 * Conventions: every entry point returns START_OK (0) or a negative status and
 * never throws; start_last_error() gives the text.  One host thread per
 * context; calls are stream ordered and only start_read_* and start_step's
 * statistics synchronise.  All arrays are plain host pointers with the sizes
 * stated; nothing here takes device memory.
 */
#ifndef START_GPU_H
#define START_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define START_OK 0
#define START_ERR_ARG -1      /* bad argument / size                     */
#define START_ERR_HIP -2      /* HIP runtime error (text in last_error)  */
#define START_ERR_MESH -3     /* degenerate or non-conforming mesh       */
#define START_ERR_STATE -4    /* call out of order (e.g. no particles)   */
#define START_ERR_KERNEL -5   /* a kernel reported an internal limit     */
#define START_ERR_NOMEM -6    /* device allocation failed                */

/* Collision model, chosen by name in the reference (CollisionModelFactory.cpp:23-36). */
#define START_MODEL_NONE 0
#define START_MODEL_HS 1
#define START_MODEL_VHS 2
#define START_MODEL_VSS 3

/* Arithmetic of the step. */
#define START_PREC_FP64 0  /* FP64 throughout                                         */
#define START_PREC_MIXED 1 /* FP32 positions (stochastically rounded), tables and push increments; FP64 velocities,
                             containment near faces, collision velocity and charge */

/* Particle status as read back. */
#define START_ACTIVE 0
#define START_SETTLED 1

typedef struct start_ctx start_ctx;

/* Mesh: the tetrahedra of VolumetricMesh (VolumetricMesh.cpp:35-121) and the
 * surface triangles of SurfaceMesh (GmshUtils.cpp:116-136).  Node ids are
 * 0-based; triangle id = row index.  Triangles must be faces of the tet mesh
 * (boundary or internal); boundary faces without a triangle let particles leave. */
typedef struct start_mesh_desc {
    int64_t n_nodes;
    const double* nodes; /* 3*n_nodes, xyz */
    int64_t n_tets;
    const int32_t* tets; /* 4*n_tets, any orientation */
    int64_t n_tris;
    const int32_t* tris; /* 3*n_tris */
} start_mesh_desc;

/* Per particle species (ParticlePropertiesManager.hpp:73-217, with overrides). */
typedef struct start_species_desc {
    double mass;     /* kg */
    double charge;   /* C  */
    double diameter; /* collision diameter r_p + r_t, m (HSModel.cpp:8-12, VHSModel.cpp:14) */
} start_species_desc;

/* Background gas (Constants.hpp:20-73; ω and α are the float-rounded values). */
typedef struct start_gas_desc {
    double mass;
    double omega;
    double alpha;
    double concentration; /* m^-3, see start_gas_concentration */
    int32_t model;        /* START_MODEL_* */
    int32_t pad;
} start_gas_desc;

typedef struct start_params {
    double dt;              /* "Time Step" (ConfigParser.cpp:37)                          */
    double bbox_expansion;  /* 1.01 (Constants.hpp:55)                                     */
    uint64_t seed;          /* variate stream key                                          */
    uint32_t run;           /* run index folded into the key (batched sweeps)              */
    int32_t em_push;        /* 1: push with the owning tet's E; 0: the reference's literal
                               behaviour, where the owner lookup never matches (no push)  */
    int32_t precision;      /* START_PREC_*                                               */
    int32_t charge_every_step; /* 1: node charge of the end-of-step positions each step   */
    int32_t n_species;
    int32_t steps_per_launch; /* >1 fuses that many steps into one launch (charge then
                                 accumulates over the fused steps); 0/1 = one step each  */
    const start_species_desc* species;
    start_gas_desc gas;
    int32_t sort_interval;  /* sort particles by tet (and settled last) every this many
                               steps; 0 = never                                        */
    int32_t reserved;
} start_params;

typedef struct start_stats {
    uint64_t step;           /* steps taken so far                                   */
    uint64_t processed;      /* particle-steps processed (not settled, not frozen)   */
    uint64_t settled;        /* settled particles                                    */
    uint64_t n_particles;
    uint64_t deferred;       /* particle-steps resolved by the exact relocation path */
    uint64_t exhaustive;     /* of those, resolved by the grid search                */
    int32_t stop;            /* settled >= n_particles (ParticleSettler.cpp:24-25)   */
    int32_t pad;
} start_stats;

/* util::calculateConcentration (Utilities.cpp:132-149): n = (P/(R*T))*N_A. */
double start_gas_concentration(double pressure_pa, double temperature_k);

/* Builds the device mesh (orientation, face adjacency, surface map, barycentric
 * maps, bucket grids).  Replaces the FEMInitializer/CubicGrid/SurfaceMesh
 * set-up the step loop relies on (ModelingMainDriver.cpp:111-131). */
int start_create(start_ctx** out, const start_mesh_desc* mesh, const start_params* params);
void start_destroy(start_ctx* ctx);
/* Text of the last failure on ctx (or of the last failed start_create when ctx is NULL). */
const char* start_last_error(const start_ctx* ctx);

/* Uploads n particles (Particle state, Particle.hpp:25-30) and locates their
 * owning tets (min-id closed containment).  ids are the caller's particle ids,
 * kept for start_read_settled; variates are keyed by the load index. */
int start_load_particles(start_ctx* ctx, int64_t n, const double* x, const double* y,
                         const double* z, const double* vx, const double* vy, const double* vz,
                         const int32_t* species, const uint64_t* ids);

/* Appends m particles after the loaded ones: _spawnParticles inserts newly
 * generated particles at the end of the ParticleVector every step in
 * continuous-sputtering mode (ModelingMainDriver.cpp:37-71, :141-142).
 * Everything already loaded is kept: particle state, settled flags and hit
 * triangles, triangle counters, node charge, statistics, the step counter, and
 * the movement and settle records.  The new particles take load indices
 * n..n+m-1 (their ids default to those indices), so their variate streams are
 * new and the loaded particles' streams do not change.  They are located like
 * start_load_particles and join the movement records while fewer particles
 * than requested are tracked.  The device buffers grow geometrically, so
 * appending a few particles per step copies the population only now and then.
 * A single simulation only: START_ERR_STATE for a batch. */
int start_append_particles(start_ctx* ctx, int64_t m, const double* x, const double* y,
                           const double* z, const double* vx, const double* vy, const double* vz,
                           const int32_t* species, const uint64_t* ids);

/* ---- particle sources sampled on the device ----
 * Replaces ParticleGeneratorHost::fromPointSource / fromSurfaceSource
 * (ParticleGeneratorHost.cpp:33-159), their CUDA twins
 * (ParticleGeneratorDevice.cu:56-304) and ParticleDisplacer
 * (ParticleDisplacer.cpp:48-134): the particles are drawn straight into the
 * context's arrays (no host round trip) and located, as by
 * start_load_particles.  Ids are 0..n-1 in source order.
 *
 * Energies: START_ENERGY_MONO (every particle at energy_eV, the reference's
 * sources), START_ENERGY_MAXWELLIAN (velocity components normal with variance
 * kT/m, kT = p0 eV; the direction fields are not used),
 * START_ENERGY_THOMPSON (f(E) ~ E/(E+U)^3 on [0, E_max], U = p0, E_max = p1,
 * eV), START_ENERGY_LOGNORMAL (mean p0 eV, log-standard deviation p1).
 * Directions about the source direction d = (sin t cos p, sin t sin p, cos t)
 * (t = theta, p = phi; a surface cell's normal n for surface sources, with
 * t = acos(n_z/|n|), p = atan2(n_y, n_x) as ParticleGeneratorHost.cpp:139-146):
 * START_ANGLE_FAN (the reference: polar angle t + u*expansion with u uniform
 * in [-1, 1), azimuth p; ParticleDynamicUtils.hpp:52-96),
 * START_ANGLE_COSINE (Lambert law about d), START_ANGLE_ISOTROPIC,
 * START_ANGLE_CONE (uniform in solid angle within half-angle expansion of d).
 * displace: each particle starts at its source point + displace * d / max|d_i|
 * (ParticleDisplacer.cpp:20-31 uses 0.05 m; 0 leaves it on the source).
 * A surface source spreads count over its cells as the reference does
 * (count / n_cells each, the remainder to cells picked at random, seeded).
 * Variates: Philox4x32-10, counter {particle index, 1, source index, draw},
 * key {seed_lo, seed_hi}: disjoint from the step's variates. */
#define START_SOURCE_POINT 0
#define START_SOURCE_SURFACE 1
#define START_ENERGY_MONO 0
#define START_ENERGY_MAXWELLIAN 1
#define START_ENERGY_THOMPSON 2
#define START_ENERGY_LOGNORMAL 3
#define START_ANGLE_FAN 0
#define START_ANGLE_COSINE 1
#define START_ANGLE_ISOTROPIC 2
#define START_ANGLE_CONE 3
typedef struct start_source_desc {
    int32_t kind;          /* START_SOURCE_*                                        */
    int32_t species;       /* index into start_params.species                       */
    int64_t count;         /* particles, > 0 (ParticleGeneratorHost.cpp:55-59)      */
    int32_t energy;        /* START_ENERGY_*                                        */
    int32_t angle;         /* START_ANGLE_*                                         */
    double energy_eV;      /* START_ENERGY_MONO: the energy; != 0 (:61-66)          */
    double p0, p1;         /* distribution parameters (above)                       */
    double position[3];    /* point source                                          */
    double theta, phi;     /* point source direction, radians                       */
    double expansion;      /* fan / cone half-angle, radians                        */
    double displace;       /* start offset along d (m); 0 = none                    */
    const double* centres; /* surface source: 3 * n_cells cell centres              */
    const double* normals; /* surface source: 3 * n_cells outward normals (any length > 0) */
    int64_t n_cells;
} start_source_desc;
int start_generate(start_ctx* ctx, int32_t n_sources, const start_source_desc* sources, uint64_t seed);

/* ---- batched sweeps: independent runs on one mesh in the same launches ----
 * (no counterpart in the reference, which runs one simulation per process;
 * each run behaves exactly like a context of its own loaded with that run's
 * particles, field, concentration and variate key). */
typedef struct start_run_desc {
    int64_t n_particles;   /* rows of this run: the runs' rows follow each other in run order */
    double field_scale[2]; /* E of the run = s0 * E_0 + s1 * E_1 (field bases 0 and 1)       */
    double concentration;  /* background gas concentration of the run, m^-3                  */
    uint32_t run;          /* folded into the variate key                                     */
    int32_t pad;
} start_run_desc;

/* Uploads the particles of n_runs runs (sum of n_particles rows) and locates
 * them; replaces any particles loaded before.  Every run keeps its own node
 * charge and triangle counters. */
int start_load_batch(start_ctx* ctx, int32_t n_runs, const start_run_desc* runs, const double* x,
                     const double* y, const double* z, const double* vx, const double* vy,
                     const double* vz, const int32_t* species, const uint64_t* ids);
/* Field basis `which` (0 or 1): per-tet E or node potentials (E = -grad phi).
 * Basis 0 is the field start_set_tet_field / start_set_node_potential set. */
int start_set_tet_field_basis(start_ctx* ctx, int32_t which, const double* E);
int start_set_node_potential_basis(start_ctx* ctx, int32_t which, const double* phi);
int start_read_node_charge_runs(start_ctx* ctx, double* out /* n_runs*n_nodes, run-major */);
int start_read_tri_counts_runs(start_ctx* ctx, uint64_t* out /* n_runs*n_tris, run-major */);

/* ---- launch configuration and the autotuner ---- */
typedef struct start_launch_cfg {
    int32_t steps_per_launch; /* steps fused per launch, >= 1                                */
    int32_t block;            /* threads per workgroup: 128 or 256                            */
    int32_t persistent;       /* 1: a grid sized to the chip pulls particle tiles from a queue */
    int32_t sort_interval;    /* steps between sorts by tet; 0 = never                        */
    int32_t graph;            /* 1: the launches between two sorts are captured once in a
                                 hipGraph and replayed                                        */
    int32_t walk;             /* 1 (default): the fast path also certifies segments
                                 through several tets, so only edge/vertex passages go
                                 to the exact straggler kernel; implies block 256.
                                 0: single-crossing fast path.  Results are identical   */
} start_launch_cfg;
int start_set_launch(start_ctx* ctx, const start_launch_cfg* cfg);
int start_get_launch(start_ctx* ctx, start_launch_cfg* cfg);
/* Times each candidate configuration (steps per launch, block, persistent,
 * sort interval, then the multi-tet walk on and off for the best of those)
 * for trial_steps steps from a copy of the current state, restores the
 * state, and sets the fastest.  best and ms_per_step may be NULL. */
int start_autotune(start_ctx* ctx, int32_t trial_steps, start_launch_cfg* best, double* ms_per_step);
/* start_autotune with a cache file of results: a context of the same shape --
 * device (gcnArchName and CU count), precision, collision model, em_push,
 * charge every step, runs, tets, particle count to the next power of two, and
 * trial_steps -- takes the cached configuration without trials
 * (*from_cache = 1); otherwise the trials run and their result is appended to
 * the file (*from_cache = 0).  One line per result, "key|K block persistent
 * sort_interval walk|ms"; the last line of a key wins.  Callers that share a
 * file append without locking.  best, ms_per_step and from_cache may be NULL.
 * (No counterpart in the reference, which has no launch shapes.) */
int start_autotune_cached(start_ctx* ctx, const char* cache_path, int32_t trial_steps, start_launch_cfg* best,
                          double* ms_per_step, int32_t* from_cache);

/* ---- particle movements (the particleMovementMap argument of process) ----
 * Replaces ParticleMovementTracker::recordMovement
 * (ParticleMovementTracker.cpp:4-16) as process calls it: the position at the
 * start of every step a particle is pushed and moved in
 * (ParticleDynamicsProcessor.cpp:182-184), and the point where a settling
 * segment meets its triangle (ParticleSurfaceCollisionHandler.cpp:91-99,
 * :134-138), computed as in start_read_hits.
 *
 * start_track_movements(ctx, max_particles, max_points) starts recording for
 * the first max_particles particles, in load order, that the next step moves
 * (not settled, |v| >= 10 m/s, inside the expanded bounding box): the map
 * of the reference fills in its loop order while it holds fewer than
 * maxParticles (200,000 by default, ParticleMovementTracker.hpp:43), and no
 * particle starts moving later in the loop.  Each slot keeps its first
 * max_points points; later points are counted, not kept.  max_particles = 0
 * stops recording.  Loading or generating particles stops it too.
 * start_read_movements: with ids, counts and xyz all null, *n = the number of
 * tracked particles; otherwise (*n in = capacity) ids[s], counts[s] (points
 * recorded, possibly above max_points) and xyz[3 * (s * max_points + k)] for
 * slot s.  The JSON of saveMovementsToJson is written on the host from these
 * (start_write_movements_json, include/start_output.h).
 * start_clear_movements empties every slot and keeps the tracked particles:
 * a caller that appends each call's points to its own map (the
 * particleMovementMap of process) reads and clears after every step, with
 * max_points = 2 (a start and a hit point per step). */
int start_track_movements(start_ctx* ctx, int64_t max_particles, int32_t max_points);
int start_read_movements(start_ctx* ctx, int64_t* n, uint64_t* ids, int32_t* counts, double* xyz);
int start_clear_movements(start_ctx* ctx);

/* ---- adjoint of the collisionless push and the surface-hit map ----
 * (no counterpart in the reference; gives d(deposition moments)/d(electrode
 * potentials) for parameter optimisation).  The field of the run is
 * s0 * E_0 + s1 * E_1 with (s0, s1) the run's field_scale, so the electrode
 * potentials are the scales of two unit-potential field bases.
 *
 * start_record_settle(ctx, 1) makes the following steps record, per particle,
 * the step in which it settled and the start of that step's segment.
 * start_read_hits gives the hit point x_h = x_a + s (x_b - x_a) of each settled
 * particle (x_a, x_b the ends of the segment, s where it meets the plane of the
 * hit triangle), the triangle and the step; -1 for the others.
 * start_adjoint_push takes lambda = dL/dx_h per particle (3n, load order) and
 * returns dL/d(s0, s1) for steps first_step.. of a collisionless run, by a
 * reverse sweep that recomputes the particles' owners backwards in time. */
int start_record_settle(start_ctx* ctx, int32_t on);
int start_read_hits(start_ctx* ctx, double* hit_xyz /* 3n */, int32_t* hit_tri /* n */,
                    int32_t* settle_step /* n */);
int start_adjoint_push(start_ctx* ctx, const double* lambda /* 3n */, int32_t first_step,
                       double* grad /* 2 */);

/* Replaces the background gas (model, target, concentration) and, when
 * diameters is not NULL, the species' collision diameters (n_species values,
 * r_p + r_t depends on the target); recomputes the collision constants.
 * ParticleDynamicsProcessor::process takes the gas per call
 * (ParticleDynamicsProcessor.hpp:167-169). */
int start_set_gas(start_ctx* ctx, const start_gas_desc* gas, const double* diameters);
/* Replaces the time step (ParticleDynamicsProcessor.hpp:157, >= 0). */
int start_set_time_step(start_ctx* ctx, double dt);

/* Per-tet field E (3*n_tets), as ChargeDensityEquationSolver::solve leaves it
 * (ChargeDensityEquationSolver.cpp:42, VolumetricMesh.cpp:382-411). */
int start_set_tet_field(start_ctx* ctx, const double* E);
/* Node potentials (n_nodes) -> per-tet E = -grad(phi) on the device
 * (the FEM solve's hand-off, MatrixEquationSolver.cpp:53-64 then
 * VolumetricMesh.cpp:382-411).  Asynchronous and stream-ordered: the array is
 * copied into pinned staging before the call returns, steps already queued
 * keep the previous field and later steps see the new one, so a field refresh
 * every few steps costs no device synchronisation. */
int start_set_node_potential(start_ctx* ctx, const double* phi);
/* Reads back per-tet field basis `which` (3*n_tets, FP64) as the step kernels
 * see it, e.g. after start_set_node_potential: the field side of the hand-off
 * (VolumetricMesh.cpp:382-411 stores E per tet; no reference accessor). */
int start_read_tet_field(start_ctx* ctx, int32_t which, double* E);

/* ParticleDynamicsProcessor::process, nsteps times (ParticleDynamicsProcessor.cpp:213-277),
 * each step followed by the relocation and charge the next gather needs
 * (NodeChargeDensityProcessor.cpp:169-328).  Asynchronous. */
int start_step(start_ctx* ctx, int32_t nsteps);

/* NodeChargeDensityProcessor::gather for the current positions: barycentric
 * node charge Q_n = sum q_p*lambda_n(x_p) into the device node-charge buffer. */
int start_gather_charge(start_ctx* ctx);

int start_read_node_charge(start_ctx* ctx, double* out /* n_nodes */);

/* Node charge density (C/m^3) as NodeChargeDensityProcessor::gather leaves it
 * in NodeChargeDensitiesMap (NodeChargeDensityProcessor.cpp:60-93).
 *   START_DENSITY_BARYCENTRIC: Q_n / V_n, Q_n the latest barycentric node
 *     charge, V_n = (sum of the volumes of the tets around n) / 4; every node
 *     is written.
 *   START_DENSITY_TET_AVERAGE: the reference formula with one owner per
 *     particle at the current positions: rho_t = sum q / V_t over the tets
 *     owning at least one unsettled particle, rho_n = sum(rho_t V_t) / sum V_t
 *     over those tets around n (ascending tet id); nodes with no such tet are
 *     not written (written[n] = 0, rho[n] = 0), as the reference leaves their
 *     map entries untouched.
 * written may be NULL. */
#define START_DENSITY_BARYCENTRIC 0
#define START_DENSITY_TET_AVERAGE 1
int start_node_charge_density(start_ctx* ctx, int32_t mode, double* rho /* n_nodes */,
                              int32_t* written /* n_nodes or NULL */);
/* The same densities left in host buffers the context owns (page-locked):
 * *rho (and *written unless written is NULL) point to n_nodes values, valid
 * until the next density call on ctx or start_destroy.  Spares the caller a
 * copy when it reads them once, as the C++ gather does when it fills its map. */
int start_node_charge_density_view(start_ctx* ctx, int32_t mode, const double** rho, const int32_t** written);
int start_read_particles(start_ctx* ctx, double* x, double* y, double* z, double* vx, double* vy,
                         double* vz, int32_t* tet, int32_t* status, int32_t* hit_tri);
int start_read_tri_counts(start_ctx* ctx, uint64_t* counts /* n_tris */);
/* Settled ids (ParticlesIDSet, PICTypes.hpp:41) in load order; *n in = capacity, out = count. */
int start_read_settled(start_ctx* ctx, uint64_t* ids, int64_t* n);
/* The ids of the particles that settled after the first `first` to settle
 * (ParticlesIDSet inserts, ParticleSettler.cpp:7-26), in the order they
 * settled (within one step: device order): what a caller that has inserted
 * the first `first` still needs, with no read-back of the whole population.
 * ids NULL: *n = the number of such particles; otherwise *n in = capacity. */
int start_read_settled_since(start_ctx* ctx, int64_t first, uint64_t* ids, int64_t* n);
int start_read_stats(start_ctx* ctx, start_stats* out);
/* start_read_stats and start_read_settled_since(first) in one call, with one
 * device sync when at most 1,024 particles settled since `first` (a second
 * copy otherwise): what ParticleDynamicsProcessor::process reads after every
 * step (ParticleSettler.cpp:7-26, the stop test at :24-25).  *n in = the
 * capacity of ids, out = the number written. */
int start_read_stats_settled(start_ctx* ctx, start_stats* out, int64_t first, uint64_t* ids, int64_t* n);
int start_synchronize(start_ctx* ctx);

/* Device point location of n points (owner rule of the step), for checks. */
int start_locate(start_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                 int32_t* out_tet);
/* start_locate without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (allocation-free
 * injection or serving loops that locate a bounded batch of points each step,
 * harnesses that forbid hipMalloc).  The points and owners are staged in one
 * static device array per GPU of START_NOALLOC_LOCATE_MAX points (3.5 MiB),
 * held until the process ends; n above that returns START_ERR_ARG.  Calls
 * on one GPU share that array: calls that overlap on one GPU (other threads or
 * contexts on the same device) race. */
#define START_NOALLOC_LOCATE_MAX 131072
int start_locate_noalloc(start_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                         int32_t* out_tet);
/* The barycentric maps and certification thresholds start_create builds, on
 * the host (no device needed), for checks: tets_out (4 per tet, may be NULL)
 * the positively oriented connectivity, geo 12 doubles per tet (v0, then the
 * map A row-major: lambda_{1,2,3} = A (x - v0)), rho per tet (may be NULL) a
 * bound on the largest row sum of |A E - I| (E the exact edge matrix), tau
 * the FP64 threshold; geo32 (12 floats per tet: A, then v0) and tau32 the
 * mixed path's, may be NULL.  An infinite threshold certifies nothing. */
int start_mesh_maps(const start_mesh_desc* mesh, int32_t* tets_out, double* geo, double* rho,
                    double* tau, float* geo32, double* tau32);
/* Device exact orientation signs of n quadruples (12 doubles each), for checks. */
int start_orient3d_batch(int64_t n, const double* pts, int32_t* out_sign);
/* The device's FP64 elementary functions of the collision path and the
 * particle sources, over n arguments (host arrays), for checks against the
 * oracle's: which = 0 pow(x, y), 1 sin x (out) and cos x (out2), 2 acos x,
 * 3 atan2(y, x), 4 log x, 5 exp x.  out2 is written (zeros) for every which.
 * No counterpart in the reference, which calls the C library's functions
 * (HSModel.cpp:12-34, VSSModel.cpp:32-46, GeometryVector.hpp:448-474,
 * ParticleGeneratorHost.cpp:130-146). */
int start_math_batch(int32_t which, int64_t n, const double* x, const double* y, double* out, double* out2);

/* Timing hooks: HIP events recorded around the step kernels on the library's
 * stream; returns the summed kernel milliseconds since the last reset. */
int start_timing_enable(start_ctx* ctx, int32_t on);
int start_timing_read(start_ctx* ctx, double* step_kernel_ms, uint64_t* launches);
/* The same, launch by launch: with ms NULL, *n = the number of launches timed;
 * otherwise (*n in = capacity) ms[k] = the streaming kernel's milliseconds in
 * launch k, and the timings are reset. */
int start_timing_read_each(start_ctx* ctx, double* ms, int64_t* n);

#ifdef __cplusplus
}
#endif
#endif
