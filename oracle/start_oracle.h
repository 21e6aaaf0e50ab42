/*
 * start_oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the per-timestep particle step of ViNN280801/Start
 * (push, move, gas collision, surface hit/settle, relocation, charge scatter).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * the library built from this file, and only as the checker / the timed CPU
 * baseline.  The GPU product never links it.
 *
 * Every function cites the reference file:line it restates.  The parity pins
 * for each function are listed in the header comment of start_oracle.c.
 */
#ifndef START_ORACLE_H
#define START_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Collision model ids (CollisionModelFactory.cpp:23-36 dispatches by name). */
enum { OR_MODEL_NONE = 0, OR_MODEL_HS = 1, OR_MODEL_VHS = 2, OR_MODEL_VSS = 3 };

/* Particle status. */
enum { OR_ACTIVE = 0, OR_SETTLED = 1 };

/* Raw per-species description, as a caller supplies it. */
typedef struct or_species_desc {
    double mass;     /* kg                                   */
    double charge;   /* C                                    */
    double diameter; /* collision diameter r_p + r_t, m      */
} or_species_desc;

/* Background gas. */
typedef struct or_gas_desc {
    double mass;          /* kg                                     */
    double omega;         /* VHS/VSS viscosity index (float value)  */
    double alpha;         /* VSS deflection parameter (float value) */
    double concentration; /* m^-3                                   */
    int32_t model;        /* OR_MODEL_*                             */
    int32_t pad;
} or_gas_desc;

/* Derived per-species constants (what HSModel/VHSModel/VSSModel recompute
 * per call; here once).  10 doubles. */
typedef struct or_species_k {
    double mass, charge;
    double sigma_hs;  /* PI * pow(d, 2)                        HSModel.cpp:12        */
    double vhs_c;     /* pow(d_ref,2) / tgamma(5/2 - omega)    VHSModel.cpp:19       */
    double two_kt;    /* 2 * KT_reference                      VHSModel.cpp:20       */
    double m_r;       /* p_mass*t_mass/(p_mass+t_mass)         VHSModel.cpp:15       */
    double vhs_exp;   /* omega - 1/2                           VHSModel.cpp:16       */
    double mass_cp;   /* p_mass/(t_mass+p_mass)                HSModel.cpp:26        */
    double mass_ct;   /* t_mass/(t_mass+p_mass)                HSModel.cpp:27        */
    double inv_alpha; /* 1./alpha                              VSSModel.cpp:32       */
} or_species_k;

typedef struct or_params {
    double dt;
    double concentration;
    double bbox_lo[3], bbox_hi[3]; /* expanded surface bbox bounds */
    uint64_t seed;
    uint32_t run;
    int32_t model;
    int32_t em_push; /* 1: push with E of the owning tet; 0: literal reference (no push) */
    int32_t nspecies;
} or_params;

typedef struct or_mesh or_mesh;

/* ---- primitives (exposed for pinning tests) ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_variates(uint64_t seed, uint32_t run, uint64_t pid, uint32_t step, uint32_t draw, double u[2]);
int or_orient3d(const double* a, const double* b, const double* c, const double* d);
int or_orient3d_exact(const double* a, const double* b, const double* c, const double* d);
int or_orient2d_exact(const double* a, const double* b, const double* c);
int or_seg_tri_intersect(const double* p, const double* q, const double* a, const double* b,
                         const double* c, double* t_out);
double or_gas_concentration(double pressure, double temperature);
void or_species_constants(const or_species_desc* s, const or_gas_desc* g, or_species_k* out);
void or_surface_bbox(int64_t nnodes, const double* nodes, int64_t ntris, const int32_t* tris,
                     double expansion, double lo[3], double hi[3]);

/* single-particle collision, exposed for fixed-variate closed-form tests */
int or_collide(int model, const or_species_k* k, double concentration, double dt, double v[3],
               double u1, double u2, double u3);
void or_push(const double E[3], double q, double m, double dt, double v[3]);

/* Elementary functions of the collision path: 0 (default) the shared
 * deterministic ones the HIP kernels also compile (start_amd/csrc/detmath.h),
 * 1 the C library's, as the reference calls them.  Set before stepping. */
void or_set_libm(int on);
int or_get_libm(void);
/* which = 0 pow(x, y), 1 sin x (out) and cos x (out2), 2 acos x,
 * 3 atan2(y, x), 4 log x, 5 exp x, over n arguments, in the given mode. */
void or_math_eval(int which, int libm, int64_t n, const double* x, const double* y, double* out, double* out2);

/* ---- mesh ---- */
or_mesh* or_mesh_build(int64_t nnodes, const double* nodes, int64_t ntets, const int32_t* tets,
                       int64_t ntris, const int32_t* tris);
void or_mesh_free(or_mesh* m);
int64_t or_mesh_ntets(const or_mesh* m);
void or_mesh_tets(const or_mesh* m, int32_t* out); /* positively oriented connectivity */
int32_t or_locate(const or_mesh* m, const double* p);
void or_locate_many(const or_mesh* m, int64_t n, const double* x, const double* y, const double* z,
                    int32_t* out);
int32_t or_first_hit(const or_mesh* m, const double* p, const double* q);
void or_tet_field(const or_mesh* m, const double* phi, double* E);
void or_barycentric(const or_mesh* m, int32_t t, const double* p, double lam[4]);

/* ---- the step ---- */
int or_step(const or_mesh* m, const or_params* prm, const or_species_k* sp, const double* E,
            int64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
            int32_t* tet, int32_t* status, int32_t* hit, const int32_t* species,
            const uint64_t* pid, uint32_t step, uint64_t* tri_count, double* node_charge,
            int nthreads, uint64_t* processed);
/* or_step, and moved[i] = 1 for the particles whose position the step records
 * as a movement (pushed and moved this step), 0 for the others */
int or_step_moved(const or_mesh* m, const or_params* prm, const or_species_k* sp, const double* E,
                  int64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                  int32_t* tet, int32_t* status, int32_t* hit, const int32_t* species,
                  const uint64_t* pid, uint32_t step, uint64_t* tri_count, double* node_charge,
                  int nthreads, uint64_t* processed, uint8_t* moved);

void or_scatter_charge(const or_mesh* m, const or_species_k* sp, int64_t n, const double* x,
                       const double* y, const double* z, const int32_t* tet, const int32_t* status,
                       const int32_t* species, double* node_charge);

/* ---- node charge density (NodeChargeDensityProcessor.cpp:51-93) ---- */
void or_tet_volumes(const or_mesh* m, double* vol);
void or_node_volumes(const or_mesh* m, double* vol);
void or_node_density(const or_mesh* m, const or_species_k* sp, int64_t n, const int32_t* tet,
                     const int32_t* status, const int32_t* species, double* rho, int32_t* written);

#ifdef __cplusplus
}
#endif
#endif
