/* Race check of the oracle's parallel step (test infrastructure): built with
 * -fsanitize=thread by `make -C oracle tsan` and run by tests/test_oracle_tsan.py.
 * A Kuhn box of n^3 cells (6 tets each), surface triangles on its bottom and
 * top faces, charged particles from a fixed LCG with HS collisions at a high
 * gas density, stepped with several threads: the per-particle loop, the
 * collision and hit paths, the triangle counters and the node charge scatter
 * all run concurrently.  Prints the step totals; ThreadSanitizer reports any
 * race on stderr. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "start_oracle.h"

static uint64_t lcg = 0x9E3779B97F4A7C15ull;
static double uni(void)
{
    lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(lcg >> 11) * (1.0 / 9007199254740992.0);
}

int main(int argc, char** argv)
{
    const int n = 6, nthreads = argc > 1 ? atoi(argv[1]) : 4, steps = 25;
    const int64_t np = 4000;
    const double L = 0.06, h = L / n;
    const int64_t nn = (int64_t)(n + 1) * (n + 1) * (n + 1);
    double* nodes = malloc(sizeof(double) * 3 * nn);
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                const int64_t id = ((int64_t)k * (n + 1) + j) * (n + 1) + i;
                nodes[3 * id] = i * h;
                nodes[3 * id + 1] = j * h;
                nodes[3 * id + 2] = k * h;
            }
#define NID(i, j, k) ((int32_t)((((k) * (n + 1) + (j)) * (n + 1)) + (i)))
    /* Kuhn: the six tets of a cube along its main diagonal v0 -> v7 */
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const int64_t nt = (int64_t)6 * n * n * n;
    int32_t* tets = malloc(sizeof(int32_t) * 4 * nt);
    int64_t t = 0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int p = 0; p < 6; ++p) {
                    int c[3] = {i, j, k};
                    tets[4 * t] = NID(c[0], c[1], c[2]);
                    for (int s = 0; s < 3; ++s) {
                        ++c[perm[p][s]];
                        tets[4 * t + 1 + s] = NID(c[0], c[1], c[2]);
                    }
                    ++t;
                }
    /* bottom and top faces: two triangles per boundary square */
    const int64_t ns = (int64_t)4 * n * n;
    int32_t* tris = malloc(sizeof(int32_t) * 3 * ns);
    int64_t s = 0;
    for (int z = 0; z <= n; z += n)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const int32_t a = NID(i, j, z), b = NID(i + 1, j, z), c = NID(i + 1, j + 1, z), d = NID(i, j + 1, z);
                int32_t* q = tris + 3 * s;
                q[0] = a, q[1] = b, q[2] = c, q[3] = a, q[4] = c, q[5] = d;
                s += 2;
            }
    or_mesh* m = or_mesh_build(nn, nodes, nt, tets, ns, tris);
    if (!m) {
        fprintf(stderr, "mesh build failed\n");
        return 1;
    }
    double* phi = malloc(sizeof(double) * nn);
    for (int64_t v = 0; v < nn; ++v) phi[v] = 50.0 * nodes[3 * v + 2] / L;
    double* E = malloc(sizeof(double) * 3 * nt);
    or_tet_field(m, phi, E);

    or_species_desc sd = {6.6335209e-26, 1.602176565e-19, 3.0e-10};
    or_gas_desc gd = {6.6335209e-26, 0.81f, 1.40f, 0.0, OR_MODEL_HS, 0};
    gd.concentration = 2e22;
    or_species_k sk;
    or_species_constants(&sd, &gd, &sk);
    or_params prm = {0};
    prm.dt = 2e-7;
    prm.concentration = gd.concentration;
    or_surface_bbox(nn, nodes, ns, tris, 1.01, prm.bbox_lo, prm.bbox_hi);
    prm.seed = 77;
    prm.model = OR_MODEL_HS;
    prm.em_push = 1;
    prm.nspecies = 1;

    double *x = malloc(8 * np), *y = malloc(8 * np), *z = malloc(8 * np);
    double *vx = malloc(8 * np), *vy = malloc(8 * np), *vz = malloc(8 * np);
    int32_t *tet = malloc(4 * np), *status = calloc(np, 4), *hit = malloc(4 * np), *species = calloc(np, 4);
    uint64_t* pid = malloc(8 * np);
    for (int64_t i = 0; i < np; ++i) {
        x[i] = L * (0.05 + 0.9 * uni());
        y[i] = L * (0.05 + 0.9 * uni());
        z[i] = L * (0.05 + 0.9 * uni());
        vx[i] = 2000.0 * (uni() - 0.5);
        vy[i] = 2000.0 * (uni() - 0.5);
        vz[i] = 2000.0 * (uni() - 0.5);
        hit[i] = -1;
        pid[i] = (uint64_t)i;
    }
    or_locate_many(m, np, x, y, z, tet);
    uint64_t* tri_count = calloc(ns, 8);
    double* q = malloc(sizeof(double) * nn);
    uint64_t total = 0;
    for (int k = 0; k < steps; ++k) {
        for (int64_t v = 0; v < nn; ++v) q[v] = 0.0;
        uint64_t processed = 0;
        if (or_step(m, &prm, &sk, E, np, x, y, z, vx, vy, vz, tet, status, hit, species, pid, (uint32_t)k, tri_count,
                    q, nthreads, &processed) != 0) {
            fprintf(stderr, "step %d failed\n", k);
            return 1;
        }
        total += processed;
    }
    uint64_t settled = 0, hits = 0;
    for (int64_t i = 0; i < np; ++i) settled += status[i] == OR_SETTLED;
    for (int64_t j = 0; j < ns; ++j) hits += tri_count[j];
    double qs = 0.0;
    for (int64_t v = 0; v < nn; ++v) qs += q[v];
    printf("steps %d processed %llu settled %llu hits %llu charge %.6e\n", steps, (unsigned long long)total,
           (unsigned long long)settled, (unsigned long long)hits, qs);
    or_mesh_free(m);
    return settled == hits ? 0 : 2;
}
