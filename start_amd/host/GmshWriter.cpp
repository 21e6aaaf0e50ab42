// Gmsh 4.1 ASCII writer for the tet meshes this project generates (box and
// chamber meshes of start_amd/meshgen.py), so that they can be opened by
// Gmsh and by the reference's libgmsh reads (VolumetricMesh.cpp:35-121,
// GmshUtils.cpp:96-152).  One volume entity (tag 1) holds every node and the
// tetrahedra (type 4); one surface entity (tag 1) holds the triangles (type 2).
#include "start_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace {

thread_local std::string g_write_error;

int fail(const std::string& msg)
{
    g_write_error = msg;
    return -1;
}

struct File {
    FILE* f;
    explicit File(const char* path) : f(std::fopen(path, "w")) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

const char* start_gmsh_write_error(void) { return g_write_error.c_str(); }

int start_gmsh_write(const char* path, const start_gmsh_mesh* m)
{
    if (!path || !m || m->n_nodes <= 0 || !m->nodes || m->n_tets < 0 || m->n_tris < 0 ||
        (m->n_tets > 0 && !m->tets) || (m->n_tris > 0 && !m->tris))
        return fail("invalid mesh");
    for (int64_t i = 0; i < 4 * m->n_tets; ++i)
        if (m->tets[i] < 0 || m->tets[i] >= m->n_nodes) return fail("tet node row out of range");
    for (int64_t i = 0; i < 3 * m->n_tris; ++i)
        if (m->tris[i] < 0 || m->tris[i] >= m->n_nodes) return fail("triangle node row out of range");
    // tags: as given, else nodes 1..N, triangles 1..S, tets S+1..S+T
    auto node_tag = [&](int64_t r) { return m->node_tags ? m->node_tags[r] : (uint64_t)(r + 1); };
    auto tri_tag = [&](int64_t s) { return m->tri_tags ? m->tri_tags[s] : (uint64_t)(s + 1); };
    auto tet_tag = [&](int64_t t) { return m->tet_tags ? m->tet_tags[t] : (uint64_t)(m->n_tris + t + 1); };
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) lo[k] = hi[k] = m->nodes[k];
    for (int64_t r = 0; r < m->n_nodes; ++r)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], m->nodes[3 * r + k]);
            hi[k] = std::fmax(hi[k], m->nodes[3 * r + k]);
        }
    uint64_t nmin = node_tag(0), nmax = node_tag(0);
    for (int64_t r = 0; r < m->n_nodes; ++r) {
        nmin = std::min(nmin, node_tag(r));
        nmax = std::max(nmax, node_tag(r));
    }
    const int64_t ne = m->n_tets + m->n_tris;
    uint64_t emin = UINT64_MAX, emax = 0;
    for (int64_t s = 0; s < m->n_tris; ++s) {
        emin = std::min(emin, tri_tag(s));
        emax = std::max(emax, tri_tag(s));
    }
    for (int64_t t = 0; t < m->n_tets; ++t) {
        emin = std::min(emin, tet_tag(t));
        emax = std::max(emax, tet_tag(t));
    }
    if (ne == 0) emin = 0;
    File out(path);
    FILE* f = out.f;
    if (!f) return fail(std::string("cannot open ") + path + " for writing");
    std::fprintf(f, "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n");
    std::fprintf(f, "$Entities\n0 0 1 1\n");
    std::fprintf(f, "1 %.17g %.17g %.17g %.17g %.17g %.17g 0 0\n", lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
    std::fprintf(f, "1 %.17g %.17g %.17g %.17g %.17g %.17g 0 0\n", lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
    std::fprintf(f, "$EndEntities\n");
    std::fprintf(f, "$Nodes\n1 %lld %llu %llu\n3 1 0 %lld\n", (long long)m->n_nodes, (unsigned long long)nmin,
                 (unsigned long long)nmax, (long long)m->n_nodes);
    for (int64_t r = 0; r < m->n_nodes; ++r) std::fprintf(f, "%llu\n", (unsigned long long)node_tag(r));
    for (int64_t r = 0; r < m->n_nodes; ++r)
        std::fprintf(f, "%.17g %.17g %.17g\n", m->nodes[3 * r], m->nodes[3 * r + 1], m->nodes[3 * r + 2]);
    std::fprintf(f, "$EndNodes\n");
    const int blocks = (m->n_tris > 0) + (m->n_tets > 0);
    std::fprintf(f, "$Elements\n%d %lld %llu %llu\n", blocks, (long long)ne, (unsigned long long)emin,
                 (unsigned long long)emax);
    if (m->n_tris > 0) {
        std::fprintf(f, "2 1 2 %lld\n", (long long)m->n_tris);
        for (int64_t s = 0; s < m->n_tris; ++s) {
            const int32_t* v = m->tris + 3 * s;
            std::fprintf(f, "%llu %llu %llu %llu\n", (unsigned long long)tri_tag(s), (unsigned long long)node_tag(v[0]),
                         (unsigned long long)node_tag(v[1]), (unsigned long long)node_tag(v[2]));
        }
    }
    if (m->n_tets > 0) {
        std::fprintf(f, "3 1 4 %lld\n", (long long)m->n_tets);
        for (int64_t t = 0; t < m->n_tets; ++t) {
            const int32_t* v = m->tets + 4 * t;
            std::fprintf(f, "%llu %llu %llu %llu %llu\n", (unsigned long long)tet_tag(t),
                         (unsigned long long)node_tag(v[0]), (unsigned long long)node_tag(v[1]),
                         (unsigned long long)node_tag(v[2]), (unsigned long long)node_tag(v[3]));
        }
    }
    std::fprintf(f, "$EndElements\n");
    if (std::ferror(f)) return fail(std::string("write failed: ") + path);
    return 0;
}

}  // extern "C"
