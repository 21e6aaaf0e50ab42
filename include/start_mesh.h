/* Gmsh 4.1 mesh reader (host side, libstart_io.so; no GPU needed).
 *
 * Replaces the reference's libgmsh reads: VolumetricMesh::_readMesh
 * (VolumetricMesh.cpp:35-121: every 4-node tetrahedron, element type 4, with
 * its element tag as the tet id and its node tags) and
 * GmshUtils::getTriangleCellsMap (GmshUtils.cpp:96-152: every 3-node
 * triangle, element type 2, keyed by its element tag).  ASCII and binary
 * 4.1 files; other element types are skipped.
 *
 * Rows: nodes, tets and triangles each in ascending tag order (the device
 * breaks owner and hit ties by the smaller row, the reference by the smaller
 * element tag); element node references as node rows (0-based), element tags
 * alongside.  A repeated tag is an error.  The arrays are allocated with
 * malloc and released by start_gmsh_free. */
#ifndef START_MESH_H
#define START_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct start_gmsh_mesh {
    int64_t n_nodes;
    double* nodes;       /* 3 * n_nodes                       */
    uint64_t* node_tags; /* n_nodes, ascending                 */
    int64_t n_tets;
    int32_t* tets;       /* 4 * n_tets node rows               */
    uint64_t* tet_tags;  /* n_tets, ascending (rows sorted by tag) */
    int64_t n_tris;
    int32_t* tris;       /* 3 * n_tris node rows               */
    uint64_t* tri_tags;  /* n_tris, ascending: the reference's triangle ids */
} start_gmsh_mesh;

/* 0 on success; -1 with the reason in start_gmsh_last_error() otherwise. */
int start_gmsh_read(const char* path, start_gmsh_mesh* out);
void start_gmsh_free(start_gmsh_mesh* mesh);
const char* start_gmsh_last_error(void);

/* Writes a mesh as a Gmsh 4.1 ASCII file that Gmsh and the reference's
 * libgmsh reads (VolumetricMesh.cpp:35-121, GmshUtils.cpp:96-152) open: one
 * volume entity with every node and the tetrahedra, one surface entity with
 * the triangles, coordinates printed to 17 significant digits (read back
 * exactly).  Null tag arrays mean nodes 1..N, triangles 1..S and tets
 * S+1..S+T.  0 on success; -1 with the reason in start_gmsh_write_error(). */
int start_gmsh_write(const char* path, const start_gmsh_mesh* mesh);
const char* start_gmsh_write_error(void);

#ifdef __cplusplus
}
#endif

#endif
