/* Surface deposition counts as HDF5 (host side, libstart_h5.so; no GPU).
 *
 * The file the reference writes at the end of a run
 * (ModelingMainDriver::_updateSurfaceMesh, ModelingMainDriver.cpp:79-86:
 * <mesh file stem>.hdf5) through TriangleMeshHdf5Manager::saveMeshToHDF5
 * (TriangleMeshHdf5Manager.cpp:97-139): one group "Triangle_<id>" per surface
 * triangle, holding three one-dimensional datasets
 *
 *   Coordinates  9 x H5T_NATIVE_DOUBLE  x, y, z of vertex 0, 1, 2
 *   Area         1 x H5T_NATIVE_DOUBLE  the triangle's area
 *   Counter      1 x H5T_NATIVE_ULONG   particles settled on it (size_t)
 *
 * An existing file of that name is removed first (TriangleMeshHdf5Manager.cpp:41-50).
 * Triangles are the mesh's surface triangles: ids are their element tags
 * (TriangleCellMap keys), vertices index `nodes` (x, y, z per node), and
 * counts are what start_read_tri_counts gives for them.  The area is
 * sqrt(|(v1 - v0) x (v2 - v0)|^2 / 4), the expression of the area CGAL's
 * Compute_area_3 evaluates (TriangleCell.cpp:16-24).
 *
 * Every function returns 0 on success and -1 with the reason in
 * start_hdf5_last_error() otherwise (start_write_settled_hdf5: 1 for no
 * particles). */
#ifndef START_HDF5_H
#define START_HDF5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces TriangleMeshHdf5Manager(filename) + saveMeshToHDF5(triangleCells)
 * (TriangleMeshHdf5Manager.cpp:41-50, :97-139).  n_tris = 0 leaves an empty
 * file, as the reference does for an empty map. */
int start_write_triangle_hdf5(const char* path, int64_t n_tris, const int64_t* ids, const double* nodes,
                              const int32_t* tris, const uint64_t* counts);

/* Replaces readMeshFromHDF5 (TriangleMeshHdf5Manager.cpp:141-178): the groups
 * Triangle_<first_id> .. Triangle_<first_id + number of groups - 1>, in that
 * order.  *n in = capacity of coords (9 per triangle), areas and counts (or 0
 * to query the number), out = number of groups in the file. */
int start_read_triangle_hdf5(const char* path, int64_t first_id, int64_t* n, double* coords, double* areas,
                             uint64_t* counts);

/* Replaces SettledParticleHDF5Writer::saveParticlesToHDF5
 * (SettledParticleHDF5Writer.cpp:4-74, :76-114): one n x 3 H5T_NATIVE_DOUBLE
 * dataset `dataset` (NULL: "settled_particles") of positions, the file created
 * with H5F_ACC_TRUNC.  n = 0 writes no file and returns 1, as the reference
 * returns false with a warning. */
int start_write_settled_hdf5(const char* path, int64_t n, const double* xyz, const char* dataset);

/* Area of one triangle (9 doubles) as the writer computes it. */
double start_triangle_area(const double* tri_xyz);

const char* start_hdf5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
