/* Output files of a run (host side, libstart_io.so; no GPU needed).
 *
 * Written from what the step library reads back (start_read_movements,
 * start_read_tri_counts) and the mesh, in the formats the reference writes:
 *
 *   start_write_movements_json  ParticleMovementTracker::saveMovementsToJson
 *                               (ParticleMovementTracker.cpp:18-69): a JSON
 *                               object {"<id>": [{"x":..,"y":..,"z":..}, ...]}
 *                               of the particles with more than one point,
 *                               keys in string order, 4-space indentation.
 *   start_deposition_histogram  SputteringModel::_writeHistogramToFile
 *   start_write_histogram       (SputteringModel.cpp:183-231): triangles whose
 *                               centroid lies in the band x_min <= x <= x_max
 *                               and whose xy projection has at least 75% of its
 *                               area in the band contribute count * overlap to
 *                               the y bin floor(y / bin) * bin; lines "y total".
 *   start_write_kde             SputteringModel::_writeKDEToFile
 *                               (SputteringModel.cpp:233-252): a header line
 *                               and "x y count" of the centroid of every
 *                               triangle with count > 0.
 *
 * Triangles are given as 9 doubles each (x, y, z of their three vertices).
 * reference_centroid = 1 uses the centroid the reference computes,
 * (v0 + v1 + v0) / 3: TriangleCell::compute_centroid (TriangleCell.cpp:11-14)
 * passes vertex(3), which a CGAL triangle maps back to vertex(0).  0 uses the
 * true centroid (v0 + v1 + v2) / 3.
 *
 * Every function returns 0 on success and -1 with the reason in
 * start_output_last_error() otherwise. */
#ifndef START_OUTPUT_H
#define START_OUTPUT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Slot s holds counts[s] points (only the first max_points are stored) at
 * xyz[3 * (s * max_points + k)], as start_read_movements gives them.  With
 * n = 0 no file is written (the reference warns and returns). */
int start_write_movements_json(const char* path, int64_t n, const uint64_t* ids, const int32_t* counts,
                               int32_t max_points, const double* xyz);

/* Histogram bins in ascending y: *n_bins in = capacity (or 0 to query the
 * count), out = number of bins. */
int start_deposition_histogram(int64_t n_tris, const double* tri_xyz, const uint64_t* counts, double x_min,
                               double x_max, double bin, int32_t reference_centroid, int64_t* n_bins,
                               double* bin_y, double* bin_total);
int start_write_histogram(const char* path, int64_t n_tris, const double* tri_xyz, const uint64_t* counts,
                          double x_min, double x_max, double bin, int32_t reference_centroid);

/* Overlap ratio of one triangle's xy projection with the band (area in the
 * band / area; 0 for a degenerate projection): TriangleOverlapCalculator::
 * calculateOverlapRatio (TriangleOverlapCalculator.cpp:7-27). */
double start_band_overlap(const double* tri_xyz, double x_min, double x_max);

int start_write_kde(const char* path, int64_t n_tris, const double* tri_xyz, const uint64_t* counts,
                    int32_t reference_centroid);

const char* start_output_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
