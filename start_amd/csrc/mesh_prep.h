// Host-side mesh preparation: the data the step kernels read from HBM.
//
//  * tets re-oriented positively (containment is orientation dependent in the
//    reference, VolumetricMesh.cpp:23-33);
//  * face adjacency (face f is opposite local vertex f) with each face coded as
//    neighbour tet id (>= 0), surface triangle (-2 - tri id) or open (-1);
//  * per-tet barycentric map lambda_{1,2,3} = A (x - v0), lambda_0 = 1 - sum:
//    a long-double inverse refined once against a nearly exact residual and
//    rounded to double, plus a global certification threshold tau (the largest
//    over the tets, with a rigorous bound on each map's own error) such that
//    every computed lambda > tau proves the exact barycentric coordinate is
//    positive.  An ill-conditioned tet raises tau; nothing is rejected;
//  * uniform bucket grids of tets and of surface triangles (the role of
//    CubicGrid.cpp:10-57 and the SurfaceMesh AABB tree, SurfaceMesh.cpp:33-53);
//  * the 1.01-expanded surface bounding box (SurfaceMesh.cpp:92-134).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "predicates.h"

namespace sg {

// local vertices of face f (opposite vertex f); orient3d(face, v_f) > 0
constexpr int kFace[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};

struct GridInfo {
    double lo[3];
    double edge;
    int div[3];
    int pad;
};

__host__ __device__ inline int grid_coord(const GridInfo& g, int k, double x)
{
    double f = (x - g.lo[k]) / g.edge;
    if (!(f >= 0.0)) return 0;
    if (f >= (double)g.div[k]) return g.div[k] - 1;
    int i = (int)f;
    return i >= g.div[k] ? g.div[k] - 1 : i;
}

struct HostMesh {
    int64_t n_nodes = 0, n_tets = 0, n_tris = 0;
    std::vector<double> nodes;    // 3M
    std::vector<int32_t> tets;    // 4T, positively oriented
    std::vector<int32_t> tris;    // 3S
    std::vector<int32_t> nbr;     // 4T face codes
    std::vector<double> geo;      // 12T: v0 (3), A row-major (9)
    std::vector<double> rho;      // T: bound on the row sums of |A E - I| (map error)
    std::vector<uint8_t> tri_degenerate;
    double tau = 0.0;             // certification threshold of the map
    double node_lo[3]{}, node_hi[3]{};  // bounding box of the nodes
    GridInfo grid{};
    std::vector<int32_t> tgrid_start, tgrid_list;
    std::vector<int32_t> sgrid_start, sgrid_list;
    double bbox_lo[3]{}, bbox_hi[3]{};
};

// tau of one tet (see mesh_prep.hip): any computed lambda > tau at a point with
// sum |lambda| <= 2 proves the exact barycentric coordinate positive.
double tau_bound(double rowsum, double diam, double rho);

// Returns an empty string on success, else the reason.
std::string prepare_mesh(HostMesh& m, int64_t n_nodes, const double* nodes, int64_t n_tets,
                         const int32_t* tets, int64_t n_tris, const int32_t* tris, double bbox_expansion);

}  // namespace sg
