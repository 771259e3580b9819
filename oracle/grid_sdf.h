// Test infrastructure only (the checker, never the product): FP64 restatement of the grid signed
// distance fields the reference takes from Discregrid (github.com/Q-Minh/Discregrid @ master,
// fetched at configure time by CMakeLists.txt:95-102 and absent here):
//   * Discregrid::CubicLagrangeDiscreteGrid (Koschier, Deul, Bender 2017, "Hierarchical hp-Adaptive
//     Signed Distance Fields"): a box domain split into res[0] x res[1] x res[2] cells; every cell is
//     a 32-node cubic serendipity Lagrange element with nodes at its 8 corners and at 1/3 and 2/3
//     of its 12 edges; nodes are shared between cells. Values and gradients interpolate the node
//     values with the element's shape functions; outside the domain the value is DBL_MAX.
//     Call sites: sdf_model.cpp:66-75 (evaluate), environment_body.cpp:72-73 (build).
//   * Discregrid::MeshDistance::signedDistanceCached: the distance to the closest point of a
//     triangle mesh, negative where (x - closest point) points against the angle-weighted
//     pseudonormal of the closest feature (Baerentzen & Aanaes 2005, "Signed distance computation
//     using the angle weighted pseudonormal"). Call site: environment_body.cpp:66-70.
//   * environment_body_t's domain (environment_body.cpp:52-64), restated with its loop as written:
//     the box is grown by 1e-3 of its diagonal once per mesh vertex.
// The exact bits of Discregrid's arithmetic are not available here (parity unpinned against it);
// the tests pin this restatement by the element's interpolation properties (Kronecker delta at the
// nodes, partition of unity, exact reproduction of cubic polynomials), by brute-force distances and
// winding-number signs, and by the analytic sphere.
#pragma once
#include <cstdint>
#include <vector>

#include "linalg.h"

namespace oracle {

struct grid_sdf_t {
    vec3 dmin, dmax;  // domain
    uint32_t res[3] = {0, 0, 0};
    vec3 cell, inv_cell;
    std::vector<double> nodes;

    void init(vec3 lo, vec3 hi, const uint32_t r[3]);
    uint64_t node_count() const;
    vec3 node_position(uint64_t l) const;
    // value (DBL_MAX outside the domain) and gradient
    double interpolate(vec3 x, vec3* grad) const;
    // the 32 node indices of cell (i, j, k) in shape-function order
    void cell_nodes(uint32_t i, uint32_t j, uint32_t k, uint64_t out[32]) const;
};

// Shape functions of the 32-node element at local coordinates xi in [-1, 1]^3 and their
// derivatives (dN[j][d] = dN_j / dxi_d).
void shape_functions(vec3 xi, double N[32], double dN[32][3]);

// Signed distance of x to a closed or open triangle mesh (pseudonormal sign).
struct mesh_distance_t {
    std::vector<vec3> x;
    std::vector<uint32_t> tri;        // 3 per face
    std::vector<vec3> face_normal;    // unit
    std::vector<vec3> edge_normal;    // 3 per face: edges (0,1), (1,2), (2,0)
    std::vector<vec3> vertex_normal;  // angle weighted
    mesh_distance_t(const double* pos, uint32_t nv, const uint32_t* tris, uint32_t nt);
    double signed_distance(vec3 p) const;
};

// Closest point of triangle (a, b, c) to p; feature: 0-2 vertex a/b/c, 3-5 edge ab/bc/ca, 6 face.
// bary (optional): the weights of a, b, c that give the closest point.
vec3 closest_point_on_triangle(vec3 p, vec3 a, vec3 b, vec3 c, int* feature, double* bary = nullptr);

void environment_domain(const double* pos, uint32_t nv, vec3 lo, vec3 hi, vec3* out_lo, vec3* out_hi);

}  // namespace oracle
