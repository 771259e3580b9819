// Mesh generation and boundary extraction, FP64 oracle side.
//
// TEST INFRASTRUCTURE ONLY (see linalg.h). Restates:
//   * get_simple_bar_model          src/geometry/get_simple_bar_model.cpp:6-122
//   * tetrahedron_set_t face order   src/physics/topology.cpp:335-342 (faces_copy), :904-956
//   * extract_boundary_surface       src/physics/tetrahedral_mesh_boundary.cpp:65-120
// plus the 6-tet (Kuhn) box split the benchmark configurations use (no reference counterpart).
#pragma once
#include <cstdint>
#include <vector>

namespace oracle {

struct tet_mesh {
    std::vector<double> positions;  // 3 per vertex
    std::vector<uint32_t> tets;     // 4 per tet
};

// Vertex lattice of width*height*depth vertices at unit spacing, 5 tets per cell with the
// reference's parity-alternating split.
tet_mesh bar5(uint32_t width, uint32_t height, uint32_t depth);

// nx*ny*nz hex cells of edge `edge` starting at `origin`, each split into the 6 tets that share
// the cell's (0,0,0)-(1,1,1) diagonal. Vertex (i,j,k) has index (i*(ny+1)+j)*(nz+1)+k; tets are
// ordered by cell (i outer, k inner), then by axis permutation in lexicographic order. Every tet
// has negative signed volume det[p1-p4, p2-p4, p3-p4]/6, the same sign the bar generator yields.
tet_mesh box6(uint32_t nx, uint32_t ny, uint32_t nz, double edge, const double origin[3]);

struct boundary {
    std::vector<uint32_t> surface_to_tet;  // surface vertex -> tet-mesh vertex
    std::vector<uint32_t> triangles;       // 3 surface-vertex indices per boundary face
    std::vector<uint32_t> triangle_to_face;// boundary face -> face id in the tet mesh
};

boundary extract_boundary(uint32_t vertex_count, const uint32_t* tets, uint32_t tet_count);

}  // namespace oracle
