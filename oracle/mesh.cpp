// TEST INFRASTRUCTURE ONLY (see linalg.h). Compiled with -ffp-contract=off so that positions
// are bit-identical to the product library's generator, which is compiled the same way.
#include "mesh.h"

#include <algorithm>
#include <array>
#include <map>

namespace oracle {

// get_simple_bar_model.cpp:6-122. Positions are floats there (geometry_t::positions); the
// integer lattice coordinates are exact in float and double alike.
tet_mesh bar5(uint32_t width, uint32_t height, uint32_t depth)
{
    tet_mesh m;
    m.positions.resize(std::size_t(width) * height * depth * 3);
    for (uint32_t i = 0; i < width; ++i)
        for (uint32_t j = 0; j < height; ++j)
            for (uint32_t k = 0; k < depth; ++k) {
                std::size_t const row = std::size_t(i) * height * depth + j * depth + k;
                m.positions[row * 3 + 0] = double(float(i));
                m.positions[row * 3 + 1] = double(float(j));
                m.positions[row * 3 + 2] = double(float(k));
            }
    if (width < 2 || height < 2 || depth < 2) return m;
    std::size_t const cells = std::size_t(width - 1) * (height - 1) * (depth - 1);
    m.tets.resize(cells * 5 * 4);
    auto id = [&](uint32_t i, uint32_t j, uint32_t k) { return uint32_t(i * height * depth + j * depth + k); };
    for (uint32_t i = 0; i + 1 < width; ++i)
        for (uint32_t j = 0; j + 1 < height; ++j)
            for (uint32_t k = 0; k + 1 < depth; ++k) {
                uint32_t const p0 = id(i, j, k), p1 = id(i + 1, j, k), p2 = id(i + 1, j + 1, k),
                               p3 = id(i, j + 1, k), p4 = id(i, j, k + 1), p5 = id(i + 1, j, k + 1),
                               p6 = id(i + 1, j + 1, k + 1), p7 = id(i, j + 1, k + 1);
                std::size_t const base =
                    (std::size_t(i) * (height - 1) * (depth - 1) + j * (depth - 1) + k) * 20;
                // odd cells (i+j+k odd) and even cells use mirrored 5-tet splits (:61-114)
                std::array<uint32_t, 20> const odd = {p1, p0, p5, p2, p5, p2, p7, p6, p7, p0,
                                                      p5, p4, p2, p0, p7, p3, p5, p0, p7, p2};
                std::array<uint32_t, 20> const even = {p3, p1, p4, p0, p6, p1, p3, p2, p4, p1,
                                                       p6, p5, p6, p3, p4, p7, p3, p1, p6, p4};
                auto const& src = ((i + j + k) % 2 == 1) ? odd : even;
                std::copy(src.begin(), src.end(), m.tets.begin() + base);
            }
    return m;
}

tet_mesh box6(uint32_t nx, uint32_t ny, uint32_t nz, double edge, const double origin[3])
{
    tet_mesh m;
    uint32_t const vx = nx + 1, vy = ny + 1, vz = nz + 1;
    m.positions.resize(std::size_t(vx) * vy * vz * 3);
    for (uint32_t i = 0; i < vx; ++i)
        for (uint32_t j = 0; j < vy; ++j)
            for (uint32_t k = 0; k < vz; ++k) {
                std::size_t const v = (std::size_t(i) * vy + j) * vz + k;
                m.positions[v * 3 + 0] = origin[0] + double(i) * edge;
                m.positions[v * 3 + 1] = origin[1] + double(j) * edge;
                m.positions[v * 3 + 2] = origin[2] + double(k) * edge;
            }
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    static const bool odd[6] = {false, true, true, false, false, true};
    m.tets.reserve(std::size_t(nx) * ny * nz * 24);
    for (uint32_t i = 0; i < nx; ++i)
        for (uint32_t j = 0; j < ny; ++j)
            for (uint32_t k = 0; k < nz; ++k)
                for (int p = 0; p < 6; ++p) {
                    uint32_t c[3] = {0, 0, 0};
                    uint32_t v[4];
                    v[0] = (i * vy + j) * vz + k;
                    for (int s = 0; s < 3; ++s) {
                        c[perms[p][s]] += 1;
                        v[s + 1] = ((i + c[0]) * vy + (j + c[1])) * vz + (k + c[2]);
                    }
                    if (odd[p]) std::swap(v[1], v[2]);
                    m.tets.insert(m.tets.end(), v, v + 4);
                }
    return m;
}

// Face ids follow first insertion over tets in order, faces in faces_copy order
// {v1v2v4, v2v3v4, v3v1v4, v1v3v2} (topology.cpp:335-342, 904-956). A face keeps the vertex
// order of the tet that inserted it (topology.cpp:706-749). Boundary faces are those with an
// incident-tet count other than 2 (tetrahedral_mesh_boundary.cpp:88); surface vertices are
// numbered in first-seen order over boundary faces in face-id order (:97-116).
boundary extract_boundary(uint32_t vertex_count, const uint32_t* tets, uint32_t tet_count)
{
    struct face {
        std::array<uint32_t, 3> v;
        uint32_t incident;
    };
    std::vector<face> faces;
    std::map<std::array<uint32_t, 3>, uint32_t> index;
    for (uint32_t t = 0; t < tet_count; ++t) {
        uint32_t const v1 = tets[4 * t], v2 = tets[4 * t + 1], v3 = tets[4 * t + 2], v4 = tets[4 * t + 3];
        std::array<std::array<uint32_t, 3>, 4> const fc = {
            std::array<uint32_t, 3>{v1, v2, v4}, {v2, v3, v4}, {v3, v1, v4}, {v1, v3, v2}};
        for (auto const& f : fc) {
            auto key = f;
            std::sort(key.begin(), key.end());
            auto it = index.find(key);
            if (it == index.end()) {
                index.emplace(key, uint32_t(faces.size()));
                faces.push_back({f, 1});
            } else {
                faces[it->second].incident += 1;
            }
        }
    }
    boundary b;
    std::vector<int64_t> to_surface(vertex_count, -1);
    for (uint32_t fi = 0; fi < faces.size(); ++fi) {
        if (faces[fi].incident == 2) continue;
        b.triangle_to_face.push_back(fi);
        for (int j = 0; j < 3; ++j) {
            uint32_t const vi = faces[fi].v[j];
            if (to_surface[vi] < 0) {
                to_surface[vi] = int64_t(b.surface_to_tet.size());
                b.surface_to_tet.push_back(vi);
            }
            b.triangles.push_back(uint32_t(to_surface[vi]));
        }
    }
    return b;
}

}  // namespace oracle
