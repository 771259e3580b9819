// Host mesh utilities of the C ABI: lattice generators, boundary surface extraction and lumped
// masses. Flat arrays and sorting (no per-face node allocation), so that the 6M-tet beam is
// processed in about a second.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "abi_util.h"
#include "sbs_gpu.h"

namespace {

// Corner offsets (dx, dy, dz) of a lattice cell in the reference's cube numbering
// (get_simple_bar_model.cpp:34-41): 0 (0,0,0) 1 (1,0,0) 2 (1,1,0) 3 (0,1,0) 4 (0,0,1) 5 (1,0,1)
// 6 (1,1,1) 7 (0,1,1).
constexpr int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// The two mirrored 5-tet splits, as corner numbers (get_simple_bar_model.cpp:61-114).
constexpr int kSplitOdd[5][4] = {{1, 0, 5, 2}, {5, 2, 7, 6}, {7, 0, 5, 4}, {2, 0, 7, 3}, {5, 0, 7, 2}};
constexpr int kSplitEven[5][4] = {{3, 1, 4, 0}, {6, 1, 3, 2}, {4, 1, 6, 5}, {6, 3, 4, 7}, {3, 1, 6, 4}};

}  // namespace

extern "C" int sbs_mesh_bar5(uint32_t w, uint32_t h, uint32_t d, double* pos, uint32_t* tets)
{
    return sbs::guarded([&] {
        if (pos) {
            std::size_t v = 0;
            for (uint32_t i = 0; i < w; ++i)
                for (uint32_t j = 0; j < h; ++j)
                    for (uint32_t k = 0; k < d; ++k, ++v) {
                        pos[3 * v + 0] = double(float(i));
                        pos[3 * v + 1] = double(float(j));
                        pos[3 * v + 2] = double(float(k));
                    }
        }
        if (tets && w > 1 && h > 1 && d > 1) {
            uint32_t* out = tets;
            for (uint32_t i = 0; i + 1 < w; ++i)
                for (uint32_t j = 0; j + 1 < h; ++j)
                    for (uint32_t k = 0; k + 1 < d; ++k) {
                        auto const& split = ((i + j + k) & 1u) ? kSplitOdd : kSplitEven;
                        for (auto const& tet : split)
                            for (int c : tet)
                                *out++ = (i + kCorner[c][0]) * h * d + (j + kCorner[c][1]) * d + (k + kCorner[c][2]);
                    }
        }
    });
}

extern "C" int sbs_mesh_box6(uint32_t nx, uint32_t ny, uint32_t nz, double edge, const double origin[3], double* pos,
                             uint32_t* tets)
{
    return sbs::guarded([&] {
        sbs::require(origin != nullptr, "origin is null");
        uint32_t const vy = ny + 1, vz = nz + 1;
        if (pos) {
            std::size_t v = 0;
            for (uint32_t i = 0; i <= nx; ++i)
                for (uint32_t j = 0; j < vy; ++j)
                    for (uint32_t k = 0; k < vz; ++k, ++v) {
                        pos[3 * v + 0] = origin[0] + double(i) * edge;
                        pos[3 * v + 1] = origin[1] + double(j) * edge;
                        pos[3 * v + 2] = origin[2] + double(k) * edge;
                    }
        }
        if (tets) {
            // The 6 monotone lattice paths from corner (0,0,0) to (1,1,1), one per axis order; odd
            // axis orders list their 2nd and 3rd vertices swapped so every tet has the same
            // (negative) orientation.
            static const int order[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
            uint32_t* out = tets;
            for (uint32_t i = 0; i < nx; ++i)
                for (uint32_t j = 0; j < ny; ++j)
                    for (uint32_t k = 0; k < nz; ++k)
                        for (int p = 0; p < 6; ++p) {
                            uint32_t c[3] = {i, j, k};
                            uint32_t v[4];
                            v[0] = (c[0] * vy + c[1]) * vz + c[2];
                            for (int s = 0; s < 3; ++s) {
                                ++c[order[p][s]];
                                v[s + 1] = (c[0] * vy + c[1]) * vz + c[2];
                            }
                            bool const odd = (p == 1 || p == 2 || p == 5);
                            out[0] = v[0];
                            out[1] = odd ? v[2] : v[1];
                            out[2] = odd ? v[1] : v[2];
                            out[3] = v[3];
                            out += 4;
                        }
        }
    });
}

// Cells of an (nx, ny, nz) lattice of edge `edge` centred on the origin whose centres satisfy
// (x/rx)^2 + (y/ry)^2 + (z/rz)^2 <= 1, each split like sbs_mesh_box6; vertices are the used lattice
// corners, numbered in lattice order.
extern "C" int sbs_mesh_ellipsoid6(double rx, double ry, double rz, double edge, double* pos, uint32_t* tets,
                                   uint32_t* num_vertices, uint32_t* num_tets)
{
    return sbs::guarded([&] {
        sbs::require(rx > 0 && ry > 0 && rz > 0 && edge > 0, "radii and edge must be positive");
        sbs::require(num_vertices && num_tets, "size outputs are null");
        uint32_t const n[3] = {uint32_t(std::ceil(2 * rx / edge)), uint32_t(std::ceil(2 * ry / edge)),
                               uint32_t(std::ceil(2 * rz / edge))};
        double const r[3] = {rx, ry, rz};
        double o[3];
        for (int a = 0; a < 3; ++a) o[a] = -0.5 * double(n[a]) * edge;
        uint32_t const vy = n[1] + 1, vz = n[2] + 1;
        std::size_t const lattice = std::size_t(n[0] + 1) * vy * vz;
        auto inside = [&](uint32_t i, uint32_t j, uint32_t k) {
            double const c[3] = {o[0] + (i + 0.5) * edge, o[1] + (j + 0.5) * edge, o[2] + (k + 0.5) * edge};
            double s2 = 0;
            for (int a = 0; a < 3; ++a) s2 += (c[a] / r[a]) * (c[a] / r[a]);
            return s2 <= 1.0;
        };
        std::vector<uint32_t> id(lattice, UINT32_MAX);
        uint32_t ncells = 0;
        for (uint32_t i = 0; i < n[0]; ++i)
            for (uint32_t j = 0; j < n[1]; ++j)
                for (uint32_t k = 0; k < n[2]; ++k)
                    if (inside(i, j, k)) {
                        ++ncells;
                        for (int c = 0; c < 8; ++c) id[((i + (c & 1)) * vy + (j + ((c >> 1) & 1))) * vz + (k + (c >> 2))] = 0;
                    }
        uint32_t nv = 0;
        for (std::size_t v = 0; v < lattice; ++v)
            if (id[v] == 0) {
                id[v] = nv;
                if (pos) {
                    uint32_t const i = uint32_t(v / (std::size_t(vy) * vz)), j = uint32_t((v / vz) % vy), k = uint32_t(v % vz);
                    pos[3 * std::size_t(nv) + 0] = o[0] + double(i) * edge;
                    pos[3 * std::size_t(nv) + 1] = o[1] + double(j) * edge;
                    pos[3 * std::size_t(nv) + 2] = o[2] + double(k) * edge;
                }
                ++nv;
            }
        *num_vertices = nv;
        *num_tets = ncells * 6;
        if (!tets) return;
        static const int order[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        uint32_t* out = tets;
        // Cells are emitted by lattice parity class (i%2, j%2, k%2), lattice order inside a class.
        // Cells of one class share no vertex, so a greedy coloring of the cells in this order needs
        // exactly one color per class (8), balanced, whatever the ellipsoid cuts away.
        for (uint32_t cls = 0; cls < 8; ++cls)
        for (uint32_t i = cls & 1; i < n[0]; i += 2)
            for (uint32_t j = (cls >> 1) & 1; j < n[1]; j += 2)
                for (uint32_t k = cls >> 2; k < n[2]; k += 2) {
                    if (!inside(i, j, k)) continue;
                    for (int p = 0; p < 6; ++p) {
                        uint32_t c[3] = {i, j, k};
                        uint32_t v[4];
                        v[0] = id[(std::size_t(c[0]) * vy + c[1]) * vz + c[2]];
                        for (int s = 0; s < 3; ++s) {
                            ++c[order[p][s]];
                            v[s + 1] = id[(std::size_t(c[0]) * vy + c[1]) * vz + c[2]];
                        }
                        bool const odd = (p == 1 || p == 2 || p == 5);
                        out[0] = v[0];
                        out[1] = odd ? v[2] : v[1];
                        out[2] = odd ? v[1] : v[2];
                        out[3] = v[3];
                        out += 4;
                    }
                }
    });
}

extern "C" int sbs_mesh_boundary(uint32_t num_vertices, const uint32_t* tets, uint32_t num_tets,
                                 uint32_t* surface_to_tet, uint32_t* triangles, uint32_t* num_surface_vertices,
                                 uint32_t* num_triangles)
{
    return sbs::guarded([&] {
        sbs::require(tets != nullptr || num_tets == 0, "tets is null");
        sbs::require(num_surface_vertices && num_triangles, "size outputs are null");
        // Faces in insertion order: tet t, local face f at position 4t+f, vertex order as in
        // tetrahedron_t::faces_copy {v1v2v4, v2v3v4, v3v1v4, v1v3v2} (topology.cpp:335-342).
        static const int fv[4][3] = {{0, 1, 3}, {1, 2, 3}, {2, 0, 3}, {0, 2, 1}};
        std::size_t const nf = std::size_t(num_tets) * 4;
        struct entry {
            uint32_t a, b, c;  // sorted vertex ids
            uint64_t pos;      // insertion position
        };
        std::vector<entry> e(nf);
        for (std::size_t t = 0; t < num_tets; ++t)
            for (int f = 0; f < 4; ++f) {
                uint32_t v[3] = {tets[4 * t + fv[f][0]], tets[4 * t + fv[f][1]], tets[4 * t + fv[f][2]]};
                for (uint32_t x : v) sbs::require(x < num_vertices, "tet vertex index out of range");
                std::sort(v, v + 3);
                e[4 * t + f] = {v[0], v[1], v[2], 4 * t + f};
            }
        std::sort(e.begin(), e.end(), [](entry const& l, entry const& r) {
            if (l.a != r.a) return l.a < r.a;
            if (l.b != r.b) return l.b < r.b;
            if (l.c != r.c) return l.c < r.c;
            return l.pos < r.pos;
        });
        // Each distinct face: first insertion position and incident count. A face is on the
        // boundary unless exactly two tets share it (tetrahedral_mesh_boundary.cpp:88); face ids,
        // and hence the boundary order, follow first insertion.
        std::vector<uint64_t> boundary_first;
        for (std::size_t i = 0; i < nf;) {
            std::size_t j = i + 1;
            while (j < nf && e[j].a == e[i].a && e[j].b == e[i].b && e[j].c == e[i].c) ++j;
            if (j - i != 2) boundary_first.push_back(e[i].pos);
            i = j;
        }
        std::sort(boundary_first.begin(), boundary_first.end());
        std::vector<uint32_t> to_surface(num_vertices, UINT32_MAX);
        uint32_t ns = 0;
        std::size_t k = 0;
        for (uint64_t p : boundary_first) {
            uint64_t const t = p / 4;
            int const f = int(p % 4);
            for (int j = 0; j < 3; ++j) {
                uint32_t const vi = tets[4 * t + fv[f][j]];
                if (to_surface[vi] == UINT32_MAX) {
                    to_surface[vi] = ns;
                    if (surface_to_tet) surface_to_tet[ns] = vi;
                    ++ns;
                }
                if (triangles) triangles[k] = to_surface[vi];
                ++k;
            }
        }
        *num_surface_vertices = ns;
        *num_triangles = uint32_t(boundary_first.size());
    });
}

extern "C" int sbs_mesh_lumped_mass(uint32_t num_vertices, const double* pos, const uint32_t* tets, uint32_t num_tets,
                                    double density, double* mass)
{
    return sbs::guarded([&] {
        sbs::require(pos && mass && (tets || num_tets == 0), "null argument");
        std::fill(mass, mass + num_vertices, 0.0);
        for (std::size_t t = 0; t < num_tets; ++t) {
            uint32_t const* v = tets + 4 * t;
            for (int k = 0; k < 4; ++k) sbs::require(v[k] < num_vertices, "tet vertex index out of range");
            double e[3][3];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) e[c][r] = pos[3 * v[c] + r] - pos[3 * v[3] + r];
            double const det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                               e[1][0] * (e[0][1] * e[2][2] - e[0][2] * e[2][1]) +
                               e[2][0] * (e[0][1] * e[1][2] - e[0][2] * e[1][1]);
            double const share = density * std::fabs(det) / 6.0 / 4.0;
            for (int k = 0; k < 4; ++k) mass[v[k]] += share;
        }
    });
}
