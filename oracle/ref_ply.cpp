// Test infrastructure only: C entry points over the reference's own PLY code (src/io/ply.cpp and
// src/common/geometry.cpp, compiled where they lie by `make -C oracle ref` into oracle/_ref/), so
// the parity tests can compare the library's PLY reader and writer with the reference's bytes and
// arrays. Nothing in the product links this.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <sstream>
#include <string>
#include <vector>

#include "sbs/io/ply.h"

extern "C" {

struct ref_geometry {
    float* positions;
    uint64_t num_positions;
    int32_t* indices;
    uint64_t num_indices;
    float* normals;
    uint64_t num_normals;
    float* uvs;
    uint64_t num_uvs;
    uint8_t* colors;
    uint64_t num_colors;
    int32_t geometry_type;
};

}

namespace {

template <class T, class U>
T* dup(std::vector<U> const& v, uint64_t* n)
{
    *n = v.size();
    if (v.empty()) return nullptr;
    T* p = static_cast<T*>(std::malloc(v.size() * sizeof(T)));
    std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

template <class T>
std::vector<T> vec(T const* p, uint64_t n)
{
    return n ? std::vector<T>(p, p + n) : std::vector<T>{};
}

}  // namespace

extern "C" {

// 0: a geometry, 1: read_ply returned no value, 2: read_ply threw
int ref_ply_read_memory(const char* data, uint64_t size, ref_geometry* out)
{
    try {
        std::istringstream is(std::string(data, data + size), std::ios::binary);
        auto g = sbs::io::read_ply(is);
        if (!g) return 1;
        out->positions = dup<float>(g->positions, &out->num_positions);
        out->indices = dup<int32_t>(g->indices, &out->num_indices);
        out->normals = dup<float>(g->normals, &out->num_normals);
        out->uvs = dup<float>(g->uvs, &out->num_uvs);
        out->colors = dup<uint8_t>(g->colors, &out->num_colors);
        out->geometry_type = g->geometry_type == sbs::common::geometry_t::geometry_type_t::triangle      ? 0
                             : g->geometry_type == sbs::common::geometry_t::geometry_type_t::tetrahedron ? 1
                                                                                                          : -1;
        return 0;
    } catch (std::exception const&) {
        return 2;
    }
}

// format 0 ascii, 1 binary little endian, 2 binary big endian; *out is malloc'd
int ref_ply_write_memory(const ref_geometry* g, int format, char** out, uint64_t* size)
{
    sbs::common::geometry_t geo;
    geo.positions = vec(g->positions, g->num_positions);
    geo.indices = vec(g->indices, g->num_indices);
    geo.normals = vec(g->normals, g->num_normals);
    geo.uvs = vec(g->uvs, g->num_uvs);
    geo.colors = vec(g->colors, g->num_colors);
    geo.geometry_type = g->geometry_type == 1 ? sbs::common::geometry_t::geometry_type_t::tetrahedron
                                              : sbs::common::geometry_t::geometry_type_t::triangle;
    std::ostringstream os(std::ios::binary);
    sbs::io::write_ply(os, geo,
                       format == 1   ? sbs::io::ply_format_t::binary_little_endian
                       : format == 2 ? sbs::io::ply_format_t::binary_big_endian
                                     : sbs::io::ply_format_t::ascii);
    std::string const s = os.str();
    *size = s.size();
    *out = static_cast<char*>(std::malloc(s.size() ? s.size() : 1));
    std::memcpy(*out, s.data(), s.size());
    return 0;
}

void ref_free(void* p) { std::free(p); }

}
