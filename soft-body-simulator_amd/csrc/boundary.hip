// Boundary surface extraction on the device (tetrahedral_mesh_boundary_t::extract_boundary_surface,
// src/physics/tetrahedral_mesh_boundary.cpp:65-120, with the face numbering of
// tetrahedron_set_t::add_tetrahedron, src/physics/topology.cpp:904-956).
//
// Face q = 4 t + f of tet t has the vertex order of tetrahedron_t::faces_copy
// {v1v2v4, v2v3v4, v3v1v4, v1v3v2} (topology.cpp:335-342). The faces are sorted by their sorted
// vertex triple, packed into 3b bits with b bits per vertex id (b <= 21), by stable radix sorts of
// the low and the high key word, so equal faces keep q order; a run of
// equal keys is one triangle of the mesh, created by its first face. Runs whose length is not 2 are
// boundary triangles; sorted by that first face they come in the mesh's triangle order, with that
// face's vertex order. Surface vertices are numbered by first appearance along the boundary
// triangles: the (vertex, corner) pairs are sorted by vertex, each vertex's first corner found,
// and the vertices sorted by it.
#include "abi_util.h"
#include "scan_sort.h"
#include "sbs_gpu.h"

namespace sbs {

namespace {

unsigned grid_for(std::size_t n) { return unsigned((n + 255) / 256); }

__device__ __forceinline__ void face_vertices(const uint32_t* tets, uint32_t q, uint32_t v[3])
{
    uint32_t const t = q >> 2, f = q & 3u;
    int const a = f == 0 ? 0 : (f == 1 ? 1 : (f == 2 ? 2 : 0));
    int const b = f == 0 ? 1 : (f == 1 ? 2 : (f == 2 ? 0 : 2));
    int const c = f == 3 ? 1 : 3;
    v[0] = tets[4 * std::size_t(t) + a];
    v[1] = tets[4 * std::size_t(t) + b];
    v[2] = tets[4 * std::size_t(t) + c];
}

// sorted vertex triple packed with b bits per id (3b <= 63)
__device__ __forceinline__ uint64_t face_key(const uint32_t* tets, uint32_t q, int b)
{
    uint32_t v[3];
    face_vertices(tets, q, v);
    uint32_t const lo = min(v[0], min(v[1], v[2])), hi = max(v[0], max(v[1], v[2]));
    uint32_t const mid = v[0] + v[1] + v[2] - lo - hi;
    return (uint64_t(lo) << (2 * b)) | (uint64_t(mid) << b) | uint64_t(hi);
}

__global__ void k_face_keys_lo(const uint32_t* __restrict__ tets, uint32_t nf, int b, uint32_t* __restrict__ keys,
                               uint32_t* __restrict__ vals)
{
    uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nf) return;
    keys[q] = uint32_t(face_key(tets, q, b));
    vals[q] = q;
}

__global__ void k_face_keys_hi(const uint32_t* __restrict__ tets, const uint32_t* __restrict__ perm, uint32_t nf,
                               int b, uint32_t* __restrict__ keys)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    keys[i] = uint32_t(face_key(tets, perm[i], b) >> 32);
}

// run starts of the sorted faces
__global__ void k_run_starts(const uint32_t* __restrict__ tets, const uint32_t* __restrict__ order, uint32_t nf,
                             int b, uint32_t* __restrict__ flag)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    flag[i] = (i == 0 || face_key(tets, order[i], b) != face_key(tets, order[i - 1], b)) ? 1u : 0u;
}

__global__ void k_compact_index(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, uint32_t n,
                                uint32_t* __restrict__ out)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) out[pos[i]] = i;
}

// boundary flag of run r (length != 2), with the run's first face
__global__ void k_boundary_runs(const uint32_t* __restrict__ starts, uint32_t nruns, uint32_t nf,
                                const uint32_t* __restrict__ order, uint32_t* __restrict__ flag,
                                uint32_t* __restrict__ first)
{
    uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    uint32_t const b = starts[r], e = r + 1 < nruns ? starts[r + 1] : nf;
    flag[r] = (e - b) != 2 ? 1u : 0u;
    first[r] = order[b];
}

__global__ void k_compact_values(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                 const uint32_t* __restrict__ val, uint32_t n, uint32_t* __restrict__ out)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) out[pos[i]] = val[i];
}

// (vertex, corner) pairs of the boundary triangles (faces in triangle order)
__global__ void k_corners(const uint32_t* __restrict__ tets, const uint32_t* __restrict__ faces, uint32_t nt,
                          uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    uint32_t const j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nt) return;
    uint32_t v[3];
    face_vertices(tets, faces[j], v);
    for (int c = 0; c < 3; ++c) {
        keys[3 * j + c] = v[c];
        vals[3 * j + c] = 3 * j + c;
    }
}

__global__ void k_vertex_starts(const uint32_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ flag)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// surface vertex r = vertex v: tet vertex -> surface index
__global__ void k_scatter_rank(const uint32_t* __restrict__ verts, uint32_t ns, uint32_t* __restrict__ s_of)
{
    uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < ns) s_of[verts[r]] = r;
}

__global__ void k_triangles(const uint32_t* __restrict__ tets, const uint32_t* __restrict__ faces, uint32_t nt,
                            const uint32_t* __restrict__ s_of, uint32_t* __restrict__ tri)
{
    uint32_t const j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nt) return;
    uint32_t v[3];
    face_vertices(tets, faces[j], v);
    for (int c = 0; c < 3; ++c) tri[3 * j + c] = s_of[v[c]];
}

uint32_t read_word(hipStream_t s, const uint32_t* p)
{
    uint32_t h = 0;
    hip_check(hipMemcpyAsync(&h, p, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read count");
    hip_check(hipStreamSynchronize(s), "read count");
    return h;
}

int bits_for(uint64_t n)
{
    int b = 1;
    while ((uint64_t(1) << b) < n) ++b;
    return b;
}

}  // namespace

}  // namespace sbs

using namespace sbs;

extern "C" int sbs_gpu_extract_boundary(int device, uint32_t num_vertices, const uint32_t* tets, uint32_t num_tets,
                                        uint32_t* surface_to_tet, uint32_t* triangles, uint32_t* num_surface_vertices,
                                        uint32_t* num_triangles)
{
    return guarded([&] {
        require(tets != nullptr || num_tets == 0, "tets is null");
        require(num_surface_vertices && num_triangles, "size outputs are null");
        require(num_vertices <= (1u << 21), "at most 2^21 vertices");
        require(num_tets <= (1u << 28), "at most 2^28 tets");
        for (std::size_t i = 0; i < 4 * std::size_t(num_tets); ++i)
            require(tets[i] < num_vertices, "tet vertex index out of range");
        *num_surface_vertices = *num_triangles = 0;
        if (num_tets == 0) return;
        hip_check(hipSetDevice(device), "hipSetDevice");
        hipStream_t s;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        try {
            uint32_t const nf = 4 * num_tets;
            int const b = bits_for(num_vertices);
            dev_buf<uint32_t> dt, keys, order, flag, pos, starts, first, bfaces, ws;
            dt.upload(tets, 4 * std::size_t(num_tets), s);
            keys.alloc(nf);
            order.alloc(nf);
            flag.alloc(nf);
            pos.alloc(std::size_t(nf) + 1);
            ws.alloc(std::max(sort_workspace(nf), scan_workspace(nf)));
            // faces sorted by their key, equal faces in q order
            k_face_keys_lo<<<grid_for(nf), 256, 0, s>>>(dt.p, nf, b, keys.p, order.p);
            radix_sort_pairs(s, keys.p, order.p, nf, std::min(3 * b, 32), ws.p);
            if (3 * b > 32) {
                k_face_keys_hi<<<grid_for(nf), 256, 0, s>>>(dt.p, order.p, nf, b, keys.p);
                radix_sort_pairs(s, keys.p, order.p, nf, 3 * b - 32, ws.p);
            }
            // runs of equal faces; boundary runs (length != 2) with their first face
            k_run_starts<<<grid_for(nf), 256, 0, s>>>(dt.p, order.p, nf, b, flag.p);
            exclusive_scan(s, flag.p, pos.p, nf, ws.p);
            starts.alloc(nf);
            k_compact_index<<<grid_for(nf), 256, 0, s>>>(flag.p, pos.p, nf, starts.p);
            hip_check(hipGetLastError(), "boundary face kernels");
            uint32_t const nr = read_word(s, pos.p + nf);
            first.alloc(nr);
            bfaces.alloc(nr + 1);
            k_boundary_runs<<<grid_for(nr), 256, 0, s>>>(starts.p, nr, nf, order.p, flag.p, first.p);
            exclusive_scan(s, flag.p, pos.p, nr, ws.p);
            k_compact_values<<<grid_for(nr), 256, 0, s>>>(flag.p, pos.p, first.p, nr, bfaces.p);
            uint32_t const nt = read_word(s, pos.p + nr);
            if (nt > 0) {
                // boundary triangles in the mesh's triangle order (by first face)
                radix_sort_pairs(s, bfaces.p, first.p, nt, bits_for(nf), ws.p);
                // surface vertices by first appearance along the triangles
                std::size_t const nc = 3 * std::size_t(nt);
                dev_buf<uint32_t> ck, cv, cflag, cpos, vfirst, vverts, s_of, tri, ws2;
                ck.alloc(nc);
                cv.alloc(nc);
                cflag.alloc(nc);
                cpos.alloc(nc + 1);
                vfirst.alloc(nc);
                vverts.alloc(nc);
                s_of.alloc(num_vertices);
                tri.alloc(nc);
                ws2.alloc(std::max(sort_workspace(nc), scan_workspace(nc)));
                k_corners<<<grid_for(nt), 256, 0, s>>>(dt.p, bfaces.p, nt, ck.p, cv.p);
                radix_sort_pairs(s, ck.p, cv.p, nc, b, ws2.p);
                k_vertex_starts<<<grid_for(nc), 256, 0, s>>>(ck.p, uint32_t(nc), cflag.p);
                exclusive_scan(s, cflag.p, cpos.p, nc, ws2.p);
                k_compact_values<<<grid_for(nc), 256, 0, s>>>(cflag.p, cpos.p, cv.p, uint32_t(nc), vfirst.p);
                k_compact_values<<<grid_for(nc), 256, 0, s>>>(cflag.p, cpos.p, ck.p, uint32_t(nc), vverts.p);
                uint32_t const ns = read_word(s, cpos.p + nc);
                radix_sort_pairs(s, vfirst.p, vverts.p, ns, bits_for(nc), ws2.p);
                k_scatter_rank<<<grid_for(ns), 256, 0, s>>>(vverts.p, ns, s_of.p);
                k_triangles<<<grid_for(nt), 256, 0, s>>>(dt.p, bfaces.p, nt, s_of.p, tri.p);
                hip_check(hipGetLastError(), "boundary kernels");
                if (surface_to_tet)
                    hip_check(hipMemcpyAsync(surface_to_tet, vverts.p, ns * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                              "read surface vertices");
                if (triangles)
                    hip_check(hipMemcpyAsync(triangles, tri.p, nc * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                              "read triangles");
                hip_check(hipStreamSynchronize(s), "boundary extraction");
                *num_surface_vertices = ns;
                *num_triangles = nt;
            }
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        hip_check(hipStreamDestroy(s), "hipStreamDestroy");
    });
}
