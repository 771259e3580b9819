// Per-frame refresh of the boundary surface on the device (tetrahedral_body_t::update_visual_model,
// src/physics/tetrahedral_body.cpp:157-165, and tetrahedral_mesh_boundary_t::compute_normals,
// src/physics/tetrahedral_mesh_boundary.cpp:122-145): surface vertex positions are the particles'
// committed positions, and every vertex normal accumulates the unnormalized face normals
// (p2 - p1) x (p3 - p1) of its triangles in triangle order and is then normalized. As in the
// reference the accumulator is not cleared between frames: a frame adds its face normals to the
// previous frame's unit normal. FP64 without contraction into FMAs, like the reference's arithmetic.
#include "kernels.h"

namespace sbs {

namespace {

__global__ void k_surface_refresh(const uint32_t* __restrict__ surf, const pos4* __restrict__ x,
                                  const uint4* __restrict__ tris, const uint32_t* __restrict__ off,
                                  const uint32_t* __restrict__ inc, double4* __restrict__ normal, uint32_t S)
{
#pragma clang fp contract(off)
    uint32_t const s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    double4 nv = normal[s];
    for (uint32_t j = off[s]; j < off[s + 1]; ++j) {
        uint4 const t = tris[inc[j]];
        pos4 const p1 = x[surf[t.x]], p2 = x[surf[t.y]], p3 = x[surf[t.z]];
        double const a0 = p2.x - p1.x, a1 = p2.y - p1.y, a2 = p2.z - p1.z;
        double const b0 = p3.x - p1.x, b1 = p3.y - p1.y, b2 = p3.z - p1.z;
        nv.x += a1 * b2 - a2 * b1;
        nv.y += a2 * b0 - a0 * b2;
        nv.z += a0 * b1 - a1 * b0;
    }
    double const z = nv.x * nv.x + nv.y * nv.y + nv.z * nv.z;
    if (z > 0.0) {
        double const r = sqrt(z);
        nv.x /= r;
        nv.y /= r;
        nv.z /= r;
    }
    normal[s] = nv;
}

}  // namespace

void launch_surface_refresh(hipStream_t s, const uint32_t* surf, const pos4* x, const uint4* tris, const uint32_t* off,
                            const uint32_t* inc, double4* normal, uint32_t S)
{
    if (S) k_surface_refresh<<<(S + 255) / 256, 256, 0, s>>>(surf, x, tris, off, inc, normal, S);
}

}  // namespace sbs
