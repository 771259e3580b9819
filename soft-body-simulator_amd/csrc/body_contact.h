// Contacts between the surfaces of deformable bodies (body against body, and a body against its own
// non-incident surface triangles), implemented in body_contact.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "kernels.h"

namespace sbs {

// Cells of the triangle hash one triangle's box may cover; a triangle covering more (one stretched
// beyond the cell edge) goes to the list of large triangles that every vertex tests.
constexpr int kTriCells = 8;

// The surface triangles as contact obstacles: per triangle its 3 surface vertices (global surface
// numbering, outward orientation) and its body; the hash of one frame.
struct body_contact_dev {
    const uint4* tris;      // (a, b, c, body) surface vertex indices, outward orientation
    uint32_t ntris;
    double inv_cell;        // 1 / cell edge
    double thickness;       // h: the contact plane's height above the triangle
    double radius;          // r: contacts form for vertices closer than r (h < r < cell edge)
    uint32_t mask;          // table size - 1 (a power of two)
    uint32_t sdf_base;      // contact index of triangle t: sdf_base + t (after the SDF indices)
    uint32_t* keys;         // [kTriCells * ntris] slot of each (cell, triangle) pair, sorted
    uint32_t* vals;         // [kTriCells * ntris] its triangle
    uint32_t* start;        // [mask + 1] slot ranges into keys/vals
    uint32_t* end;
    uint32_t* big;          // [ntris] this frame's triangles covering more than kTriCells cells
    uint32_t* nbig;         // their number (one word)
    const uint32_t* sbody;  // body of each surface vertex
    const uint32_t* toff;   // [bodies + 1] body k's triangles are toff[k] .. toff[k + 1] - 1
    const uint4* nbr;       // [ntris] the triangle across edge k (corner k to corner k+1) in x, y, z;
                            // 0xffffffff where the edge has no neighbour
    const uint32_t* vt_off; // [S + 1] per surface vertex, its range of vt
    const uint32_t* vt;     // 3 t + k for each triangle t having the vertex as corner k, t increasing
    pos4* tn;               // [ntris] per frame: the triangle's unit normal (zero when degenerate)
    pos4* sph;              // [ntris] per frame: a sphere holding the triangle (center, radius)
    int4* tlo;              // [ntris] per frame: the lowest cell of the triangle's box
    uint4* tpx;             // [ntris] per frame: the corners' particles and the body
    uint4* erec;            // [kTriCells * ntris] per frame, hashed entry e: its triangle and that triangle's tlo
    float4* esph;           // [kTriCells * ntris] per frame: entry e's sphere relative to the corner of cell tlo
    pos4* vn;               // [S] per frame: the vertex's angle-weighted pseudo-normal
};

// Workspace (uint32 words) of the per-frame hash build for ntris triangles and a table of 2^bits.
std::size_t body_contact_workspace(uint32_t ntris);

// The hash of the triangles at positions x: every triangle's box emits its (slot, triangle) pairs
// (a triangle covering more than kTriCells cells goes to the large-triangle list instead), a stable
// radix sort groups them by slot (triangles in increasing order inside a slot) and each slot's range
// is marked. ws: body_contact_workspace words.
void build_body_hash(hipStream_t s, body_contact_dev const& b, const uint32_t* surf, const pos4* x, int bits,
                     uint32_t* ws);

// Per surface vertex s (at x), after the contacts already in c (its SDF contacts): one contact
// with the nearest triangle of another body with |p - closest point|^2 < r^2 (among the triangles
// hashed into the cells that the box p +- r overlaps, and the large ones), then one with the nearest such triangle of its own surface that does
// not have s as a corner and that p lies more than h / 10 outside of along the closest feature's
// pseudo-normal (the lowest index on ties): the closest point's barycentric weights in the triangle
// (in c.q) and the normal n along the line from the closest point c to p, pointing out of the
// surface (the triangle's normal when p lies on it). Computes the triangles' unit normals and the
// vertices' pseudo-normals first.
void launch_detect_bodies(hipStream_t s, contact_dev c, body_contact_dev const& b, const pos4* x, int block);

// Candidate pairs of the broad phase: per surface vertex the triangle entries of the 27 slots around
// its cell and the large triangles, those not incident to it (repeats included); count[s] (s < S) first, then with offsets
// (exclusive prefix of the counts) the pairs (s, t) into out[2 * (off[s] + i)].
void launch_body_candidates(hipStream_t s, contact_dev c, body_contact_dev const& b, const pos4* x,
                            const uint32_t* offsets, uint32_t* counts, uint32_t* out);

}  // namespace sbs
