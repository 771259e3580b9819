// Contacts between deformable bodies on gfx950: a spatial hash of the surface triangles rebuilt
// every frame, and a nearest-triangle test per surface vertex.
//
// The reference detects contacts once per frame, from the positions at the frame's start, for
// every pair of collision models (brute_force_cd_system.cpp:17-25); its only contact is a surface
// particle against an SDF, whose contact point q and normal n stay fixed for the frame's substeps
// (contact_handler.cpp:14-54, collision_constraint.cpp:21-41), and a pair of point-cloud models
// yields nothing (bvh_model.cpp:30-34). Here the surface triangles of the bodies take the SDF's
// part for the pairs it leaves empty. Once per frame, a surface vertex p gets up to two contacts: with
// the nearest triangle of another body closer than a radius r (p inside or outside that body), and
// with the nearest triangle of its own surface closer than r that does not have p as a corner and
// that p lies more than h / 10 outside of (a fold of the surface approaching or resting on it). Inside and outside come
// from the angle-weighted pseudo-normal N of the feature holding the closest point c (the face's
// normal, the sum of the two face normals of an edge, the angle-weighted sum of a corner's face
// normals; Baerentzen and Aanaes, "Signed distance computation using the angle weighted
// pseudonormal", IEEE TVCG 11(3), 2005): p lies inside when (p - c).N < 0, right at concave and
// convex edges and corners where the side of one triangle's plane is not, and more than h / 10
// outside when (p - c).N > (h / 10) |N| (never for the flat or convex surface around p itself; a
// fold held at the thickness h stays well above that). A contact keeps
// c's barycentric weights in the triangle and the normal n along the line from c to p, pointing
// out of the surface (the triangle's normal when p lies on the triangle). Every contact launch
// takes c at the triangle's corners as they are when it starts (a copy of the surface positions,
// so that the projections of one launch stay independent), and the constraint (x - q) . n >= 0
// with q = c + h n keeps p at least the thickness h < r away from the moving triangle; it acts only
// when p would cross that plane (a vertex approaching by less than r - h per frame is caught before
// it reaches the surface), and p moves by its share of the two-sided projection (the triangle
// point's inverse mass in the denominator). The contacts follow p's SDF contacts (another body's
// first, the lowest triangle index on ties) and are projected with them in the contact slot of each
// iteration.
//
// Broad phase: cells of edge H (twice the mean surface edge by default, r < H) hashed into a table
// of 2^b slots (Teschner et al. 2003). A triangle whose box covers at most kTriCells cells writes
// its (slot, triangle) pairs in triangle order, padding pairs carrying the slot key 2^b (sorted past
// every slot); a triangle covering more (stretched beyond the cell edge) goes to a list of large
// triangles that every vertex tests. A stable radix sort groups the pairs by slot and each slot's
// range is marked. A triangle closer than r < H to a vertex overlaps a cell of the 3 x 3 x 3 block
// around the vertex's cell, so the vertex finds it in one of those 27 slots or in the large list:
// the nearest triangle found this way is the brute force's. All FP64 without contraction, so the
// contacts equal the oracle's bit for bit. Every launch has a fixed size (graph capture).
#include "body_contact.h"
#include "contact_hash.h"
#include "scan_sort.h"
#include "sdf.h"

namespace sbs {

namespace {

unsigned grid_for(std::size_t n, int block) { return unsigned((n + std::size_t(block) - 1) / std::size_t(block)); }

__device__ __forceinline__ void tri_corners(body_contact_dev const& b, const uint32_t* surf, const pos4* x,
                                            uint4 t, double c[9])
{
    pos4 const pa = x[surf[t.x]], pb = x[surf[t.y]], pc = x[surf[t.z]];
    c[0] = pa.x, c[1] = pa.y, c[2] = pa.z;
    c[3] = pb.x, c[4] = pb.y, c[5] = pb.z;
    c[6] = pc.x, c[7] = pc.y, c[8] = pc.z;
}

// Unnormalized normal (b - a) x (c - a) of a triangle's corners.
__device__ __forceinline__ void tri_normal(const double* corners, double n[3])
{
#pragma clang fp contract(off)
    double const e1[3] = {corners[3] - corners[0], corners[4] - corners[1], corners[5] - corners[2]};
    double const e2[3] = {corners[6] - corners[0], corners[7] - corners[1], corners[8] - corners[2]};
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
}

__global__ void k_tri_emit(body_contact_dev b, const uint32_t* __restrict__ surf, const pos4* __restrict__ x)
{
#pragma clang fp contract(off)
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.ntris) return;
    double c[9];
    tri_corners(b, surf, x, b.tris[t], c);
    int lo[3], hi[3];
    uint64_t cells = 1;
    for (int d = 0; d < 3; ++d) {
        double const mn = fmin(fmin(c[d], c[3 + d]), c[6 + d]), mx = fmax(fmax(c[d], c[3 + d]), c[6 + d]);
        lo[d] = hash_cell(mn, b.inv_cell);
        hi[d] = hash_cell(mx, b.inv_cell);
        cells *= uint64_t(uint32_t(hi[d] - lo[d]) + 1u);
    }
    uint32_t* keys = b.keys + std::size_t(t) * kTriCells;
    uint32_t* vals = b.vals + std::size_t(t) * kTriCells;
    uint32_t const pad = b.mask + 1u;
    int j = 0;
    if (cells <= uint64_t(kTriCells)) {
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int jj = lo[1]; jj <= hi[1]; ++jj)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    keys[j] = hash_slot(i, jj, k, b.mask);
                    vals[j] = t;
                    ++j;
                }
    } else {  // a triangle stretched beyond the cell edge: every vertex tests it
        b.big[atomicAdd(b.nbig, 1u)] = t;
    }
    for (; j < kTriCells; ++j) {
        keys[j] = pad;
        vals[j] = t;
    }
}

__global__ void k_fill2(uint32_t* __restrict__ a, uint32_t* __restrict__ b, uint32_t n, uint32_t* __restrict__ c)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = b[i] = 0u;
    if (i == 0) *c = 0u;
}

__global__ void k_body_slot_ranges(body_contact_dev b, uint32_t n)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t const k = b.keys[i];
    if (k > b.mask) return;
    if (i == 0 || b.keys[i - 1] != k) b.start[k] = i;
    if (i + 1 == n || b.keys[i + 1] != k) b.end[k] = i + 1;
}

// Every triangle entry of the slots of the 27 cells around p's cell, then every large triangle:
// visit(t) (a triangle may come more than once, from several slots).
template <class Visit>
__device__ __forceinline__ void for_neighbour_triangles(body_contact_dev const& b, pos4 p, Visit&& visit)
{
    int const ci = hash_cell(p.x, b.inv_cell), cj = hash_cell(p.y, b.inv_cell), ck = hash_cell(p.z, b.inv_cell);
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                uint32_t const slot = hash_slot(ci + di, cj + dj, ck + dk, b.mask);
                uint32_t prev = 0xffffffffu;
                for (uint32_t i = b.start[slot], e = b.end[slot]; i < e; ++i) {
                    uint32_t const t = b.vals[i];
                    if (t == prev) continue;  // two cells of one triangle in one slot
                    prev = t;
                    visit(t);
                }
            }
    for (uint32_t i = 0, e = *b.nbig; i < e; ++i) visit(b.big[i]);
}

__global__ void k_tri_unit_normals(body_contact_dev b, const uint32_t* __restrict__ surf, const pos4* __restrict__ x)
{
#pragma clang fp contract(off)
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.ntris) return;
    double c[9], n[3];
    tri_corners(b, surf, x, b.tris[t], c);
    tri_normal(c, n);
    double const l2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    double const l = sqrt(l2);
    b.tn[t] = l2 > 0.0 ? make_double4(n[0] / l, n[1] / l, n[2] / l, 0.0) : make_double4(0.0, 0.0, 0.0, 0.0);
}

// N(s) = sum over the triangles t having s as corner k, in increasing t, of the corner's angle
// atan2(|e1 x e2|, e1 . e2) times t's unit normal (e1, e2: the edges from the corner to the next
// two corners).
__global__ void k_vertex_pseudonormals(body_contact_dev b, const uint32_t* __restrict__ surf,
                                       const pos4* __restrict__ x, uint32_t S)
{
#pragma clang fp contract(off)
    uint32_t const s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    double N[3] = {0.0, 0.0, 0.0};
    for (uint32_t i = b.vt_off[s], e = b.vt_off[s + 1]; i < e; ++i) {
        uint32_t const t = b.vt[i] / 3u, k = b.vt[i] % 3u;
        double c[9];
        tri_corners(b, surf, x, b.tris[t], c);
        double const* o = c + 3 * k;
        double const* u = c + 3 * ((k + 1u) % 3u);
        double const* v = c + 3 * ((k + 2u) % 3u);
        double const e1[3] = {u[0] - o[0], u[1] - o[1], u[2] - o[2]};
        double const e2[3] = {v[0] - o[0], v[1] - o[1], v[2] - o[2]};
        double const cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                     cz = e1[0] * e2[1] - e1[1] * e2[0];
        double const ang = atan2(sqrt(cx * cx + cy * cy + cz * cz), e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]);
        pos4 const n = b.tn[t];
        N[0] = N[0] + ang * n.x;
        N[1] = N[1] + ang * n.y;
        N[2] = N[2] + ang * n.z;
    }
    b.vn[s] = make_double4(N[0], N[1], N[2], 0.0);
}

// The pseudo-normal of feature f of triangle t (tri_closest_point's numbering).
__device__ __forceinline__ void pseudo_normal(body_contact_dev const& b, uint32_t t, int f, double N[3])
{
#pragma clang fp contract(off)
    pos4 const n = b.tn[t];
    uint4 const tr = b.tris[t];
    uint4 const nb = b.nbr[t];
    pos4 v;
    uint32_t other = 0xffffffffu;
    switch (f) {
    case 0: v = b.vn[tr.x]; break;
    case 1: v = b.vn[tr.y]; break;
    case 2: v = b.vn[tr.z]; break;
    case 3: other = nb.x; v = n; break;
    case 4: other = nb.y; v = n; break;
    case 5: other = nb.z; v = n; break;
    default: v = n; break;
    }
    N[0] = v.x, N[1] = v.y, N[2] = v.z;
    if (other != 0xffffffffu) {
        pos4 const m = b.tn[other];
        N[0] = n.x + m.x, N[1] = n.y + m.y, N[2] = n.z + m.z;
    }
}

// The nearest non-incident triangle of surface vertex s closer than r among another body's triangles
// (own false) or its own surface's (own true, with p more than h / 10 outside along the
// pseudo-normal N of the closest feature: (p - c).N > (h / 10) |N|), the lowest index on ties
// whatever the visiting order;
// returns the triangle (or 0xffffffff), its closest point, the point's barycentric weights and its
// feature (tri_closest_point's numbering).
__device__ __forceinline__ uint32_t nearest_triangle(contact_dev const& c, body_contact_dev const& b, const pos4* x,
                                                     uint32_t s, pos4 p, bool own, double best_q[3],
                                                     double best_w[3], int& best_f)
{
#pragma clang fp contract(off)
    double best = b.radius * b.radius;
    uint32_t bt = 0xffffffffu;
    double const pp[3] = {p.x, p.y, p.z};
    uint32_t const body = b.sbody[s];
    for_neighbour_triangles(b, p, [&](uint32_t t) {
        uint4 const tr = b.tris[t];
        if ((tr.w == body) != own) return;
        if (tr.x == s || tr.y == s || tr.z == s) return;
        double corners[9], q[3], wts[3];
        tri_corners(b, c.surf, x, tr, corners);
        int feature;
        tri_closest_point(pp, corners, q, &feature, wts);
        double const dx = pp[0] - q[0], dy = pp[1] - q[1], dz = pp[2] - q[2];
        double const d2 = dx * dx + dy * dy + dz * dz;
        if (!(d2 < best || (d2 == best && bt != 0xffffffffu && t < bt))) return;
        if (own) {
            double N[3];
            pseudo_normal(b, t, feature, N);
            double const ln = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
            if (!(dx * N[0] + dy * N[1] + dz * N[2] > 0.1 * b.thickness * ln)) return;
        }
        best = d2;
        bt = t;
        best_f = feature;
        best_q[0] = q[0], best_q[1] = q[1], best_q[2] = q[2];
        best_w[0] = wts[0], best_w[1] = wts[1], best_w[2] = wts[2];
    });
    return bt;
}

__global__ void k_detect_bodies(contact_dev c, body_contact_dev b, const pos4* __restrict__ x)
{
#pragma clang fp contract(off)
    uint32_t const s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c.S) return;
    pos4 const p = x[c.surf[s]];
    for (int own = 0; own < 2; ++own) {  // another body's nearest triangle, then its own surface's
        double cq[3], wts[3];
        int feature = 6;
        uint32_t const t = nearest_triangle(c, b, x, s, p, own != 0, cq, wts, feature);
        if (t == 0xffffffffu) continue;
        // along the line from the closest point to the vertex, away from the surface (towards the
        // closest point when the vertex lies inside another body); the triangle's normal when the
        // vertex lies on the triangle
        pos4 const tn = b.tn[t];
        double n[3] = {tn.x, tn.y, tn.z};
        double const d[3] = {p.x - cq[0], p.y - cq[1], p.z - cq[2]};
        double const dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (dd > 0.0) {
            double N[3];
            pseudo_normal(b, t, feature, N);
            double const l = d[0] * N[0] + d[1] * N[1] + d[2] * N[2] < 0.0 ? -sqrt(dd) : sqrt(dd);
            n[0] = d[0] / l, n[1] = d[1] / l, n[2] = d[2] / l;
        }
        if (!(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] > 0.0)) continue;  // a degenerate triangle
        uint32_t const k = c.count[s];
        if (k == c.slots) {
            atomicAdd(c.overflow, 1u);
            return;
        }
        // the contact keeps the closest point's barycentric weights: its plane follows the triangle
        store_contact(c, s, k, b.sdf_base + t, wts, n);
        c.count[s] = k + 1;
    }
}

// The non-incident triangle entries of the 27 slots around the vertex (repeats included):
// counted (out null) or written as (s, t) pairs.
__global__ void k_body_candidates(contact_dev c, body_contact_dev b, const pos4* __restrict__ x,
                                  const uint32_t* __restrict__ offsets, uint32_t* __restrict__ counts,
                                  uint32_t* __restrict__ out)
{
    uint32_t const s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c.S) return;
    pos4 const p = x[c.surf[s]];
    uint32_t n = 0;
    for_neighbour_triangles(b, p, [&](uint32_t t) {
        uint4 const tr = b.tris[t];
        if (tr.x == s || tr.y == s || tr.z == s) return;
        if (out) {
            out[2 * (std::size_t(offsets[s]) + n)] = s;
            out[2 * (std::size_t(offsets[s]) + n) + 1] = t;
        }
        ++n;
    });
    if (!out) counts[s] = n;
}

}  // namespace

std::size_t body_contact_workspace(uint32_t ntris) { return sort_workspace(std::size_t(kTriCells) * ntris); }

void build_body_hash(hipStream_t s, body_contact_dev const& b, const uint32_t* surf, const pos4* x, int bits,
                     uint32_t* ws)
{
    uint32_t const T = b.mask + 1u;
    k_fill2<<<grid_for(T, 256), 256, 0, s>>>(b.start, b.end, T, b.nbig);
    if (!b.ntris) return;
    std::size_t const n = std::size_t(kTriCells) * b.ntris;
    k_tri_emit<<<grid_for(b.ntris, 256), 256, 0, s>>>(b, surf, x);
    radix_sort_pairs(s, b.keys, b.vals, n, bits + 1, ws);  // the padding key 2^bits needs bit `bits`
    k_body_slot_ranges<<<grid_for(n, 256), 256, 0, s>>>(b, uint32_t(n));
}

void launch_detect_bodies(hipStream_t s, contact_dev c, body_contact_dev const& b, const pos4* x, int block)
{
    if (!c.S || !b.ntris) return;
    k_tri_unit_normals<<<grid_for(b.ntris, 256), 256, 0, s>>>(b, c.surf, x);
    k_vertex_pseudonormals<<<grid_for(c.S, 256), 256, 0, s>>>(b, c.surf, x, c.S);
    k_detect_bodies<<<grid_for(c.S, block), block, 0, s>>>(c, b, x);
}

void launch_body_candidates(hipStream_t s, contact_dev c, body_contact_dev const& b, const pos4* x,
                            const uint32_t* offsets, uint32_t* counts, uint32_t* out)
{
    if (c.S) k_body_candidates<<<grid_for(c.S, 256), 256, 0, s>>>(c, b, x, offsets, counts, out);
}

}  // namespace sbs
