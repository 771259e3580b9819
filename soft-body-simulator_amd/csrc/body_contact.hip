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
// around the vertex's cell (the candidate pairs list those 27 slots), and more precisely one of the
// cells that the box p +- r overlaps (the search visits those), or it is in the large list: the
// nearest triangle found this way is the brute force's. All FP64 without contraction, so the
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
    b.tlo[t] = make_int4(lo[0], lo[1], lo[2], 0);
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
    // the centroid and the largest corner distance from it, widened by 1e-9 of itself: a bound that
    // rounding cannot undercut, for the search's early rejections
    double const cx = (c[0] + c[3] + c[6]) / 3.0, cy = (c[1] + c[4] + c[7]) / 3.0, cz = (c[2] + c[5] + c[8]) / 3.0;
    double r2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double const dx = c[3 * k] - cx, dy = c[3 * k + 1] - cy, dz = c[3 * k + 2] - cz;
        r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
    }
    b.sph[t] = make_double4(cx, cy, cz, sqrt(r2) * (1.0 + 1e-9));
    uint4 const tr = b.tris[t];
    b.tpx[t] = make_uint4(surf[tr.x], surf[tr.y], surf[tr.z], tr.w);
}

// Per hashed entry e (key below the padding slot), the search's record of it: the triangle and the
// lowest cell lo of its box (erec), and a sphere holding the triangle in FP32 relative to that
// cell's corner O = lo * cell edge (esph: the center minus O rounded to FP32, the radius widened by
// twice the largest per-axis error of that rounding and of the FP64 subtraction before it, then
// rounded up; twice covers the sqrt(3) of three axes).
__global__ void k_entry_records(body_contact_dev b, uint32_t n)
{
#pragma clang fp contract(off)
    uint32_t const e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n || b.keys[e] > b.mask) return;
    uint32_t const t = b.vals[e];
    int4 const lo = b.tlo[t];
    pos4 const sp = b.sph[t];
    double const cs = 1.0 / b.inv_cell;
    double const d[3] = {sp.x - double(lo.x) * cs, sp.y - double(lo.y) * cs, sp.z - double(lo.z) * cs};
    float const f[3] = {float(d[0]), float(d[1]), float(d[2])};
    double err = 0.0;
    for (int k = 0; k < 3; ++k) err = fmax(err, fabs(double(f[k]) - d[k]) + 1e-15 * fabs(d[k]));
    double const R = sp.w + 2.0 * err;
    float Rf = float(R);
    if (double(Rf) < R) Rf = nextafterf(Rf, INFINITY);
    b.erec[e] = make_uint4(t, uint32_t(lo.x), uint32_t(lo.y), uint32_t(lo.z));
    b.esph[e] = make_float4(f[0], f[1], f[2], Rf);
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

constexpr int kDetectBlock = 256;

// Records the contact of surface vertex s (at p) with triangle t, whose closest point q (barycentric
// weights w) lies on feature f: the normal runs along the line from q to p, away from the surface
// (towards q when the vertex lies inside another body), the triangle's normal when p lies on it.
// Returns false when the vertex has no free slot (counted in c.overflow).
__device__ __forceinline__ bool emit_triangle_contact(contact_dev const& c, body_contact_dev const& b, uint32_t s,
                                                      pos4 p, uint32_t t, int f, const double q[3], const double w[3])
{
#pragma clang fp contract(off)
    pos4 const tn = b.tn[t];
    double n[3] = {tn.x, tn.y, tn.z};
    double const d[3] = {p.x - q[0], p.y - q[1], p.z - q[2]};
    double const dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (dd > 0.0) {
        double N[3];
        pseudo_normal(b, t, f, N);
        double const l = d[0] * N[0] + d[1] * N[1] + d[2] * N[2] < 0.0 ? -sqrt(dd) : sqrt(dd);
        n[0] = d[0] / l, n[1] = d[1] / l, n[2] = d[2] / l;
    }
    if (!(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] > 0.0)) return true;  // a degenerate triangle
    uint32_t const slot = c.count[s];
    if (slot == c.slots) {
        atomicAdd(c.overflow, 1u);
        return false;
    }
    // the contact keeps the closest point's barycentric weights: its plane follows the triangle
    store_contact(c, s, slot, b.sdf_base + t, w, n);
    c.count[s] = slot + 1;
    return true;
}

constexpr int kDetectQueue = 32;  // candidates a lane holds before it computes their closest points
constexpr int kDeferQueue = 16;   // own-surface candidates farther than kNear r, held back
constexpr double kNear = 0.25;
constexpr uint32_t kDenseEntries = 512;  // hashed entries in a vertex's cells above which it holds back

// x as a float no smaller than x.
__device__ __forceinline__ float float_up(double x)
{
    float const f = float(x);
    return double(f) < x ? nextafterf(f, INFINITY) : f;
}

// Per surface vertex s at p, in one search, the nearest triangle of another body and the nearest
// of its own surface (not having s as a corner, with p more than h / 10 outside along the
// pseudo-normal N of the closest feature: (p - c).N > (h / 10) |N|), each closer than r, the lowest
// index on ties whatever the visiting order; a contact for each, another body's first.
// The search visits the triangles hashed into the slots of the cells that the box p +- r (widened
// by 1e-6 of r against rounding) overlaps, then the large triangles: every triangle with a point
// within r of p is among them, since that point lies in the box and in the triangle's own box of
// cells (a subset of the 27-cell neighbourhood of the candidate pairs). A hashed triangle is taken
// only from the first of those cells (in visiting order) that its own box covers, the corner
// max(lo, lowest query cell): once, and never lost. A triangle whose bounding sphere lies farther
// from p than its class's best so far (the class from the body's triangle range) cannot be nearer
// and is dropped before its closest point is computed. The scan tests that in FP32 relative to the
// corner of the cell being scanned, with a margin of 1e-5 of the cell edge and of the bound (the
// FP32 roundings of a center within 8 cells come to ~1e-7 m); the queued candidates are tested
// again in FP64 (margin 1e-9) just before their closest points.
// Each lane scans its entries (one 32-byte record each, no dependent loads) into a queue in LDS
// until it holds kDetectQueue candidates or the entries end, then computes the queued closest
// points, each after testing its sphere again against the class's best as it is by then: the FP64
// closest-point code runs with most of the wave's lanes busy instead of the fraction whose current
// entry passed. The two bests are plain scalars of the kernel (registers).
// Where a vertex's cells hold more than kDenseEntries entries (a surface folded onto itself),
// own-surface triangles farther than r1 = kNear r from p go last: they wait in a second queue (or,
// when more of them come than it holds, a second scan collects them), and are skipped altogether
// when the own surface's best ends within r1, which no triangle farther than r1 can beat; there the
// nearest accepted triangle is usually close, and most triangles within r are never computed.
// Blocks go round robin over the 8 XCDs in launch order; XCD k takes a contiguous run of vertices so
// that neighbouring vertices, which read the same cells, share its L2.
__global__ void __launch_bounds__(kDetectBlock) __attribute__((amdgpu_waves_per_eu(3))) k_detect_bodies(contact_dev c, body_contact_dev b, const pos4* __restrict__ x)
{
#pragma clang fp contract(off)
    __shared__ uint32_t queue[kDetectQueue * kDetectBlock];
    __shared__ uint32_t deferred[kDeferQueue * kDetectBlock];
    uint32_t const xcd = blockIdx.x % 8u, per = gridDim.x / 8u, extra = gridDim.x % 8u;
    uint32_t const logical = xcd * per + min(xcd, extra) + blockIdx.x / 8u;
    uint32_t const s = logical * kDetectBlock + threadIdx.x;
    if (s >= c.S) return;
    uint32_t const ps = c.surf[s];
    pos4 const p = x[ps];
    double const pp[3] = {p.x, p.y, p.z};
    uint32_t const body = b.sbody[s];
    uint32_t const own0 = b.toff[body], nown = b.toff[body + 1] - own0;  // the own surface's triangles
    // another body's nearest (o) and the own surface's (m): triangle, squared distance and its root
    // (the closest point is computed again for the two winners at the end)
    uint32_t to = 0xffffffffu, tm = 0xffffffffu;
    double bo = b.radius * b.radius, bm = bo, ro = b.radius, rm = b.radius;
    double const cs = 1.0 / b.inv_cell;
    double const rr = b.radius * (1.0 + 1e-6);
    int const i0 = hash_cell(p.x - rr, b.inv_cell), ni = hash_cell(p.x + rr, b.inv_cell) - i0 + 1;
    int const j0 = hash_cell(p.y - rr, b.inv_cell), nj = hash_cell(p.y + rr, b.inv_cell) - j0 + 1;
    int const k0 = hash_cell(p.z - rr, b.inv_cell), nk = hash_cell(p.z + rr, b.inv_cell) - k0 + 1;
    int const ncells = ni * nj * nk;
    uint32_t* const qu = queue + threadIdx.x;     // this lane's queue, stride kDetectBlock
    uint32_t* const dq = deferred + threadIdx.x;  // its held-back own-surface candidates
    double const r1 = kNear * b.radius;
    float const cs32 = float(cs), slack = 1e-5f * cs32, r1_32 = float_up(r1);
    float rm32 = float_up(rm), ro32 = rm32;  // the bounds rounded up
    float pc[3] = {0.0f, 0.0f, 0.0f};       // p relative to the corner of the cell being scanned
    uint32_t entries = 0;
    for (int k = 0; k < ncells; ++k) {
        uint32_t const slot = hash_slot(i0 + k % ni, j0 + (k / ni) % nj, k0 + k / (ni * nj), b.mask);
        entries += b.end[slot] - b.start[slot];
    }
    bool const dense = entries > kDenseEntries;
    int dcnt = 0;
    bool dover = false;  // more were held back than dq holds
    // mode 0: every candidate, the far own-surface ones held back; 1: a second scan for those;
    // 2: the held-back queue is being computed
    int mode = 0;
    // the stream: cell `cell` (ncells: the large triangles), entries [e, n), the last triangle taken
    int cell = -1, ci = 0, cj = 0, ck = 0;
    uint32_t e = 0, n = 0, prev = 0xffffffffu;
    bool more = true;
    for (;;) {
        int cnt = 0;
        const uint32_t* src = qu;
        if (!more) {
            if (mode != 0 || !(rm > r1)) break;  // done, or nothing held back can win
            if (dover) {  // scan again for the held-back candidates
                mode = 1, more = true, cell = -1, e = 0, n = 0;
                continue;
            }
            src = dq, cnt = dcnt, mode = 2;
        }
        while (more && cnt <= kDetectQueue - 4) {  // scan: up to 4 entries a step, into the queue
            if (e == n) {
                if (++cell > ncells) {
                    more = false;
                    break;
                }
                prev = 0xffffffffu;
                if (cell < ncells) {
                    ci = i0 + cell % ni, cj = j0 + (cell / ni) % nj, ck = k0 + cell / (ni * nj);
                    uint32_t const slot = hash_slot(ci, cj, ck, b.mask);
                    e = b.start[slot], n = b.end[slot];
                    pc[0] = float(pp[0] - double(ci) * cs), pc[1] = float(pp[1] - double(cj) * cs);
                    pc[2] = float(pp[2] - double(ck) * cs);
                } else {
                    e = 0, n = *b.nbig;
                }
                continue;
            }
            if (cell < ncells) {
                uint32_t const m = min(n - e, 4u);
                uint4 r[4];
                float4 f[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (uint32_t(u) < m) r[u] = b.erec[e + u], f[u] = b.esph[e + u];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (uint32_t(u) >= m) break;
                    uint32_t const t = r[u].x;
                    if (t == prev) continue;  // two cells of one triangle in one slot
                    prev = t;
                    int const lx = int(r[u].y), ly = int(r[u].z), lz = int(r[u].w);
                    if (max(lx, i0) != ci || max(ly, j0) != cj || max(lz, k0) != ck) continue;  // not its first cell
                    // the sphere's center relative to this cell's corner: lo - cell is 0 or, on an
                    // axis where the triangle's box starts below the query box, negative
                    float const ex = pc[0] - fmaf(float(lx - ci), cs32, f[u].x);
                    float const ey = pc[1] - fmaf(float(ly - cj), cs32, f[u].y);
                    float const ez = pc[2] - fmaf(float(lz - ck), cs32, f[u].z);
                    float const e2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
                    bool const own = t - own0 < nown;
                    float const l = fmaf(f[u].w + (own ? rm32 : ro32), 1.0f + 1e-5f, slack);
                    if (e2 > l * l) continue;
                    float const l1 = fmaf(f[u].w + r1_32, 1.0f + 1e-5f, slack);
                    bool const far = dense && own && e2 > l1 * l1;  // every point farther than r1
                    if (mode == 1) {
                        if (!far) continue;
                    } else if (far) {
                        if (dcnt < kDeferQueue)
                            dq[dcnt++ * kDetectBlock] = e + u;
                        else
                            dover = true;
                        continue;
                    }
                    qu[cnt * kDetectBlock] = e + u;
                    ++cnt;
                }
                e += m;
            } else if (mode == 1) {  // the large triangles were all computed in the first scan
                e = n;
            } else {
                uint32_t const t = b.big[e++];
                pos4 const sp = b.sph[t];
                double const ex = pp[0] - sp.x, ey = pp[1] - sp.y, ez = pp[2] - sp.z;
                double const l = (sp.w + (t - own0 < nown ? rm : ro)) * (1.0 + 1e-9);
                if (!(ex * ex + ey * ey + ez * ez > l * l)) {
                    qu[cnt * kDetectBlock] = 0x80000000u | t;
                    ++cnt;
                }
            }
        }
        for (int i = 0; i < cnt; ++i) {  // the queued candidates' closest points
            uint32_t const qe = src[i * kDetectBlock];
            uint32_t t;
            if (qe & 0x80000000u) {  // a large triangle
                t = qe & 0x7fffffffu;
            } else {  // a hashed entry: tested again against the class's best as it is now
                uint4 const r = b.erec[qe];
                float4 const f = b.esph[qe];
                t = r.x;
                double const ex = (pp[0] - double(int(r.y)) * cs) - double(f.x);
                double const ey = (pp[1] - double(int(r.z)) * cs) - double(f.y);
                double const ez = (pp[2] - double(int(r.w)) * cs) - double(f.z);
                double const l = (double(f.w) + (t - own0 < nown ? rm : ro)) * (1.0 + 1e-9);
                if (ex * ex + ey * ey + ez * ez > l * l) continue;
            }
            uint4 const tp = b.tpx[t];
            bool const mine = tp.w == body;
            if (mine && (tp.x == ps || tp.y == ps || tp.z == ps)) continue;
            pos4 const pa = x[tp.x], pb = x[tp.y], pc = x[tp.z];
            double const corners[9] = {pa.x, pa.y, pa.z, pb.x, pb.y, pb.z, pc.x, pc.y, pc.z};
            double q[3], wts[3];
            int feature;
            tri_closest_point(pp, corners, q, &feature, wts);
            double const dx = pp[0] - q[0], dy = pp[1] - q[1], dz = pp[2] - q[2];
            double const d2 = dx * dx + dy * dy + dz * dz;
            double const best = mine ? bm : bo;
            uint32_t const bt = mine ? tm : to;
            if (!(d2 < best || (d2 == best && bt != 0xffffffffu && t < bt))) continue;
            if (mine) {
                double N[3];
                pseudo_normal(b, t, feature, N);
                double const ln = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
                if (!(dx * N[0] + dy * N[1] + dz * N[2] > 0.1 * b.thickness * ln)) continue;
                bm = d2, rm = sqrt(d2), tm = t, rm32 = float_up(rm);
            } else {
                bo = d2, ro = sqrt(d2), to = t, ro32 = float_up(ro);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {  // another body's nearest triangle, then its own surface's
        uint32_t const t = k == 0 ? to : tm;
        if (t == 0xffffffffu) continue;
        uint4 const tp = b.tpx[t];
        pos4 const pa = x[tp.x], pb = x[tp.y], pc = x[tp.z];
        double const corners[9] = {pa.x, pa.y, pa.z, pb.x, pb.y, pb.z, pc.x, pc.y, pc.z};
        double q[3], wts[3];
        int feature;
        tri_closest_point(pp, corners, q, &feature, wts);
        if (!emit_triangle_contact(c, b, s, p, t, feature, q, wts)) return;
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
    uint32_t const n = uint32_t(std::size_t(kTriCells) * b.ntris);
    k_entry_records<<<grid_for(n, 256), 256, 0, s>>>(b, n);
    k_vertex_pseudonormals<<<grid_for(c.S, 256), 256, 0, s>>>(b, c.surf, x, c.S);
    k_detect_bodies<<<grid_for(c.S, kDetectBlock), kDetectBlock, 0, s>>>(c, b, x);
}

void launch_body_candidates(hipStream_t s, contact_dev c, body_contact_dev const& b, const pos4* x,
                            const uint32_t* offsets, uint32_t* counts, uint32_t* out)
{
    if (c.S) k_body_candidates<<<grid_for(c.S, 256), 256, 0, s>>>(c, b, x, offsets, counts, out);
}

}  // namespace sbs
