// Grid SDF construction on gfx950: the signed distance of a triangle mesh at every node of a cubic
// Lagrange grid (Discregrid::MeshDistance::signedDistanceCached evaluated by
// CubicLagrangeDiscreteGrid::addFunction, environment_body.cpp:66-73).
//
// Distance: the closest point of every triangle (Voronoi-region test of the triangle's vertices,
// edges and face; Ericson, "Real-Time Collision Detection", 5.1.5), the smallest squared distance
// winning (the first triangle on ties). Sign: negative where (x - closest point) points against the
// angle-weighted pseudonormal of the closest feature (Baerentzen & Aanaes 2005): the vertex normal
// sum of incident face normals weighted by the incident angles, the edge normal sum of the normals
// of the faces sharing the edge, or the face normal. Zero-area triangles are left out.
//
// One thread per grid node; the triangles stream through LDS in tiles shared by the 256 threads of
// a block, tiles that cannot hold a nearer triangle for any node of the block skipped (below). The
// work is FP64 arithmetic (about 60 operations per node and triangle), bounded by the FP64 vector
// rate, not by memory.
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <utility>
#include <vector>

#include "abi_util.h"
#include "grid_sdf.h"
#include "sdf.h"

namespace sbs {

namespace {

constexpr int kBlock = 256, kTile = 128;

struct grid_params {
    uint32_t res[3];
    double dmin[3], cell[3];
};

// Squared distance from p to the box lo..hi (0 inside).
__device__ __forceinline__ double box_dist2(const double p[3], const double* box)
{
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double const e = p[d] < box[d] ? box[d] - p[d] : (p[d] > box[3 + d] ? p[d] - box[3 + d] : 0.0);
        s += e * e;
    }
    return s;
}

// One thread per node. The triangles are in Morton order of their centroids, in tiles of kTile
// with bounding boxes. A block starts at the tile nearest its first node and visits the others
// cyclically; a tile is loaded into LDS only when some node of the block may have a nearer
// triangle in it (box distance^2 <= best * (1 + 1e-9), a margin well above the rounding of both
// sides), and a node skips the tile otherwise. The nearest triangle is the smallest (squared
// distance, original index) pair, so the visiting order does not change the result: it equals the
// scan of every triangle in index order keeping the first nearest.
__global__ void __launch_bounds__(kBlock) k_grid_nodes(mesh_dev mesh, grid_params gp, uint64_t n,
                                                       double* __restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ double tc[kTile * 9];
    __shared__ uint32_t to[kTile];
    __shared__ double red_d[kBlock];
    __shared__ uint32_t red_i[kBlock];
    uint64_t const l = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    bool const valid = l < n;
    grid_dims const g(gp.res);
    double p[3] = {0.0, 0.0, 0.0};
    if (valid) grid_node_position(g, gp.dmin, gp.cell, l, p);
    // the tile nearest the block's first node
    double p0[3];
    grid_node_position(g, gp.dmin, gp.cell, uint64_t(blockIdx.x) * kBlock, p0);
    double md = DBL_MAX;
    uint32_t mi = 0;
    for (uint32_t i = threadIdx.x; i < mesh.ntiles; i += kBlock) {
        double const dd = box_dist2(p0, mesh.tile_box + 6 * std::size_t(i));
        if (dd < md) {
            md = dd;
            mi = i;
        }
    }
    red_d[threadIdx.x] = md;
    red_i[threadIdx.x] = mi;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            double const od = red_d[threadIdx.x + w];
            uint32_t const oi = red_i[threadIdx.x + w];
            if (od < red_d[threadIdx.x] || (od == red_d[threadIdx.x] && oi < red_i[threadIdx.x])) {
                red_d[threadIdx.x] = od;
                red_i[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    uint32_t const start = red_i[0];
    double best = DBL_MAX, bq[3] = {0.0, 0.0, 0.0};
    uint32_t bf = 0xffffffffu;
    int bfeat = 6;
    for (uint32_t u = 0; u < mesh.ntiles; ++u) {
        uint32_t const tile = start + u < mesh.ntiles ? start + u : start + u - mesh.ntiles;
        bool const need = valid && box_dist2(p, mesh.tile_box + 6 * std::size_t(tile)) <= best * (1.0 + 1e-9);
        if (!__syncthreads_or(need)) continue;
        uint32_t const t0 = tile * kTile, cnt = min(uint32_t(kTile), mesh.m - t0);
        for (uint32_t w = threadIdx.x; w < cnt * 9; w += kBlock) tc[w] = mesh.corners[std::size_t(t0) * 9 + w];
        for (uint32_t w = threadIdx.x; w < cnt; w += kBlock) to[w] = mesh.orig[t0 + w];
        __syncthreads();
        if (need)
            for (uint32_t t = 0; t < cnt; ++t) {
                double q[3];
                int feat;
                tri_closest_point(p, tc + 9 * t, q, &feat);
                double const e0 = p[0] - q[0], e1 = p[1] - q[1], e2 = p[2] - q[2];
                double const d2 = e0 * e0 + e1 * e1 + e2 * e2;
                uint32_t const o = to[t];
                if (d2 < best || (d2 == best && o < bf)) {
                    best = d2;
                    bf = o;
                    bfeat = feat;
                    bq[0] = q[0];
                    bq[1] = q[1];
                    bq[2] = q[2];
                }
            }
        __syncthreads();
    }
    if (!valid) return;
    if (mesh.m == 0) {
        out[l] = DBL_MAX;
        return;
    }
    const double* nrm = bfeat < 3   ? mesh.vnormal + 3 * std::size_t(mesh.tri[3 * std::size_t(bf) + bfeat])
                        : bfeat < 6 ? mesh.enormal + 9 * std::size_t(bf) + 3 * (bfeat - 3)
                                    : mesh.fnormal + 3 * std::size_t(bf);
    double const s = (p[0] - bq[0]) * nrm[0] + (p[1] - bq[1]) * nrm[1] + (p[2] - bq[2]) * nrm[2];
    double const d = sqrt(best);
    out[l] = s < 0.0 ? -d : d;
}

}  // namespace

void launch_grid_nodes(hipStream_t s, mesh_dev mesh, const uint32_t res[3], const double dmin[3],
                       const double cell[3], double* nodes)
{
    grid_params gp{};
    for (int d = 0; d < 3; ++d) {
        gp.res[d] = res[d];
        gp.dmin[d] = dmin[d];
        gp.cell[d] = cell[d];
    }
    uint64_t const n = grid_dims(res).nodes();
    k_grid_nodes<<<unsigned((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(mesh, gp, n, nodes);
}

// Host side of the mesh: non-degenerate triangles, corners and pseudonormals (FP64, no FMA
// contraction), uploaded to the device.
struct mesh_upload {
    dev_buf<uint32_t> tri, orig;
    dev_buf<double> corners, vnormal, enormal, fnormal, tile_box;
    uint32_t m = 0, ntiles = 0;

    mesh_upload(hipStream_t s, const double* pos, uint32_t nv, const uint32_t* tris, uint32_t nt)
    {
#pragma clang fp contract(off)
        auto at = [&](uint32_t v, int d) { return pos[3 * std::size_t(v) + d]; };
        auto cross = [](const double a[3], const double b[3], double o[3]) {
            o[0] = a[1] * b[2] - a[2] * b[1];
            o[1] = a[2] * b[0] - a[0] * b[2];
            o[2] = a[0] * b[1] - a[1] * b[0];
        };
        auto dot = [](const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        std::vector<uint32_t> t;
        for (uint32_t f = 0; f < nt; ++f) {
            double e1[3], e2[3], n[3];
            for (int d = 0; d < 3; ++d) {
                e1[d] = at(tris[3 * f + 1], d) - at(tris[3 * f], d);
                e2[d] = at(tris[3 * f + 2], d) - at(tris[3 * f], d);
            }
            cross(e1, e2, n);
            if (dot(n, n) > 0.0)
                for (int k = 0; k < 3; ++k) t.push_back(tris[3 * f + k]);
        }
        m = uint32_t(t.size() / 3);
        std::vector<double> c(9 * std::size_t(m)), vn(3 * std::size_t(nv), 0.0), en(9 * std::size_t(m)),
            fn(3 * std::size_t(m));
        std::map<std::pair<uint32_t, uint32_t>, std::array<double, 3>> edge_sum;
        for (uint32_t f = 0; f < m; ++f) {
            double p[3][3];
            for (int k = 0; k < 3; ++k)
                for (int d = 0; d < 3; ++d) {
                    p[k][d] = at(t[3 * f + k], d);
                    c[9 * std::size_t(f) + 3 * k + d] = p[k][d];
                }
            double e1[3], e2[3], n[3];
            for (int d = 0; d < 3; ++d) {
                e1[d] = p[1][d] - p[0][d];
                e2[d] = p[2][d] - p[0][d];
            }
            cross(e1, e2, n);
            double const ln = std::sqrt(dot(n, n));
            for (int d = 0; d < 3; ++d) {
                n[d] = n[d] / ln;
                fn[3 * std::size_t(f) + d] = n[d];
            }
            for (int k = 0; k < 3; ++k) {
                double a1[3], a2[3];
                for (int d = 0; d < 3; ++d) {
                    a1[d] = p[(k + 1) % 3][d] - p[k][d];
                    a2[d] = p[(k + 2) % 3][d] - p[k][d];
                }
                double const cs = dot(a1, a2) / (std::sqrt(dot(a1, a1)) * std::sqrt(dot(a2, a2)));
                double const alpha = std::acos(std::max(-1.0, std::min(1.0, cs)));
                uint32_t const v = t[3 * f + k];
                for (int d = 0; d < 3; ++d) vn[3 * std::size_t(v) + d] = vn[3 * std::size_t(v) + d] + alpha * n[d];
                uint32_t const w = t[3 * f + (k + 1) % 3];
                auto& es = edge_sum.try_emplace({std::min(v, w), std::max(v, w)}, std::array<double, 3>{0, 0, 0})
                               .first->second;
                for (int d = 0; d < 3; ++d) es[d] = es[d] + n[d];
            }
        }
        for (uint32_t f = 0; f < m; ++f)
            for (int k = 0; k < 3; ++k) {
                uint32_t const v = t[3 * f + k], w = t[3 * f + (k + 1) % 3];
                auto const& es = edge_sum[{std::min(v, w), std::max(v, w)}];
                for (int d = 0; d < 3; ++d) en[9 * std::size_t(f) + 3 * k + d] = es[d];
            }
        // triangles in Morton order of their centroids (10 bits per axis over the mesh's box),
        // corners in that order, each with its original index; bounding boxes of tiles of kTile
        std::vector<uint32_t> ord(m);
        {
            double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
            for (std::size_t i = 0; i < 9 * std::size_t(m); ++i) {
                lo[i % 3] = std::min(lo[i % 3], c[i]);
                hi[i % 3] = std::max(hi[i % 3], c[i]);
            }
            auto spread = [](uint32_t v) {
                uint64_t x = v & 0x3ffu;
                x = (x | (x << 16)) & 0x30000ffull;
                x = (x | (x << 8)) & 0x300f00full;
                x = (x | (x << 4)) & 0x30c30c3ull;
                x = (x | (x << 2)) & 0x9249249ull;
                return x;
            };
            std::vector<std::pair<uint64_t, uint32_t>> key(m);
            for (uint32_t f = 0; f < m; ++f) {
                uint64_t code = 0;
                for (int d = 0; d < 3; ++d) {
                    double const cen = (c[9 * std::size_t(f) + d] + c[9 * std::size_t(f) + 3 + d] +
                                        c[9 * std::size_t(f) + 6 + d]) / 3.0;
                    double const ext = hi[d] - lo[d];
                    double const u = ext > 0.0 ? (cen - lo[d]) / ext : 0.0;
                    uint32_t const q = uint32_t(std::min(1023.0, std::max(0.0, u * 1024.0)));
                    code |= spread(q) << d;
                }
                key[f] = {code, f};
            }
            std::sort(key.begin(), key.end());
            for (uint32_t i = 0; i < m; ++i) ord[i] = key[i].second;
        }
        std::vector<double> cs(9 * std::size_t(m));
        for (uint32_t i = 0; i < m; ++i)
            std::copy(c.begin() + 9 * std::size_t(ord[i]), c.begin() + 9 * std::size_t(ord[i]) + 9,
                      cs.begin() + 9 * std::size_t(i));
        ntiles = (m + kTile - 1) / kTile;
        std::vector<double> box(6 * std::size_t(ntiles));
        for (uint32_t tl = 0; tl < ntiles; ++tl) {
            double* b = box.data() + 6 * std::size_t(tl);
            b[0] = b[1] = b[2] = DBL_MAX;
            b[3] = b[4] = b[5] = -DBL_MAX;
            for (uint32_t i = tl * kTile; i < std::min(m, (tl + 1) * kTile); ++i)
                for (int k = 0; k < 9; ++k) {
                    b[k % 3] = std::min(b[k % 3], cs[9 * std::size_t(i) + k]);
                    b[3 + k % 3] = std::max(b[3 + k % 3], cs[9 * std::size_t(i) + k]);
                }
        }
        tri.upload(t.data(), t.size(), s);
        corners.upload(cs.data(), cs.size(), s);
        orig.upload(ord.data(), ord.size(), s);
        tile_box.upload(box.data(), box.size(), s);
        vnormal.upload(vn.data(), vn.size(), s);
        enormal.upload(en.data(), en.size(), s);
        fnormal.upload(fn.data(), fn.size(), s);
        hip_check(hipStreamSynchronize(s), "mesh upload");
    }
    mesh_dev view() const { return {tri.p, corners.p, vnormal.p, enormal.p, fnormal.p, orig.p, tile_box.p, m, ntiles}; }
};

void grid_cell_size(const double dmin[3], const double dmax[3], const uint32_t res[3], double cell[3],
                    double inv_cell[3])
{
    for (int d = 0; d < 3; ++d) {
        cell[d] = (dmax[d] - dmin[d]) / double(res[d]);
        inv_cell[d] = 1.0 / cell[d];
    }
}

void check_grid(const double dmin[3], const double dmax[3], const uint32_t res[3])
{
    require(dmin && dmax && res, "null grid argument");
    for (int d = 0; d < 3; ++d) {
        require(res[d] >= 1, "grid resolution must be at least 1 cell per axis");
        require(std::isfinite(dmin[d]) && std::isfinite(dmax[d]) && dmax[d] > dmin[d],
                "grid domain must be a finite box with max > min");
    }
    require(grid_dims(res).nodes() < (uint64_t(1) << 32), "grid has too many nodes");
}

void build_mesh_grid(hipStream_t s, const double* pos, uint32_t nv, const uint32_t* tris, uint32_t nt,
                     const double dmin[3], const double dmax[3], const uint32_t res[3], double* dev_nodes)
{
    require((pos || nv == 0) && (tris || nt == 0), "null mesh argument");
    for (std::size_t i = 0; i < 3 * std::size_t(nt); ++i) require(tris[i] < nv, "triangle vertex index out of range");
    check_grid(dmin, dmax, res);
    mesh_upload const mesh(s, pos, nv, tris, nt);
    double cell[3], inv_cell[3];
    grid_cell_size(dmin, dmax, res, cell, inv_cell);
    launch_grid_nodes(s, mesh.view(), res, dmin, cell, dev_nodes);
    hip_check(hipGetLastError(), "grid node kernel");
    hip_check(hipStreamSynchronize(s), "grid node kernel");
}

void environment_domain(const double* pos, uint32_t nv, const double lo_in[3], const double hi_in[3], double lo[3],
                        double hi[3])
{
#pragma clang fp contract(off)
    for (int d = 0; d < 3; ++d) {
        lo[d] = lo_in[d];
        hi[d] = hi_in[d];
    }
    for (uint32_t i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], pos[3 * std::size_t(i) + d]);
            hi[d] = std::max(hi[d], pos[3 * std::size_t(i) + d]);
        }
    // environment_body.cpp:54-64: once per vertex, max grows by 1e-3 |diagonal|, then min by
    // 1e-3 |diagonal| of the grown box (re-extending by the vertices inside the loop changes nothing
    // after the first pass)
    auto diag_norm = [&] {
        double const a = hi[0] - lo[0], b = hi[1] - lo[1], c = hi[2] - lo[2];
        return std::sqrt(a * a + b * b + c * c);
    };
    for (uint32_t v = 0; v < nv; ++v) {
        double const g1 = 1.0e-3 * diag_norm();
        for (int d = 0; d < 3; ++d) hi[d] += g1 * 1.0;
        double const g2 = 1.0e-3 * diag_norm();
        for (int d = 0; d < 3; ++d) lo[d] -= g2 * 1.0;
    }
}

namespace {

__global__ void k_eval_sdf(sdf_dev f, const double* __restrict__ pts, uint32_t n, double* __restrict__ sd,
                           double* __restrict__ grad)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos4 const p = make_double4(pts[3 * std::size_t(i)], pts[3 * std::size_t(i) + 1], pts[3 * std::size_t(i) + 2], 0.0);
    double v, g[3];
    sdf_eval(f, p, v, g);
    sd[i] = v;
    grad[3 * std::size_t(i)] = g[0];
    grad[3 * std::size_t(i) + 1] = g[1];
    grad[3 * std::size_t(i) + 2] = g[2];
}

}  // namespace

void evaluate_sdf(hipStream_t s, sdf_dev const& f, const double* pts, uint32_t n, double* sd, double* grad)
{
    if (n == 0) return;
    dev_buf<double> dp, dsd, dg;
    dp.upload(pts, 3 * std::size_t(n), s);
    dsd.alloc(n);
    dg.alloc(3 * std::size_t(n));
    k_eval_sdf<<<(n + 255) / 256, 256, 0, s>>>(f, dp.p, n, dsd.p, dg.p);
    hip_check(hipGetLastError(), "sdf evaluation");
    hip_check(hipMemcpyAsync(sd, dsd.p, n * sizeof(double), hipMemcpyDeviceToHost, s), "read sdf values");
    hip_check(hipMemcpyAsync(grad, dg.p, 3 * std::size_t(n) * sizeof(double), hipMemcpyDeviceToHost, s),
              "read sdf gradients");
    hip_check(hipStreamSynchronize(s), "sdf evaluation");
}

}  // namespace sbs

using namespace sbs;

extern "C" int sbs_sdf_grid_node_count(const uint32_t resolution[3], uint64_t* count)
{
    return guarded([&] {
        require(resolution && count, "null argument");
        *count = grid_dims(resolution).nodes();
    });
}

extern "C" int sbs_sdf_grid_node_positions(const double domain_min[3], const double domain_max[3],
                                           const uint32_t resolution[3], double* positions)
{
    return guarded([&] {
        check_grid(domain_min, domain_max, resolution);
        require(positions != nullptr, "positions is null");
        double cell[3], inv_cell[3];
        grid_cell_size(domain_min, domain_max, resolution, cell, inv_cell);
        grid_dims const g(resolution);
        for (uint64_t l = 0; l < g.nodes(); ++l) grid_node_position(g, domain_min, cell, l, positions + 3 * l);
    });
}

extern "C" int sbs_sdf_environment_domain(const double* positions, uint32_t num_vertices, const double domain_min[3],
                                          const double domain_max[3], double out_min[3], double out_max[3])
{
    return guarded([&] {
        require((positions || num_vertices == 0) && domain_min && domain_max && out_min && out_max, "null argument");
        environment_domain(positions, num_vertices, domain_min, domain_max, out_min, out_max);
    });
}

extern "C" int sbs_sdf_grid_from_mesh(int device, const double* positions, uint32_t num_vertices,
                                      const uint32_t* triangles, uint32_t num_triangles, const double domain_min[3],
                                      const double domain_max[3], const uint32_t resolution[3], double* nodes)
{
    return guarded([&] {
        require(nodes != nullptr, "nodes is null");
        check_grid(domain_min, domain_max, resolution);
        hip_check(hipSetDevice(device), "hipSetDevice");
        hipStream_t s;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        try {
            uint64_t const n = grid_dims(resolution).nodes();
            dev_buf<double> dn;
            dn.alloc(n);
            build_mesh_grid(s, positions, num_vertices, triangles, num_triangles, domain_min, domain_max, resolution,
                            dn.p);
            hip_check(hipMemcpyAsync(nodes, dn.p, n * sizeof(double), hipMemcpyDeviceToHost, s), "read grid nodes");
            hip_check(hipStreamSynchronize(s), "read grid nodes");
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        hip_check(hipStreamDestroy(s), "hipStreamDestroy");
    });
}
