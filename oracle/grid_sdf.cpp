// Test infrastructure only: the checker's grid SDFs (see grid_sdf.h for what is restated and how
// it is pinned).
#include "grid_sdf.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <map>
#include <utility>


namespace oracle {

void grid_sdf_t::init(vec3 lo, vec3 hi, const uint32_t r[3])
{
    dmin = lo;
    dmax = hi;
    for (int d = 0; d < 3; ++d) {
        res[d] = r[d];
        cell[d] = (hi[d] - lo[d]) / double(r[d]);  // domain.diagonal() / resolution
        inv_cell[d] = 1.0 / cell[d];
    }
    nodes.assign(node_count(), 0.0);
}

uint64_t grid_sdf_t::node_count() const
{
    uint64_t const nx = res[0], ny = res[1], nz = res[2];
    uint64_t const nv = (nx + 1) * (ny + 1) * (nz + 1);
    uint64_t const ne = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz;
    return nv + 2 * ne;
}

// Node numbering: corners (i fastest, then j, then k); then the x edges, two nodes each (at 1/3
// and 2/3 of the edge), edge (i, j, k) numbered i fastest, then j, then k; then the y edges, j
// fastest, then k, then i; then the z edges, k fastest, then i, then j.
vec3 grid_sdf_t::node_position(uint64_t l) const
{
    uint64_t const nx = res[0], ny = res[1], nz = res[2];
    uint64_t const nv = (nx + 1) * (ny + 1) * (nz + 1);
    uint64_t const nex = nx * (ny + 1) * (nz + 1), ney = (nx + 1) * ny * (nz + 1);
    uint64_t ijk[3];
    int axis = -1;
    uint64_t half = 0;
    if (l < nv) {
        ijk[2] = l / ((ny + 1) * (nx + 1));
        uint64_t const t = l % ((ny + 1) * (nx + 1));
        ijk[1] = t / (nx + 1);
        ijk[0] = t % (nx + 1);
    } else if (l < nv + 2 * nex) {
        uint64_t const m = l - nv, e = m / 2;
        half = m % 2;
        axis = 0;
        ijk[2] = e / ((ny + 1) * nx);
        uint64_t const t = e % ((ny + 1) * nx);
        ijk[1] = t / nx;
        ijk[0] = t % nx;
    } else if (l < nv + 2 * (nex + ney)) {
        uint64_t const m = l - nv - 2 * nex, e = m / 2;
        half = m % 2;
        axis = 1;
        ijk[0] = e / ((nz + 1) * ny);
        uint64_t const t = e % ((nz + 1) * ny);
        ijk[2] = t / ny;
        ijk[1] = t % ny;
    } else {
        uint64_t const m = l - nv - 2 * (nex + ney), e = m / 2;
        half = m % 2;
        axis = 2;
        ijk[1] = e / ((nx + 1) * nz);
        uint64_t const t = e % ((nx + 1) * nz);
        ijk[0] = t / nz;
        ijk[2] = t % nz;
    }
    vec3 x;
    for (int d = 0; d < 3; ++d) x[d] = dmin[d] + cell[d] * double(ijk[d]);
    if (axis >= 0) x[axis] += (1.0 + double(half)) / 3.0 * cell[axis];
    return x;
}

void grid_sdf_t::cell_nodes(uint32_t i, uint32_t j, uint32_t k, uint64_t out[32]) const
{
    uint64_t const nx = res[0], ny = res[1], nz = res[2];
    uint64_t const nv = (nx + 1) * (ny + 1) * (nz + 1);
    uint64_t const nex = nx * (ny + 1) * (nz + 1), ney = (nx + 1) * ny * (nz + 1);
    auto corner = [&](uint64_t a, uint64_t b, uint64_t c) { return (nx + 1) * (ny + 1) * c + (nx + 1) * b + a; };
    auto xe = [&](uint64_t a, uint64_t b, uint64_t c) { return nv + 2 * (nx * (ny + 1) * c + nx * b + a); };
    auto ye = [&](uint64_t a, uint64_t b, uint64_t c) { return nv + 2 * nex + 2 * (ny * (nz + 1) * a + ny * c + b); };
    auto ze = [&](uint64_t a, uint64_t b, uint64_t c) {
        return nv + 2 * (nex + ney) + 2 * (nz * (nx + 1) * b + nz * a + c);
    };
    // corners in shape-function order: x fastest, then y, then z
    for (int c = 0; c < 8; ++c) out[c] = corner(i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2));
    // x edges at (y, z) = (j, k), (j, k+1), (j+1, k), (j+1, k+1)
    uint64_t const ex[4] = {xe(i, j, k), xe(i, j, k + 1), xe(i, j + 1, k), xe(i, j + 1, k + 1)};
    // y edges at (x, z) = (i, k), (i+1, k), (i, k+1), (i+1, k+1)
    uint64_t const ey[4] = {ye(i, j, k), ye(i + 1, j, k), ye(i, j, k + 1), ye(i + 1, j, k + 1)};
    // z edges at (x, y) = (i, j), (i, j+1), (i+1, j), (i+1, j+1)
    uint64_t const ez[4] = {ze(i, j, k), ze(i, j + 1, k), ze(i + 1, j, k), ze(i + 1, j + 1, k)};
    for (int e = 0; e < 4; ++e) {
        out[8 + 2 * e] = ex[e];
        out[9 + 2 * e] = ex[e] + 1;
        out[16 + 2 * e] = ey[e];
        out[17 + 2 * e] = ey[e] + 1;
        out[24 + 2 * e] = ez[e];
        out[25 + 2 * e] = ez[e] + 1;
    }
}

// Corner node c at (sx, sy, sz) = (+-1, +-1, +-1), bit 0 of c: x, bit 1: y, bit 2: z:
//   N = (9 (x^2 + y^2 + z^2) - 19) / 64 (1 + sx x)(1 + sy y)(1 + sz z).
// Edge node at xi_a = -1/3 (first) or +1/3 (second) of an edge along axis a, the other two
// coordinates at +-1:  N = 9/64 (1 - xi_a^2)(1 -+ 3 xi_a)(1 + s_b xi_b)(1 + s_c xi_c).
void shape_functions(vec3 xi, double N[32], double dN[32][3])
{
    double const x = xi.x, y = xi.y, z = xi.z;
    double const x2 = x * x, y2 = y * y, z2 = z * z;
    double const mx = 1.0 - x, my = 1.0 - y, mz = 1.0 - z;
    double const px = 1.0 + x, py = 1.0 + y, pz = 1.0 + z;
    double const fx[2] = {mx, px}, fy[2] = {my, py}, fz[2] = {mz, pz};
    double const sg[2] = {-1.0, 1.0};
    double const r = 9.0 * (x2 + y2 + z2) - 19.0;
    double const fac = 1.0 / 64.0 * r;
    for (int c = 0; c < 8; ++c) {
        int const a = c & 1, b = (c >> 1) & 1, e = c >> 2;
        N[c] = fac * fx[a] * fy[b] * fz[e];
        // d/dx [r (1 + sx x)] = 18 x (1 + sx x) + sx r
        dN[c][0] = 1.0 / 64.0 * (18.0 * x * fx[a] + sg[a] * r) * fy[b] * fz[e];
        dN[c][1] = 1.0 / 64.0 * (18.0 * y * fy[b] + sg[b] * r) * fx[a] * fz[e];
        dN[c][2] = 1.0 / 64.0 * (18.0 * z * fz[e] + sg[e] * r) * fx[a] * fy[b];
    }
    // edges along axis a: u = xi_a, (v, w) the other two coordinates with their factor pairs
    for (int ax = 0; ax < 3; ++ax) {
        double const u = ax == 0 ? x : (ax == 1 ? y : z);
        double const u2 = u * u;
        double const g = 9.0 / 64.0 * (1.0 - u2);
        double const lin[2] = {1.0 - 3.0 * u, 1.0 + 3.0 * u};
        double const dlin[2] = {-3.0, 3.0};
        // the 4 edges: (v side, w side) for shape-function slots 0..3 of this axis
        // x edges: (y, z) = (-,-), (-,+), (+,-), (+,+); y edges: (x, z) = (-,-), (+,-), (-,+), (+,+);
        // z edges: (x, y) = (-,-), (-,+), (+,-), (+,+)
        static const int side[3][4][2] = {{{0, 0}, {0, 1}, {1, 0}, {1, 1}},
                                          {{0, 0}, {1, 0}, {0, 1}, {1, 1}},
                                          {{0, 0}, {0, 1}, {1, 0}, {1, 1}}};
        int const av = ax == 0 ? 1 : 0, aw = ax == 2 ? 1 : 2;  // the other axes in increasing order
        double const* fv = av == 0 ? fx : fy;
        double const* fw = aw == 1 ? fy : fz;
        for (int e = 0; e < 4; ++e) {
            int const sv = side[ax][e][0], sw = side[ax][e][1];
            double const vw = fv[sv] * fw[sw];
            for (int h = 0; h < 2; ++h) {
                int const j = 8 + 8 * ax + 2 * e + h;
                N[j] = g * lin[h] * vw;
                double d[3];
                // d/du [(1 - u^2)(1 +- 3u)] = -2u (1 +- 3u) +- 3 (1 - u^2)
                d[ax] = 9.0 / 64.0 * (-2.0 * u * lin[h] + dlin[h] * (1.0 - u2)) * vw;
                d[av] = g * lin[h] * sg[sv] * fw[sw];
                d[aw] = g * lin[h] * fv[sv] * sg[sw];
                dN[j][0] = d[0];
                dN[j][1] = d[1];
                dN[j][2] = d[2];
            }
        }
    }
}

double grid_sdf_t::interpolate(vec3 x, vec3* grad) const
{
    for (int d = 0; d < 3; ++d)
        if (!(dmin[d] <= x[d] && x[d] <= dmax[d])) {
            if (grad) *grad = vec3{0.0, 0.0, 0.0};
            return DBL_MAX;
        }
    uint32_t mi[3];
    vec3 lo, hi;
    for (int d = 0; d < 3; ++d) {
        mi[d] = uint32_t((x[d] - dmin[d]) * inv_cell[d]);
        if (mi[d] >= res[d]) mi[d] = res[d] - 1;
        lo[d] = dmin[d] + double(mi[d]) * cell[d];
        hi[d] = lo[d] + cell[d];
    }
    vec3 xi, c0;
    for (int d = 0; d < 3; ++d) {
        double const den = hi[d] - lo[d];
        c0[d] = 2.0 * (1.0 / den);
        double const c1 = (hi[d] + lo[d]) / den;
        xi[d] = c0[d] * x[d] - c1;
    }
    uint64_t cn[32];
    cell_nodes(mi[0], mi[1], mi[2], cn);
    double N[32], dN[32][3];
    shape_functions(xi, N, dN);
    double phi = 0.0;
    vec3 g{0.0, 0.0, 0.0};
    for (int j = 0; j < 32; ++j) {
        double const c = nodes[cn[j]];
        phi += c * N[j];
        g.x += c * dN[j][0];
        g.y += c * dN[j][1];
        g.z += c * dN[j][2];
    }
    if (grad) *grad = vec3{g.x * c0.x, g.y * c0.y, g.z * c0.z};
    return phi;
}

vec3 closest_point_on_triangle(vec3 p, vec3 a, vec3 b, vec3 c, int* feature, double* bary)
{
    vec3 const ab = b - a, ac = c - a, ap = p - a;
    double const d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        *feature = 0;
        if (bary) bary[0] = 1.0, bary[1] = 0.0, bary[2] = 0.0;
        return a;
    }
    vec3 const bp = p - b;
    double const d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        *feature = 1;
        if (bary) bary[0] = 0.0, bary[1] = 1.0, bary[2] = 0.0;
        return b;
    }
    double const vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double const v = d1 / (d1 - d3);
        *feature = 3;
        if (bary) bary[0] = 1.0 - v, bary[1] = v, bary[2] = 0.0;
        return a + v * ab;
    }
    vec3 const cp = p - c;
    double const d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        *feature = 2;
        if (bary) bary[0] = 0.0, bary[1] = 0.0, bary[2] = 1.0;
        return c;
    }
    double const vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double const w = d2 / (d2 - d6);
        *feature = 5;
        if (bary) bary[0] = 1.0 - w, bary[1] = 0.0, bary[2] = w;
        return a + w * ac;
    }
    double const va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double const w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *feature = 4;
        if (bary) bary[0] = 0.0, bary[1] = 1.0 - w, bary[2] = w;
        return b + w * (c - b);
    }
    double const den = 1.0 / (va + vb + vc);
    double const v = vb * den, w = vc * den;
    *feature = 6;
    if (bary) bary[0] = 1.0 - v - w, bary[1] = v, bary[2] = w;
    return a + ab * v + ac * w;
}

mesh_distance_t::mesh_distance_t(const double* pos, uint32_t nv, const uint32_t* tris, uint32_t nt)
{
    x.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) x[i] = vec3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    // zero-area triangles have no normal and no interior: they are left out
    for (uint32_t f = 0; f < nt; ++f) {
        vec3 const n = cross(x[tris[3 * f + 1]] - x[tris[3 * f]], x[tris[3 * f + 2]] - x[tris[3 * f]]);
        if (norm2(n) > 0.0)
            for (int k = 0; k < 3; ++k) tri.push_back(tris[3 * f + k]);
    }
    uint32_t const m = uint32_t(tri.size() / 3);
    face_normal.resize(m);
    vertex_normal.assign(nv, vec3{0.0, 0.0, 0.0});
    std::map<std::pair<uint32_t, uint32_t>, vec3> edge_sum;
    for (uint32_t f = 0; f < m; ++f) {
        vec3 const a = x[tri[3 * f]], b = x[tri[3 * f + 1]], c = x[tri[3 * f + 2]];
        vec3 n = cross(b - a, c - a);
        n = n / norm(n);
        face_normal[f] = n;
        vec3 const p[3] = {a, b, c};
        for (int k = 0; k < 3; ++k) {
            vec3 const e1 = p[(k + 1) % 3] - p[k], e2 = p[(k + 2) % 3] - p[k];
            double const cs = dot(e1, e2) / (norm(e1) * norm(e2));
            double const alpha = std::acos(std::max(-1.0, std::min(1.0, cs)));
            vertex_normal[tri[3 * f + k]] += alpha * n;
            uint32_t const u = tri[3 * f + k], v = tri[3 * f + (k + 1) % 3];
            edge_sum[{std::min(u, v), std::max(u, v)}] += n;
        }
    }
    edge_normal.resize(3 * size_t(m));
    for (uint32_t f = 0; f < m; ++f)
        for (int k = 0; k < 3; ++k) {
            uint32_t const u = tri[3 * f + k], v = tri[3 * f + (k + 1) % 3];
            edge_normal[3 * size_t(f) + k] = edge_sum[{std::min(u, v), std::max(u, v)}];
        }
}

double mesh_distance_t::signed_distance(vec3 p) const
{
    double best = DBL_MAX;
    uint32_t bf = 0;
    int bfeat = 6;
    vec3 bq;
    uint32_t const m = uint32_t(tri.size() / 3);
    for (uint32_t f = 0; f < m; ++f) {
        int feat;
        vec3 const q = closest_point_on_triangle(p, x[tri[3 * f]], x[tri[3 * f + 1]], x[tri[3 * f + 2]], &feat);
        double const d2 = norm2(p - q);
        if (d2 < best) {
            best = d2;
            bf = f;
            bfeat = feat;
            bq = q;
        }
    }
    if (m == 0) return DBL_MAX;
    vec3 const n = bfeat < 3   ? vertex_normal[tri[3 * bf + bfeat]]
                   : bfeat < 6 ? edge_normal[3 * size_t(bf) + (bfeat - 3)]
                               : face_normal[bf];
    double const d = std::sqrt(best);
    return dot(p - bq, n) < 0.0 ? -d : d;
}

void environment_domain(const double* pos, uint32_t nv, vec3 lo, vec3 hi, vec3* out_lo, vec3* out_hi)
{
    // environment_body.cpp:52-64: for every vertex, extend by all vertices (a no-op after the
    // first pass), then grow max by 1e-3 |diagonal|, then min by 1e-3 |diagonal| (recomputed)
    for (uint32_t i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], pos[3 * i + d]);
            hi[d] = std::max(hi[d], pos[3 * i + d]);
        }
    for (uint32_t v = 0; v < nv; ++v) {
        vec3 diag = hi - lo;
        double const g1 = 1.0e-3 * std::sqrt(diag.x * diag.x + diag.y * diag.y + diag.z * diag.z);
        for (int d = 0; d < 3; ++d) hi[d] += g1 * 1.0;
        diag = hi - lo;
        double const g2 = 1.0e-3 * std::sqrt(diag.x * diag.x + diag.y * diag.y + diag.z * diag.z);
        for (int d = 0; d < 3; ++d) lo[d] -= g2 * 1.0;
    }
    *out_lo = lo;
    *out_hi = hi;
}

}  // namespace oracle
