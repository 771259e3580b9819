// Grid signed distance fields: the cubic Lagrange grid of Discregrid (CubicLagrangeDiscreteGrid,
// used by sdf_model.cpp:66-75 and built by environment_body.cpp:66-76) -- a box split into
// res[0] x res[1] x res[2] cells, each a 32-node cubic serendipity element with nodes at its 8
// corners and at 1/3 and 2/3 of its 12 edges, nodes shared between neighbouring cells.
//
// Node numbering: corners (i fastest, then j, then k); then x edges, two nodes each (at 1/3 and 2/3
// of the edge), edge (i, j, k) with i fastest, then j, then k; then y edges with j fastest, then k,
// then i; then z edges with k fastest, then i, then j. Node l lies at
// domain_min + cell * (i, j, k), plus (1 + h) / 3 * cell along its edge's axis for edge node h.
//
// Shape functions at local coordinates (x, y, z) in [-1, 1]^3:
//   corner (sx, sy, sz): (9 (x^2 + y^2 + z^2) - 19) / 64 (1 + sx x)(1 + sy y)(1 + sz z)
//   edge node at u = -1/3 or +1/3 of an edge along axis u, the others at sides sv, sw:
//                        9/64 (1 - u^2)(1 -+ 3 u)(1 + sv v)(1 + sw w)
// Value and gradient are the node values weighted by the shape functions and their derivatives
// (scaled by 2 / cell); outside the domain the value is DBL_MAX and the gradient zero.
// All of this is FP64 without contraction into FMAs.
#pragma once
#include <cfloat>
#include <cstdint>

#include "kernels.h"

namespace sbs {

struct grid_dims {
    uint64_t nx, ny, nz, nv, nex, ney, nez;
    __host__ __device__ explicit grid_dims(const uint32_t r[3])
        : nx(r[0]), ny(r[1]), nz(r[2]), nv((nx + 1) * (ny + 1) * (nz + 1)), nex(nx * (ny + 1) * (nz + 1)),
          ney((nx + 1) * ny * (nz + 1)), nez((nx + 1) * (ny + 1) * nz)
    {
    }
    __host__ __device__ uint64_t nodes() const { return nv + 2 * (nex + ney + nez); }
};

__host__ __device__ inline void grid_node_position(grid_dims const& g, const double dmin[3], const double cell[3],
                                                   uint64_t l, double x[3])
{
#pragma clang fp contract(off)
    uint64_t ijk[3];
    int axis = -1;
    uint64_t half = 0;
    if (l < g.nv) {
        ijk[2] = l / ((g.ny + 1) * (g.nx + 1));
        uint64_t const t = l % ((g.ny + 1) * (g.nx + 1));
        ijk[1] = t / (g.nx + 1);
        ijk[0] = t % (g.nx + 1);
    } else if (l < g.nv + 2 * g.nex) {
        uint64_t const m = l - g.nv, e = m / 2;
        half = m % 2;
        axis = 0;
        ijk[2] = e / ((g.ny + 1) * g.nx);
        uint64_t const t = e % ((g.ny + 1) * g.nx);
        ijk[1] = t / g.nx;
        ijk[0] = t % g.nx;
    } else if (l < g.nv + 2 * (g.nex + g.ney)) {
        uint64_t const m = l - g.nv - 2 * g.nex, e = m / 2;
        half = m % 2;
        axis = 1;
        ijk[0] = e / ((g.nz + 1) * g.ny);
        uint64_t const t = e % ((g.nz + 1) * g.ny);
        ijk[2] = t / g.ny;
        ijk[1] = t % g.ny;
    } else {
        uint64_t const m = l - g.nv - 2 * (g.nex + g.ney), e = m / 2;
        half = m % 2;
        axis = 2;
        ijk[1] = e / ((g.nx + 1) * g.nz);
        uint64_t const t = e % ((g.nx + 1) * g.nz);
        ijk[0] = t / g.nz;
        ijk[2] = t % g.nz;
    }
    for (int d = 0; d < 3; ++d) x[d] = dmin[d] + cell[d] * double(ijk[d]);
    if (axis >= 0) x[axis] += (1.0 + double(half)) / 3.0 * cell[axis];
}

// Value and gradient of grid SDF f at p (see the file comment).
__device__ inline double grid_eval(sdf_dev const& f, double px, double py, double pz, double grad[3])
{
#pragma clang fp contract(off)
    double const x[3] = {px, py, pz};
    grad[0] = grad[1] = grad[2] = 0.0;
    for (int d = 0; d < 3; ++d)
        if (!(f.gmin[d] <= x[d] && x[d] <= f.gmax[d])) return DBL_MAX;
    grid_dims const g(f.res);
    uint64_t mi[3];
    double xi[3], c0[3];
    for (int d = 0; d < 3; ++d) {
        uint32_t m = uint32_t((x[d] - f.gmin[d]) * f.inv_cell[d]);
        if (m >= f.res[d]) m = f.res[d] - 1;
        mi[d] = m;
        double const lo = f.gmin[d] + double(m) * f.cell[d];
        double const hi = lo + f.cell[d];
        double const den = hi - lo;
        c0[d] = 2.0 * (1.0 / den);
        double const c1 = (hi + lo) / den;
        xi[d] = c0[d] * x[d] - c1;
    }
    uint64_t const i = mi[0], j = mi[1], k = mi[2];
    double const xx = xi[0], yy = xi[1], zz = xi[2];
    double const x2 = xx * xx, y2 = yy * yy, z2 = zz * zz;
    double const fx[2] = {1.0 - xx, 1.0 + xx}, fy[2] = {1.0 - yy, 1.0 + yy}, fz[2] = {1.0 - zz, 1.0 + zz};
    double const sg[2] = {-1.0, 1.0};
    double const r = 9.0 * (x2 + y2 + z2) - 19.0;
    double const fac = 1.0 / 64.0 * r;
    const double* __restrict__ c = f.nodes;
    double phi = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
    for (int cn = 0; cn < 8; ++cn) {
        int const a = cn & 1, b = (cn >> 1) & 1, e = cn >> 2;
        double const v = c[(g.nx + 1) * (g.ny + 1) * (k + e) + (g.nx + 1) * (j + b) + (i + a)];
        double const N = fac * fx[a] * fy[b] * fz[e];
        double const d0 = 1.0 / 64.0 * (18.0 * xx * fx[a] + sg[a] * r) * fy[b] * fz[e];
        double const d1 = 1.0 / 64.0 * (18.0 * yy * fy[b] + sg[b] * r) * fx[a] * fz[e];
        double const d2 = 1.0 / 64.0 * (18.0 * zz * fz[e] + sg[e] * r) * fx[a] * fy[b];
        phi += v * N;
        g0 += v * d0;
        g1 += v * d1;
        g2 += v * d2;
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        double const u = ax == 0 ? xx : (ax == 1 ? yy : zz);
        double const u2 = u * u;
        double const gq = 9.0 / 64.0 * (1.0 - u2);
        double const lin[2] = {1.0 - 3.0 * u, 1.0 + 3.0 * u};
        double const dlin[2] = {-3.0, 3.0};
        // the other two axes (v < w) and, per edge slot e, their sides
        double const* fv = ax == 0 ? fy : fx;
        double const* fw = ax == 2 ? fy : fz;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int const sv = ax == 1 ? (e & 1) : (e >> 1), sw = ax == 1 ? (e >> 1) : (e & 1);
            uint64_t base;
            if (ax == 0)  // x edge (i, j + sv, k + sw)
                base = g.nv + 2 * (g.nx * (g.ny + 1) * (k + sw) + g.nx * (j + sv) + i);
            else if (ax == 1)  // y edge (i + sv, j, k + sw)
                base = g.nv + 2 * g.nex + 2 * (g.ny * (g.nz + 1) * (i + sv) + g.ny * (k + sw) + j);
            else  // z edge (i + sv, j + sw, k)
                base = g.nv + 2 * (g.nex + g.ney) + 2 * (g.nz * (g.nx + 1) * (j + sw) + g.nz * (i + sv) + k);
            double const vw = fv[sv] * fw[sw];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                double const v = c[base + h];
                double const N = gq * lin[h] * vw;
                double const du = 9.0 / 64.0 * (-2.0 * u * lin[h] + dlin[h] * (1.0 - u2)) * vw;
                double const dv = gq * lin[h] * sg[sv] * fw[sw];
                double const dw = gq * lin[h] * fv[sv] * sg[sw];
                double const d0 = ax == 0 ? du : dv;
                double const d1 = ax == 0 ? dv : (ax == 1 ? du : dw);
                double const d2 = ax == 2 ? du : dw;
                phi += v * N;
                g0 += v * d0;
                g1 += v * d1;
                g2 += v * d2;
            }
        }
    }
    grad[0] = g0 * c0[0];
    grad[1] = g1 * c0[1];
    grad[2] = g2 * c0[2];
    return phi;
}

// Mesh data of the node-distance kernel: non-degenerate triangles (3 vertex ids) and the
// pseudonormals of vertices, edges (3 per triangle: (0,1), (1,2), (2,0)) and faces, indexed by
// triangle; the triangles' corner positions (9 doubles each) in a spatial order with the original
// index of each, and the bounding boxes (lo xyz, hi xyz) of tiles of consecutive triangles in
// that order.
struct mesh_dev {
    const uint32_t* tri;
    const double* corners;
    const double* vnormal;
    const double* enormal;
    const double* fnormal;
    const uint32_t* orig;
    const double* tile_box;
    uint32_t m, ntiles;
};

void launch_grid_nodes(hipStream_t s, mesh_dev mesh, const uint32_t res[3], const double dmin[3],
                       const double cell[3], double* nodes);

// Host side: argument checks, cell size, the node values of a triangle mesh's signed distance
// written to device memory, environment_body_t's extended domain, and evaluation of one SDF at
// host points.
void check_grid(const double dmin[3], const double dmax[3], const uint32_t res[3]);
void grid_cell_size(const double dmin[3], const double dmax[3], const uint32_t res[3], double cell[3],
                    double inv_cell[3]);
void build_mesh_grid(hipStream_t s, const double* pos, uint32_t nv, const uint32_t* tris, uint32_t nt,
                     const double dmin[3], const double dmax[3], const uint32_t res[3], double* dev_nodes);
void environment_domain(const double* pos, uint32_t nv, const double lo_in[3], const double hi_in[3], double lo[3],
                        double hi[3]);
void evaluate_sdf(hipStream_t s, sdf_dev const& f, const double* pts, uint32_t n, double* sd, double* grad);

}  // namespace sbs
