// SDF evaluation on the device (shared by the contact kernels): analytic planes and spheres, and
// cubic Lagrange grids.
#pragma once
#include "grid_sdf.h"
#include "kernels.h"

namespace sbs {

// Signed distance and gradient in FP64 (the contact test and contact point).
// FP64 without contraction into FMAs, so that the contacts equal the oracle's bit for bit.
__device__ __forceinline__ void sdf_eval(sdf_dev const& f, pos4 p, double& sd, double grad[3])
{
#pragma clang fp contract(off)
    if (f.kind == 0) {
        grad[0] = f.a[0];
        grad[1] = f.a[1];
        grad[2] = f.a[2];
        sd = grad[0] * p.x + grad[1] * p.y + grad[2] * p.z + f.a[3];
    } else if (f.kind == 2) {
        sd = grid_eval(f, p.x, p.y, p.z, grad);
    } else {
        double const d[3] = {p.x - f.a[0], p.y - f.a[1], p.z - f.a[2]};
        double const len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        sd = len - f.a[3];
        if (len > 0.0) {
            grad[0] = d[0] / len;
            grad[1] = d[1] / len;
            grad[2] = d[2] / len;
        } else {
            grad[0] = 0.0;
            grad[1] = 1.0;
            grad[2] = 0.0;
        }
    }
}

// The contact test shared by every detection path: vertex p penetrates SDF f (sd < 0) and lies in
// its englobing volume. On a hit, q (contact point p + |sd| n) and n (unit normal) are set.
__device__ __forceinline__ bool sdf_contact(sdf_dev const& f, pos4 p, double q[3], double n[3])
{
#pragma clang fp contract(off)
    double sd, grad[3];
    sdf_eval(f, p, sd, grad);
    if (!(sd < 0.0)) return false;
    if (!(p.x >= f.vmin[0] && p.x <= f.vmax[0] && p.y >= f.vmin[1] && p.y <= f.vmax[1] && p.z >= f.vmin[2] &&
          p.z <= f.vmax[2]))
        return false;
    // Eigen's normalized(): a zero gradient stays zero
    double const gl2 = grad[0] * grad[0] + grad[1] * grad[1] + grad[2] * grad[2];
    double const gl = gl2 > 0.0 ? sqrt(gl2) : 1.0;
    n[0] = grad[0] / gl;
    n[1] = grad[1] / gl;
    n[2] = grad[2] / gl;
    double const a = fabs(sd);
    q[0] = p.x + a * n[0];
    q[1] = p.y + a * n[1];
    q[2] = p.z + a * n[2];
    return true;
}

// The point q of triangle t (9 values: its corners a, b, c) closest to p, by the Voronoi regions of
// the corners, edges and face (Ericson, "Real-Time Collision Detection", 5.1.5); *feature: 0-2
// corner a, b, c, 3 edge ab, 4 edge bc, 5 edge ca, 6 face; bary (optional): the weights of a, b, c
// that give q. FP64 without contraction into FMAs.
__device__ __forceinline__ void tri_closest_point(const double p[3], const double* t, double q[3], int* feature,
                                                  double* bary = nullptr)
{
#pragma clang fp contract(off)
    double ab[3], ac[3], ap[3], bp[3], cp[3];
    for (int d = 0; d < 3; ++d) {
        ab[d] = t[3 + d] - t[d];
        ac[d] = t[6 + d] - t[d];
        ap[d] = p[d] - t[d];
    }
    double const d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
    double const d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
    if (d1 <= 0.0 && d2 <= 0.0) {
        *feature = 0;
        for (int d = 0; d < 3; ++d) q[d] = t[d];
        if (bary) bary[0] = 1.0, bary[1] = 0.0, bary[2] = 0.0;
        return;
    }
    for (int d = 0; d < 3; ++d) bp[d] = p[d] - t[3 + d];
    double const d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
    double const d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
    if (d3 >= 0.0 && d4 <= d3) {
        *feature = 1;
        for (int d = 0; d < 3; ++d) q[d] = t[3 + d];
        if (bary) bary[0] = 0.0, bary[1] = 1.0, bary[2] = 0.0;
        return;
    }
    double const vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double const v = d1 / (d1 - d3);
        *feature = 3;
        for (int d = 0; d < 3; ++d) q[d] = t[d] + v * ab[d];
        if (bary) bary[0] = 1.0 - v, bary[1] = v, bary[2] = 0.0;
        return;
    }
    for (int d = 0; d < 3; ++d) cp[d] = p[d] - t[6 + d];
    double const d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
    double const d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
    if (d6 >= 0.0 && d5 <= d6) {
        *feature = 2;
        for (int d = 0; d < 3; ++d) q[d] = t[6 + d];
        if (bary) bary[0] = 0.0, bary[1] = 0.0, bary[2] = 1.0;
        return;
    }
    double const vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double const w = d2 / (d2 - d6);
        *feature = 5;
        for (int d = 0; d < 3; ++d) q[d] = t[d] + w * ac[d];
        if (bary) bary[0] = 1.0 - w, bary[1] = 0.0, bary[2] = w;
        return;
    }
    double const va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double const w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *feature = 4;
        for (int d = 0; d < 3; ++d) q[d] = t[3 + d] + w * (t[6 + d] - t[3 + d]);
        if (bary) bary[0] = 0.0, bary[1] = 1.0 - w, bary[2] = w;
        return;
    }
    double const den = 1.0 / (va + vb + vc);
    double const v = vb * den, w = vc * den;
    *feature = 6;
    for (int d = 0; d < 3; ++d) q[d] = t[d] + ab[d] * v + ac[d] * w;
    if (bary) bary[0] = 1.0 - v - w, bary[1] = v, bary[2] = w;
}

// Contact `slot` of surface vertex s: point q and SDF index k in c.q, normal n in c.n.
__device__ __forceinline__ void store_contact(contact_dev const& c, uint32_t s, uint32_t slot, uint32_t k,
                                              const double q[3], const double n[3])
{
    std::size_t const i = std::size_t(slot) * c.S + s;
    c.q[i] = make_double4(q[0], q[1], q[2], double(k));
    c.n[i] = make_float4(float(n[0]), float(n[1]), float(n[2]), 0.0f);
}

}  // namespace sbs
