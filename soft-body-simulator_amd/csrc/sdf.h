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

// Contact `slot` of surface vertex s: point q and SDF index k in c.q, normal n in c.n.
__device__ __forceinline__ void store_contact(contact_dev const& c, uint32_t s, uint32_t slot, uint32_t k,
                                              const double q[3], const double n[3])
{
    std::size_t const i = std::size_t(slot) * c.S + s;
    c.q[i] = make_double4(q[0], q[1], q[2], double(k));
    c.n[i] = make_float4(float(n[0]), float(n[1]), float(n[2]), 0.0f);
}

}  // namespace sbs
