// gfx950 kernels of the XPBD time step: prediction, per-color constraint projection (green strain,
// stable neo-Hookean, distance), clustered Gauss-Seidel, SDF contact detection and projection,
// velocity update, FP64 residual / energy reductions and Jones-Plassmann graph coloring.
//
// Particle state lives in HBM as SoA streams: xi (working positions, w = inverse mass) and xn
// (substep start = committed positions, w = inverse mass) as double4, velocities as float4.
// Positions are stored in FP64 so that a tet's edge vectors keep FP32 relative precision wherever
// the body is: every constraint forms its edge differences in FP64, does all of its arithmetic in
// FP32 on those differences, and adds its FP32 correction back in FP64. Constraint data of a group
// is stored in execution (color) order so that one color is one contiguous, coalesced range.
#include <algorithm>

#include "device_math.h"
#include "kernels.h"
#include "mlp_device.h"
#include "sdf.h"

namespace sbs {

namespace {

inline unsigned grid_for(uint32_t n, int block) { return (n + uint32_t(block) - 1) / uint32_t(block); }

// a - b of two positions, rounded to FP32
__device__ __forceinline__ v3 sub(pos4 a, pos4 b)
{
    return {float(a.x - b.x), float(a.y - b.y), float(a.z - b.z)};
}
// p += d (an FP32 correction) in FP64
__device__ __forceinline__ void add_to(pos4& p, v3 d)
{
    p.x += double(d.x);
    p.y += double(d.y);
    p.z += double(d.z);
}
__device__ __forceinline__ float wof(pos4 p) { return float(p.w); }

// ---------------------------------------------------------------- prediction / commit

// timestep.cpp:35-43: v += a dt, xi = x + v dt. With gravity_is_force the acceleration is
// g * w (the reference's f = g, a = f w); otherwise g for every particle with w > 0.
__device__ __forceinline__ void predict_one(pos4& xi, pos4 const& x, float4& v, float gx, float gy, float gz,
                                            int force, float dt)
{
    float const w = wof(x);
    float ax, ay, az;
    if (force) {
        ax = gx * w;
        ay = gy * w;
        az = gz * w;
    } else if (w > 0.0f) {
        ax = gx;
        ay = gy;
        az = gz;
    } else {
        ax = ay = az = 0.0f;
    }
    v.x = v.x + ax * dt;
    v.y = v.y + ay * dt;
    v.z = v.z + az * dt;
    double const d = dt;
    xi = make_double4(x.x + double(v.x) * d, x.y + double(v.y) * d, x.z + double(v.z) * d, x.w);
}

__global__ void k_predict(pos4* __restrict__ xi, const pos4* __restrict__ xn, float4* __restrict__ v, uint32_t n,
                          float gx, float gy, float gz, int force, float dt)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos4 const x = xn[i];
    float4 vel = v[i];
    pos4 p;
    predict_one(p, x, vel, gx, gy, gz, force, dt);
    v[i] = vel;
    xi[i] = p;
}

// timestep.cpp:48-57: x = xi, v = (x - xn)/dt, xn = x.
__device__ __forceinline__ float4 velocity(pos4 p, pos4 q, float dt)
{
    double const d = dt;
    return make_float4(float((p.x - q.x) / d), float((p.y - q.y) / d), float((p.z - q.z) / d), 0.0f);
}

__global__ void k_commit(const pos4* __restrict__ xi, pos4* __restrict__ xn, float4* __restrict__ v, uint32_t n,
                         float dt)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos4 const p = xi[i], q = xn[i];
    v[i] = velocity(p, q, dt);
    xn[i] = p;
}

// Commit of one substep fused with the prediction of the next (same per-particle order). With prow,
// thread r handles particle prow[r] (rows [0, n); padding rows hold 0xffffffff), else particle r.
__global__ void k_commit_predict(pos4* __restrict__ xi, pos4* __restrict__ xn, float4* __restrict__ v, uint32_t n,
                                 float gx, float gy, float gz, int force, float dt, const uint32_t* __restrict__ prow)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (prow) {
        i = prow[i];
        if (i == 0xffffffffu) return;
    }
    pos4 const p = xi[i], q = xn[i];
    float4 vel = velocity(p, q, dt);
    pos4 np;
    predict_one(np, p, vel, gx, gy, gz, force, dt);
    xn[i] = p;
    v[i] = vel;
    xi[i] = np;
}

// ---------------------------------------------------------------- green strain constraint

struct tet_rows {
    v3 r0, r1, r2;
};

__device__ __forceinline__ tet_rows load_rows(float4 a, float4 b, float4 c)
{
    return {mk(a.x, a.y, a.z), mk(a.w, b.x, b.y), mk(b.z, b.w, c.x)};
}

// green_constraint.cpp:61-133 on the edge vectors e_k = p_k - p_4. Computes C = |V0| psi and
// f1..f4 = -grad C, with the reference's inversion flip of the smallest singular value (:92-96),
// the 0.577 clamp (:99-102), the StVK first Piola stress of the clamped F^ mapped back with U, V
// (:104-112), and psi from E = U E^ V^T: |E|_F^2 = sum e_i^2, tr(E) = sum e_i (u_i . v_i) (:108-110).
__device__ __forceinline__ float green_core(v3 e0, v3 e1, v3 e2, tet_rows const& R, float V0, float mu, float la,
                                            v3 f[4])
{
    float const det = dot(e0, cross(e1, e2));
    bool const inverted = (det >= 0.0f) != (V0 >= 0.0f);
    m3 const F = deformation_gradient(e0, e1, e2, R.r0, R.r1, R.r2);
    m3 U, V;
    float s[3];
    svd3(F, U, s, V);
    // reference convention: non-negative singular values, the smallest negated iff inverted
    float s2 = inverted ? -fabsf(s[2]) : fabsf(s[2]);
    if ((s2 < 0.0f) != (s[2] < 0.0f)) U.c2 = -U.c2;
    float const f0 = fmaxf(s[0], 0.577f), f1 = fmaxf(s[1], 0.577f), f2 = fmaxf(s2, 0.577f);
    float const h0 = 0.5f * (f0 * f0 - 1.0f), h1 = 0.5f * (f1 * f1 - 1.0f), h2 = 0.5f * (f2 * f2 - 1.0f);
    float const tr = h0 + h1 + h2;
    float const ph0 = f0 * (2.0f * mu * h0 + la * tr);
    float const ph1 = f1 * (2.0f * mu * h1 + la * tr);
    float const ph2 = f2 * (2.0f * mu * h2 + la * tr);
    float const trE = h0 * dot(U.c0, V.c0) + h1 * dot(U.c1, V.c1) + h2 * dot(U.c2, V.c2);
    float const psi = mu * (h0 * h0 + h1 * h1 + h2 * h2) + 0.5f * la * trE * trE;
    float const aV0 = fabsf(V0);
    // f_k = -|V0| P r_k, P = sum_i ph_i u_i v_i^T
    float const a0 = -aV0 * ph0, a1 = -aV0 * ph1, a2 = -aV0 * ph2;
    auto force = [&](v3 r) { return (a0 * dot(V.c0, r)) * U.c0 + (a1 * dot(V.c1, r)) * U.c1 + (a2 * dot(V.c2, r)) * U.c2; };
    f[0] = force(R.r0);
    f[1] = force(R.r1);
    f[2] = force(R.r2);
    f[3] = -(f[0] + f[1] + f[2]);
    return aV0 * psi;
}

// One green projection of the tet (P0..P3) with material (a, b, c) and multiplier lm; xn_of(k) is
// the substep start position of vertex k in P's frame, read only for the damping term. Returns
// false on the epsilon early-out.
template <class Pos, class XnOf>
__device__ __forceinline__ bool green_project(Pos P[4], XnOf&& xn_of, float4 a, float4 b, float4 c, float alpha_tilde,
                                              float gamma, float& lm)
{
    v3 f[4];
    float const C = green_core(sub(P[0], P[3]), sub(P[1], P[3]), sub(P[2], P[3]), load_rows(a, b, c), c.y, c.z, c.w, f);
    float const wsum = wof(P[0]) * dot(f[0], f[0]) + wof(P[1]) * dot(f[1], f[1]) + wof(P[2]) * dot(f[2], f[2]) +
                       wof(P[3]) * dot(f[3], f[3]);
    if (wsum < 1e-20f) return false;  // green_constraint.cpp:130-131
    float gdd = 0.0f;
    if (gamma != 0.0f) {  // damping term (:140-144)
        gdd = dot(f[0], sub(P[0], xn_of(0))) + dot(f[1], sub(P[1], xn_of(1))) + dot(f[2], sub(P[2], xn_of(2))) +
              dot(f[3], sub(P[3], xn_of(3)));
    }
    float const dl = hw_div(-(C + alpha_tilde * lm) + gamma * gdd, (1.0f + gamma) * wsum + alpha_tilde);
    lm += dl;
    // xi += w (-f) dl (:154-157)
#pragma unroll
    for (int k = 0; k < 4; ++k) add_to(P[k], -(wof(P[k]) * dl) * f[k]);
    return true;
}

__global__ void __launch_bounds__(256) k_green(tet_group_dev g, pos4* xi, const pos4* __restrict__ xn, uint32_t begin,
                                               uint32_t count, float alpha_tilde, float gamma, int first)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t const t = begin + i;
    int4 const id = g.idx[t];
    float4 const a = g.m0[t], b = g.m1[t], c = g.m2[t];
    float lm = first ? 0.0f : g.lambda[t];
    pos4 P[4] = {xi[id.x], xi[id.y], xi[id.z], xi[id.w]};
    auto xn_of = [&](int k) { return xn[k == 0 ? id.x : (k == 1 ? id.y : (k == 2 ? id.z : id.w))]; };
    if (!green_project(P, xn_of, a, b, c, alpha_tilde, gamma, lm)) {
        if (first) g.lambda[t] = 0.0f;
        return;
    }
    g.lambda[t] = lm;
    xi[id.x] = P[0];
    xi[id.y] = P[1];
    xi[id.z] = P[2];
    xi[id.w] = P[3];
}

// ---------------------------------------------------------------- stable neo-Hookean constraint

// A tet in FP32 relative form for a sequence of constraint updates: edges e_k = p_k - p_3, the
// displacements dx_k = p_k - xn_k since the substep start (only with damping), the inverse masses
// and the corrections d_k accumulated so far. The FP64 differences are taken once; each update
// then moves the FP32 quantities, and the positions receive the sum of the corrections once.
struct tet_frame {
    v3 e[3], dx[4], d[4];
    float w[4];
};

template <class Pos, class XnOf>
__device__ __forceinline__ tet_frame make_frame(Pos const P[4], XnOf&& xn_of, bool damped)
{
    tet_frame f;
#pragma unroll
    for (int k = 0; k < 3; ++k) f.e[k] = sub(P[k], P[3]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f.w[k] = wof(P[k]);
        f.d[k] = mk(0.0f, 0.0f, 0.0f);
        f.dx[k] = damped ? sub(P[k], xn_of(k)) : mk(0.0f, 0.0f, 0.0f);
    }
    return f;
}

// Rest data of a neo-Hookean tet: with R = DmInv (rows r_k), Ds = [e0 e1 e2] and F = Ds R,
//   ||F||_F^2 = sum_mn (e_m . e_n) Q_mn and F r_k = sum_m Q_mk e_m, Q = R R^T (symmetric);
//   cof(F) r_k = cof(Ds) cof(R) r_k = det(R) (e_{k+1} x e_{k+2}), det F = det(R) e0 . (e1 x e2)
// (cof(R) = det(R) R^-T and R^-T r_k = Dm^T r_k is the unit vector k). So the deviatoric gradient
// dC/dx_k = F r_k / ||F|| and the hydrostatic one cof(F) r_k need no F and no cofactor matrix.
// Material layout (40 B per tet): m0 = (Q00, Q11, Q22, Q01), m1 = (Q02, Q12, det R, 1/(mu|V0|)),
// and m2 read as float2 per tet = (1/(lambda|V0|), 1 + mu/lambda).
struct nh_rest {
    float q00, q11, q22, q01, q02, q12, detr;
};

__device__ __forceinline__ nh_rest load_nh(float4 a, float4 b) { return {a.x, a.y, a.z, a.w, b.x, b.y, b.z}; }

__device__ __forceinline__ float2 nh_m2(tet_group_dev const& g, uint32_t t)
{
    return reinterpret_cast<const float2*>(g.m2)[t];
}

// The neo-Hookean updates below spell out every fused multiply-add and contract nothing else, so
// their rounding is fixed by the source and not by the kernel they are inlined into: the hex
// kernel's launch variants (material prefetched or streamed) give bit-identical positions.
#pragma clang fp contract(on)

__device__ __forceinline__ float fdot(v3 a, v3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }

__device__ __forceinline__ v3 fcross(v3 a, v3 b)
{
    return {fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x))};
}

// a e0 + b e1 + c e2
__device__ __forceinline__ v3 fcomb(float a, v3 e0, float b, v3 e1, float c, v3 e2)
{
    return {fmaf(c, e2.x, fmaf(b, e1.x, a * e0.x)), fmaf(c, e2.y, fmaf(b, e1.y, a * e0.y)),
            fmaf(c, e2.z, fmaf(b, e1.z, a * e0.z))};
}

// x + s a
__device__ __forceinline__ v3 faxpy(float s, v3 a, v3 x) { return {fmaf(s, a.x, x.x), fmaf(s, a.y, x.y), fmaf(s, a.z, x.z)}; }

// -(a + b + c)
__device__ __forceinline__ v3 fnegsum(v3 a, v3 b, v3 c)
{
    return {-((a.x + b.x) + c.x), -((a.y + b.y) + c.y), -((a.z + b.z) + c.z)};
}

// Deviatoric constraint C = ||F||_F at edges e; gradients s h_k with s = 1/C (0 when C = 0).
__device__ __forceinline__ float nh_deviatoric(v3 const e[3], nh_rest const& q, v3 h[4], float& s)
{
    float const g00 = fdot(e[0], e[0]), g11 = fdot(e[1], e[1]), g22 = fdot(e[2], e[2]);
    float const g01 = fdot(e[0], e[1]), g02 = fdot(e[0], e[2]), g12 = fdot(e[1], e[2]);
    float const off = fmaf(q.q01, g01, fmaf(q.q02, g02, q.q12 * g12));
    float const c2 = fmaf(q.q00, g00, fmaf(q.q11, g11, fmaf(q.q22, g22, 2.0f * off)));
    float const C = hw_sqrt(fmaxf(c2, 0.0f));
    s = C > 0.0f ? hw_rcp(C) : 0.0f;
    h[0] = fcomb(q.q00, e[0], q.q01, e[1], q.q02, e[2]);
    h[1] = fcomb(q.q01, e[0], q.q11, e[1], q.q12, e[2]);
    h[2] = fcomb(q.q02, e[0], q.q12, e[1], q.q22, e[2]);
    h[3] = fnegsum(h[0], h[1], h[2]);
    return C;
}

// Hydrostatic constraint C = det F - gamma0 at edges e; gradients s h_k with s = det R.
__device__ __forceinline__ float nh_hydrostatic(v3 const e[3], nh_rest const& q, float gamma0, v3 h[4], float& s)
{
    h[0] = fcross(e[1], e[2]);
    h[1] = fcross(e[2], e[0]);
    h[2] = fcross(e[0], e[1]);
    h[3] = fnegsum(h[0], h[1], h[2]);
    s = q.detr;
    return fmaf(q.detr, fdot(e[0], h[0]), -gamma0);
}

// Sum over the tet's vertices of w_k |s h_k|^2.
__device__ __forceinline__ float grad_norm2(float const w[4], v3 const h[4], float s)
{
    float const n = fmaf(w[3], fdot(h[3], h[3]), fmaf(w[2], fdot(h[2], h[2]), fmaf(w[1], fdot(h[1], h[1]), w[0] * fdot(h[0], h[0]))));
    return (s * s) * n;
}

// One XPBD update of a tet constraint with gradients g_k = s h_k: dl from the multiplier and the
// damping term, then d_k += w_k s dl h_k (and dx_k), e_k += (corr_k - corr_3).
__device__ __forceinline__ void tet_update(tet_frame& f, v3 const h[4], float s, float C, float alpha_tilde,
                                           float gamma, float& lm)
{
    float const wsum = grad_norm2(f.w, h, s);
    if (wsum < 1e-20f) return;
    float gdd = 0.0f;
    if (gamma != 0.0f)
        gdd = s * (((fdot(h[0], f.dx[0]) + fdot(h[1], f.dx[1])) + fdot(h[2], f.dx[2])) + fdot(h[3], f.dx[3]));
    float const dl = hw_div(fmaf(gamma, gdd, -fmaf(alpha_tilde, lm, C)), fmaf(1.0f + gamma, wsum, alpha_tilde));
    lm += dl;
    float const sd = s * dl;
    float ws[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        ws[k] = f.w[k] * sd;
        f.d[k] = faxpy(ws[k], h[k], f.d[k]);
        f.dx[k] = faxpy(ws[k], h[k], f.dx[k]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) f.e[k] = faxpy(ws[k], h[k], faxpy(-ws[3], h[3], f.e[k]));
}

// Deviatoric then hydrostatic projection of one tet (P0..P3), multipliers lm = (deviatoric,
// hydrostatic), material (a, b, c) as laid out at nh_rest. xn_of(k) is the substep start position
// of vertex k, read only with damping (beta_dt != 0).
template <class Pos, class XnOf>
__device__ __forceinline__ void neohookean_project(Pos P[4], XnOf&& xn_of, float4 a, float4 b, float2 c, float inv_dt2,
                                                   float beta_dt, float2& lm)
{
    nh_rest const q = load_nh(a, b);
    tet_frame f = make_frame(P, xn_of, beta_dt != 0.0f);
    {  // deviatoric: C = ||F||_F, compliance 1/(mu |V0|)
        v3 h[4];
        float s;
        float const C = nh_deviatoric(f.e, q, h, s);
        float const at = b.w * inv_dt2;
        tet_update(f, h, s, C, at, at * beta_dt, lm.x);
    }
    {  // hydrostatic: C = det F - (1 + mu/lambda), compliance 1/(lambda |V0|)
        v3 h[4];
        float s;
        float const C = nh_hydrostatic(f.e, q, c.y, h, s);
        float const at = c.x * inv_dt2;
        tet_update(f, h, s, C, at, at * beta_dt, lm.y);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) add_to(P[k], f.d[k]);
}

// Warm-start displacement of one tet at positions P (the predicted positions): both gradients at
// P, d_k = w_k (grad_k C_dev lm.x + grad_k C_hyd lm.y). With explicit the multipliers are first
// set to -C(P) / alpha~, the pair that zeroes the XPBD residual at P; then they are multiplied by
// scale (x: deviatoric, y: hydrostatic; the caller stores them back). A constraint whose
// sum w|grad C|^2 is below the projection's epsilon contributes nothing, as it would not project.
template <class Pos>
__device__ __forceinline__ void neohookean_displacement(Pos const P[4], float4 a, float4 b, float2 c, float inv_dt2,
                                                        bool expl, float2 scale, float2& lm, v3 d[4])
{
    nh_rest const q = load_nh(a, b);
    tet_frame const f = make_frame(P, [](int) { return pos4{}; }, false);
    v3 hd[4], hh[4];
    float sd, sh;
    float const Cd = nh_deviatoric(f.e, q, hd, sd);
    float const Ch = nh_hydrostatic(f.e, q, c.y, hh, sh);
    if (expl) lm = make_float2(-Cd / (b.w * inv_dt2), -Ch / (c.x * inv_dt2));
    lm = make_float2(scale.x * lm.x, scale.y * lm.y);
    float const wd = grad_norm2(f.w, hd, sd), wh = grad_norm2(f.w, hh, sh);
    float const ld = wd < 1e-20f ? 0.0f : lm.x * sd, lh = wh < 1e-20f ? 0.0f : lm.y * sh;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float const a = f.w[k] * ld, b = f.w[k] * lh;
        d[k] = faxpy(b, hh[k], mk(a * hd[k].x, a * hd[k].y, a * hd[k].z));
    }
}

// The contacts of surface vertex s (all its slots in SDF order) projected onto P, the vertex's
// position (collision_constraint.cpp:21-41): C = (p - q).n; when C < 0, dl = -(C + alpha~ lm) /
// (w + alpha~), lm += dl, p += w n dl. first: the first iteration (multipliers start at zero, and
// are stored as zero for slots without penetration). For a contact with a surface triangle, q is the
// triangle's point of the stored barycentric weights b at the copied surface positions, plus h n, and
// the denominator adds that point's inverse mass sum_k b_k^2 w_k: the vertex moves by its share of
// the two-sided projection (the triangle's corners take theirs through their own contacts).
__device__ __forceinline__ void contact_vertex_at(contact_dev const& c, pos4& P, uint32_t s, float alpha_tilde,
                                                  bool first)
{
    uint32_t const cnt = c.count[s];
    float const w = wof(P);
    for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t const slot = k * c.S + s;
        pos4 q = c.q[slot];
        v3 const n = xyz(c.n[slot]);
        float wt = 0.0f;
        if (c.snap && q.w >= double(c.nsdf)) {  // a surface triangle: its closest point as it is now
            uint4 const tr = c.btris[uint32_t(q.w) - c.nsdf];
            pos4 const a = c.snap[tr.x], b = c.snap[tr.y], d = c.snap[tr.z];
            wt = float(q.x * q.x * a.w + q.y * q.y * b.w + q.z * q.z * d.w);
            double const hx = c.body_h * double(n.x), hy = c.body_h * double(n.y), hz = c.body_h * double(n.z);
            q = make_double4(q.x * a.x + q.y * b.x + q.z * d.x + hx, q.x * a.y + q.y * b.y + q.z * d.y + hy,
                             q.x * a.z + q.y * b.z + q.z * d.z + hz, q.w);
        }
        float const lm = first ? 0.0f : c.lambda[slot];
        float const C = fdot(sub(P, q), n);
        if (C >= 0.0f) {
            if (first) c.lambda[slot] = 0.0f;
            continue;
        }
        float const dl = -fmaf(alpha_tilde, lm, C) / (w + wt + alpha_tilde);
        c.lambda[slot] = lm + dl;
        float const sd = w * dl;
        P.x += double(sd * n.x);
        P.y += double(sd * n.y);
        P.z += double(sd * n.z);
    }
}

#pragma clang fp contract(fast)

// x += d in FP64 (read, add, write) for the x, y, z of one particle.
__device__ __forceinline__ void add_displacement(pos4* x, v3 d)
{
    double* q = reinterpret_cast<double*>(x);
    q[0] += double(d.x);
    q[1] += double(d.y);
    q[2] += double(d.z);
}

__global__ void __launch_bounds__(256) k_neohookean(tet_group_dev g, pos4* xi, const pos4* __restrict__ xn,
                                                    uint32_t begin, uint32_t count, float inv_dt2, float beta_dt,
                                                    int mode, float hyd_scale)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t const t = begin + i;
    int4 const id = g.idx[t];
    float4 const a = g.m0[t], b = g.m1[t];
    float2 const c = nh_m2(g, t);
    float2 lm = mode == kSolveFirst ? make_float2(0.0f, 0.0f) : reinterpret_cast<const float2*>(g.lambda)[t];
    if (mode >= kApplyWarmStart) {  // gradients at the predicted positions (xn holds a copy of them)
        pos4 const P[4] = {xn[id.x], xn[id.y], xn[id.z], xn[id.w]};
        v3 d[4];
        // in the warm-start modes beta_dt carries the deviatoric multiplier's scale (no damping term
        // applies), hyd_scale the hydrostatic one's
        neohookean_displacement(P, a, b, c, inv_dt2, mode == kApplyExplicit, make_float2(beta_dt, hyd_scale), lm, d);
        reinterpret_cast<float2*>(g.lambda)[t] = lm;
        add_displacement(&xi[id.x], d[0]);
        add_displacement(&xi[id.y], d[1]);
        add_displacement(&xi[id.z], d[2]);
        add_displacement(&xi[id.w], d[3]);
        return;
    }
    pos4 P[4] = {xi[id.x], xi[id.y], xi[id.z], xi[id.w]};
    auto xn_of = [&](int k) { return xn[k == 0 ? id.x : (k == 1 ? id.y : (k == 2 ? id.z : id.w))]; };
    neohookean_project(P, xn_of, a, b, c, inv_dt2, beta_dt, lm);
    reinterpret_cast<float2*>(g.lambda)[t] = lm;
    xi[id.x] = P[0];
    xi[id.y] = P[1];
    xi[id.z] = P[2];
    xi[id.w] = P[3];
}

// ---------------------------------------------------------------- multiplier warm start

__global__ void k_copy_pos(const pos4* __restrict__ a, pos4* __restrict__ b, uint32_t n)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// ---------------------------------------------------------------- clustered Gauss-Seidel

// One thread per cluster: the cluster's particles are staged in LDS (slot-major, so that the
// threads of a wave hit consecutive words), its constraints run in order against the staged
// copies, and the particles are written back once. Clusters of one color share no particle, so
// this equals the sequential sweep over the color's clusters.
template <int BLOCK>
struct cluster_stage {
    pos4 p[kClusterSlots][BLOCK];
};

__device__ __forceinline__ int slot_id(int4 va, int4 vb, uint32_t s)
{
    int const lo = s == 0 ? va.x : (s == 1 ? va.y : (s == 2 ? va.z : va.w));
    int const hi = s == 4 ? vb.x : (s == 5 ? vb.y : (s == 6 ? vb.z : vb.w));
    return s < 4 ? lo : hi;
}

template <int BLOCK, class Project>
__device__ __forceinline__ void run_cluster(cluster_dev const& cl, pos4* xi, const pos4* __restrict__ xn,
                                            uint32_t begin, uint32_t count, Project&& project)
{
    __shared__ cluster_stage<BLOCK> st;
    uint32_t const i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= count) return;
    uint32_t const c = begin + i;
    int4 const va = cl.v0[c], vb = cl.v1[c];
    int const ids[kClusterSlots] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
    pos4 (*P)[BLOCK] = st.p;
#pragma unroll
    for (int k = 0; k < kClusterSlots; ++k)
        if (ids[k] >= 0) P[k][threadIdx.x] = xi[ids[k]];
    uint32_t const t0 = cl.off[c], t1 = cl.off[c + 1];
    for (uint32_t t = t0; t < t1; ++t) {
        uint32_t const loc = cl.loc[t];
        uint32_t const s0 = loc & 0xffu, s1 = (loc >> 8) & 0xffu, s2 = (loc >> 16) & 0xffu, s3 = loc >> 24;
        pos4 Q[4] = {P[s0][threadIdx.x], P[s1][threadIdx.x], P[s2][threadIdx.x], P[s3][threadIdx.x]};
        int const id[4] = {slot_id(va, vb, s0), slot_id(va, vb, s1), slot_id(va, vb, s2), slot_id(va, vb, s3)};
        auto xn_of = [&](int k) { return xn[id[k]]; };
        project(t, Q, xn_of);
        P[s0][threadIdx.x] = Q[0];
        P[s1][threadIdx.x] = Q[1];
        P[s2][threadIdx.x] = Q[2];
        P[s3][threadIdx.x] = Q[3];
    }
#pragma unroll
    for (int k = 0; k < kClusterSlots; ++k)
        if (ids[k] >= 0) xi[ids[k]] = P[k][threadIdx.x];
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_green_clusters(tet_group_dev g, cluster_dev cl, pos4* xi,
                                                          const pos4* __restrict__ xn, uint32_t begin, uint32_t count,
                                                          float alpha_tilde, float gamma, int first)
{
    run_cluster<BLOCK>(cl, xi, xn, begin, count, [&](uint32_t t, pos4 P[4], auto&& xn_of) {
        float lm = first ? 0.0f : g.lambda[t];
        if (green_project(P, xn_of, g.m0[t], g.m1[t], g.m2[t], alpha_tilde, gamma, lm) || first) g.lambda[t] = lm;
    });
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_neohookean_clusters(tet_group_dev g, cluster_dev cl, pos4* xi,
                                                               const pos4* __restrict__ xn, uint32_t begin,
                                                               uint32_t count, float inv_dt2, float beta_dt, int mode)
{
    run_cluster<BLOCK>(cl, xi, xn, begin, count, [&](uint32_t t, pos4 P[4], auto&& xn_of) {
        float2 lm = mode == kSolveFirst ? make_float2(0.0f, 0.0f) : reinterpret_cast<const float2*>(g.lambda)[t];
        neohookean_project(P, xn_of, g.m0[t], g.m1[t], nh_m2(g, t), inv_dt2, beta_dt, lm);
        reinterpret_cast<float2*>(g.lambda)[t] = lm;
    });
}

// Warm-start displacement of the clusters of one color: every tet's gradients at the predicted
// positions xt, its displacement added to xi at once (clusters of a color share no particle, and
// one thread runs a cluster's tets in order, so the additions do not race).
__global__ void __launch_bounds__(128) k_neohookean_clusters_apply(tet_group_dev g, cluster_dev cl, pos4* xi,
                                                                   const pos4* __restrict__ xt, uint32_t begin,
                                                                   uint32_t count, float inv_dt2, int expl,
                                                                   float2 scale)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t const c = begin + i;
    int4 const va = cl.v0[c], vb = cl.v1[c];
    for (uint32_t t = cl.off[c]; t < cl.off[c + 1]; ++t) {
        uint32_t const loc = cl.loc[t];
        int const id[4] = {slot_id(va, vb, loc & 0xffu), slot_id(va, vb, (loc >> 8) & 0xffu),
                           slot_id(va, vb, (loc >> 16) & 0xffu), slot_id(va, vb, loc >> 24)};
        pos4 const P[4] = {xt[id[0]], xt[id[1]], xt[id[2]], xt[id[3]]};
        float2 lm = reinterpret_cast<const float2*>(g.lambda)[t];
        v3 d[4];
        neohookean_displacement(P, g.m0[t], g.m1[t], nh_m2(g, t), inv_dt2, expl != 0, scale, lm, d);
        reinterpret_cast<float2*>(g.lambda)[t] = lm;
#pragma unroll
        for (int k = 0; k < 4; ++k) add_displacement(&xi[id[k]], d[k]);
    }
}

// ---------------------------------------------------------------- hex-cell clusters

// Slot pattern shared by every cluster made of the 6 tets of one hex cell of the box split (slots
// numbered by first appearance). With the pattern known at compile time the 8 particles live in
// registers, and every constraint's material and multipliers are loaded before the first one
// runs. Inside a color the constraints are slot-major: tet k of the color's i-th cluster sits at
// tbase + k * count + i, so the loads of a wave coalesce.
constexpr int kHexTets = 6;

constexpr int kHexSlots[kHexTets][4] = {{0, 1, 2, 3}, {0, 4, 1, 3}, {0, 2, 5, 3}, {0, 5, 6, 3}, {0, 7, 4, 3}, {0, 6, 7, 3}};

// PREFETCH: every tet's material and multipliers are loaded before the first projection (one
// memory round trip, ~240 VGPRs, 2 waves per SIMD: for launches too small to fill the chip, where
// latency rules); otherwise each tet's are loaded just before it is projected (fewer registers,
// more waves in flight: for launches that fill the chip, where bandwidth rules).
// FIRST: the first iteration of a substep (multipliers start at zero). A compile-time switch: a
// multiplier load under a run-time condition sat in a branch, after which the compiler drained the
// memory queue, serializing the prefetch of the six tets' data into several round trips.
template <int BLOCK, bool PREFETCH, int MIN_WAVES, bool DAMPED, bool FIRST>
__global__ void __launch_bounds__(BLOCK, MIN_WAVES) k_neohookean_hex(tet_group_dev g, cluster_dev cl, pos4* xi,
                                                                     const pos4* __restrict__ xn, uint32_t begin,
                                                                     uint32_t count, uint32_t tbase, xcd_split sp,
                                                                     float inv_dt2, float beta_dt)
{
    uint32_t const x = blockIdx.x % kXcds, local = (blockIdx.x / kXcds) * BLOCK + threadIdx.x;
    if (local >= sp.off[x + 1] - sp.off[x]) return;
    uint32_t const i = sp.off[x] + local;
    uint32_t const c = begin + i;
    int4 const va = cl.v0[c], vb = cl.v1[c];
    int const ids[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
    constexpr int NP = PREFETCH ? kHexTets : 1;
    float4 A[NP], B[NP];
    float2 Cm[NP];
    float2 L[NP];
    auto load = [&](int k, int slot) {
        uint32_t const t = tbase + uint32_t(k) * count + i;
        A[slot] = g.m0[t];
        B[slot] = g.m1[t];
        Cm[slot] = nh_m2(g, t);
        L[slot] = FIRST ? make_float2(0.0f, 0.0f) : reinterpret_cast<const float2*>(g.lambda)[t];
    };
    auto store = [&](int k, int slot) { reinterpret_cast<float2*>(g.lambda)[tbase + uint32_t(k) * count + i] = L[slot]; };
    if (PREFETCH) {
#pragma unroll
        for (int k = 0; k < kHexTets; ++k) load(k, k);
    }
    pos4 P[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) P[k] = xi[ids[k]];
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) {
        int const m = PREFETCH ? k : 0;
        if (!PREFETCH) load(k, 0);
        int const s0 = kHexSlots[k][0], s1 = kHexSlots[k][1], s2 = kHexSlots[k][2], s3 = kHexSlots[k][3];
        pos4 Q[4] = {P[s0], P[s1], P[s2], P[s3]};
        int const id[4] = {ids[s0], ids[s1], ids[s2], ids[s3]};
        auto xn_of = [&](int q) { return xn[id[q]]; };
        neohookean_project(Q, xn_of, A[m], B[m], Cm[m], inv_dt2, DAMPED ? beta_dt : 0.0f, L[m]);
        P[s0] = Q[0];
        P[s1] = Q[1];
        P[s2] = Q[2];
        P[s3] = Q[3];
        if (!PREFETCH) store(k, 0);
    }
    if (PREFETCH) {
#pragma unroll
        for (int k = 0; k < kHexTets; ++k) store(k, k);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) xi[ids[k]] = P[k];
}

// Warm-start displacement of the hex clusters of one color (gradients at the predicted positions
// xt, the 6 tets' displacements summed per particle, then added to xi).
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_neohookean_hex_apply(tet_group_dev g, cluster_dev cl, pos4* xi,
                                                                const pos4* __restrict__ xt, uint32_t begin,
                                                                uint32_t count, uint32_t tbase, float inv_dt2, int expl,
                                                                float2 scale)
{
    uint32_t const i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= count) return;
    uint32_t const c = begin + i;
    int4 const va = cl.v0[c], vb = cl.v1[c];
    int const ids[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
    pos4 P[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) P[k] = xt[ids[k]];
    v3 D[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) D[k] = mk(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) {
        uint32_t const t = tbase + uint32_t(k) * count + i;
        int const s0 = kHexSlots[k][0], s1 = kHexSlots[k][1], s2 = kHexSlots[k][2], s3 = kHexSlots[k][3];
        pos4 const Q[4] = {P[s0], P[s1], P[s2], P[s3]};
        float2 lm = reinterpret_cast<const float2*>(g.lambda)[t];
        v3 d[4];
        neohookean_displacement(Q, g.m0[t], g.m1[t], nh_m2(g, t), inv_dt2, expl != 0, scale, lm, d);
        reinterpret_cast<float2*>(g.lambda)[t] = lm;
        D[s0] = D[s0] + d[0];
        D[s1] = D[s1] + d[1];
        D[s2] = D[s2] + d[2];
        D[s3] = D[s3] + d[3];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) add_displacement(&xi[ids[k]], D[k]);
}

// Constraint values of a neo-Hookean tet at edges e (the quantities nh_deviatoric and
// nh_hydrostatic return, without their gradients).
__device__ __forceinline__ float2 nh_values(v3 const e[3], nh_rest const& q, float gamma0)
{
    float const g00 = fdot(e[0], e[0]), g11 = fdot(e[1], e[1]), g22 = fdot(e[2], e[2]);
    float const g01 = fdot(e[0], e[1]), g02 = fdot(e[0], e[2]), g12 = fdot(e[1], e[2]);
    float const off = fmaf(q.q01, g01, fmaf(q.q02, g02, q.q12 * g12));
    float const c2 = fmaf(q.q00, g00, fmaf(q.q11, g11, fmaf(q.q22, g22, 2.0f * off)));
    return make_float2(hw_sqrt(fmaxf(c2, 0.0f)), fmaf(q.detr, fdot(e[0], fcross(e[1], e[2])), -gamma0));
}

// Warm-start displacements of every hex cluster of a group in one launch (all colors): nothing
// moves during the launch, so every gradient is taken at the predicted positions xi (a copy of
// them, or the working positions themselves before any displacement is added). Workgroup b (64
// rows) runs the clusters cinfo[2b].w + t, t < cinfo[2b + 1].x, of one color (cinfo[2b]: the first
// one's index in its color, the color's cluster count and first constraint); the summed
// displacement of the cluster's slot k goes to disp[k * nrows + 64 b + t] (one coalesced store per
// slot).
// MODE kWarmFeatures (capture) and kWarmLearned also form the cluster's network features (see
// launch_hex_warm_start) and store the multipliers they start from as the history; kWarmFeatures
// writes the features to raw[row], kWarmLearned runs the network on the workgroup's 64 rows (two
// 32-row matrix-core tiles, rows staged through LDS) and adds its predicted change to the
// multipliers before the displacement.
constexpr int kWarmPlain = kWarmNetNone, kWarmFeatures = kWarmNetFeatures, kWarmLearned = kWarmNetLearned;
constexpr int kFeatures = 39;  // used network inputs of a hex cluster

template <int MODE>
__device__ __forceinline__ void hex_warm_disp(tet_group_dev const& g, cluster_dev const& cl, const pos4* xi,
                                              const uint4* __restrict__ cinfo, uint32_t nrows, float inv_dt2,
                                              int expl, float2 scale, float4* __restrict__ disp,
                                              warm_net_dev const& net)
{
    uint4 const info = cinfo[2 * blockIdx.x];
    uint32_t const used = cinfo[2 * blockIdx.x + 1].x;
    if (MODE != kWarmLearned && threadIdx.x >= used) return;
    if (used == 0) return;  // a padding workgroup (uniform)
    bool const active = threadIdx.x < used;
    uint32_t const r = blockIdx.x * 64u + threadIdx.x;
    // inactive lanes (learned mode only) run the network on zero rows and touch no memory of theirs
    uint32_t const i = info.x + (active ? threadIdx.x : 0u), count = info.y, tbase = info.z;
    uint32_t const ci = info.w + (active ? threadIdx.x : 0u);
    int4 const va = cl.v0[ci], vb = cl.v1[ci];
    int const ids[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
    // every tet's material and multipliers are loaded before the first multiplier is stored back
    float4 A[kHexTets], B[kHexTets];
    float2 C[kHexTets], L[kHexTets], H[kHexTets];
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) {
        uint32_t const t = tbase + uint32_t(k) * count + i;
        A[k] = g.m0[t];
        B[k] = g.m1[t];
        C[k] = nh_m2(g, t);
        L[k] = reinterpret_cast<const float2*>(g.lambda)[t];
        if (MODE != kWarmPlain) H[k] = net.hist[t];
    }
    pos4 P[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) P[k] = xi[ids[k]];
    if (MODE != kWarmPlain) {
        // features: alpha~ lambda_prev (12), alpha~ lambda_prev2 (12), C at the predicted positions
        // (12), log alpha~ of tet 0 (deviatoric, hydrostatic), the body-size feature
        float f[kFeatures];
#pragma unroll
        for (int k = 0; k < kHexTets; ++k) {
            int const s0 = kHexSlots[k][0], s1 = kHexSlots[k][1], s2 = kHexSlots[k][2], s3 = kHexSlots[k][3];
            v3 const e[3] = {sub(P[s0], P[s3]), sub(P[s1], P[s3]), sub(P[s2], P[s3])};
            float2 const cv = nh_values(e, load_nh(A[k], B[k]), C[k].y);
            float const atd = B[k].w * inv_dt2, ath = C[k].x * inv_dt2;
            f[2 * k] = atd * L[k].x;
            f[2 * k + 1] = ath * L[k].y;
            f[12 + 2 * k] = atd * H[k].x;
            f[13 + 2 * k] = ath * H[k].y;
            f[24 + 2 * k] = cv.x;
            f[25 + 2 * k] = cv.y;
        }
        f[36] = logf(B[0].w * inv_dt2);
        f[37] = logf(C[0].x * inv_dt2);
        f[38] = net.body_size[tbase + i];
        if (active) {
#pragma unroll
            for (int k = 0; k < kHexTets; ++k) net.hist[tbase + uint32_t(k) * count + i] = L[k];
        }
        if (MODE == kWarmFeatures) {
            float* o = net.raw + std::size_t(r) * mlp::kMaxInputs;
#pragma unroll
            for (int m = 0; m < mlp::kMaxInputs; ++m) o[m] = m < kFeatures ? f[m] : 0.0f;
        } else {
            // the network corrects carried-over multipliers: a cluster whose multipliers or their
            // history are all zero (the first two substeps after a reset) keeps the previous start
            bool carried = false, history = false;
#pragma unroll
            for (int m = 0; m < 2 * kHexTets; ++m) {
                carried = carried || f[m] != 0.0f;
                history = history || f[2 * kHexTets + m] != 0.0f;
            }
            using mlp::bf16x8;
            __shared__ bf16x8 rows_lds[64 * mlp::kIn / 8];  // [row][48] BF16 inputs, then [row][12] FP32 outputs
            __shared__ float bias_lds[mlp::kBiasTotal];
            for (int m = threadIdx.x; m < int(mlp::kBiasTotal); m += 64) bias_lds[m] = net.bias[m];
            {
                bf16x8 v[mlp::kIn / 8];
#pragma unroll
                for (int m = 0; m < mlp::kIn; ++m)
                    v[m / 8][m % 8] = static_cast<mlp::bf16>(
                        active && m < kFeatures ? (f[m] - net.norm->mean[m]) * net.norm->inv_std[m] : 0.0f);
#pragma unroll
                for (int m = 0; m < mlp::kIn / 8; ++m) rows_lds[threadIdx.x * (mlp::kIn / 8) + m] = v[m];
            }
            __syncthreads();
            int const lane = threadIdx.x, h = lane >> 5, c = lane & 31;
            // the two 32-row tiles one after the other, not unrolled (each pass reads the weights
            // again, from L2): half the accumulators and activations live at a time. A tile's
            // outputs ([row][12] FP32) overwrite only inputs of rows already read (the workgroup is
            // one wave: tile t's 32 rows of inputs are in registers before its outputs are written;
            // its outputs occupy bytes [1536 t, 1536 t + 1536), tile 1's inputs start at 3072).
            float* out = reinterpret_cast<float*>(rows_lds);
#pragma unroll 1
            for (int t2 = 0; t2 < 2; ++t2) {
                bf16x8 xin[1][mlp::kQ1];
#pragma unroll
                for (int q = 0; q < mlp::kQ1; ++q) xin[0][q] = rows_lds[(32 * t2 + c) * (mlp::kIn / 8) + 2 * q + h];
                mlp::f32x16 yt[1];
                mlp::forward_tiles<1>(reinterpret_cast<const bf16x8*>(net.pack), bias_lds, xin, lane, yt);
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    int const o = mlp::acc_unit(j, h);
                    if (o < 2 * kHexTets) out[(32 * t2 + c) * (2 * kHexTets) + o] = yt[0][j];
                }
            }
            __syncthreads();
            if (!active) return;
            // the material and the positions again (the same values): read back after the network
            // instead of held across it, so that the network runs with the registers it needs
#pragma unroll
            for (int k = 0; k < kHexTets; ++k) {
                uint32_t const t = tbase + uint32_t(k) * count + i;
                A[k] = g.m0[t];
                B[k] = g.m1[t];
                C[k] = nh_m2(g, t);
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) P[k] = xi[ids[k]];
            // the predicted change of alpha~ lambda over the substep, back to multipliers
            if (carried && history) {
#pragma unroll
                for (int k = 0; k < kHexTets; ++k) {
                    float const yd = out[threadIdx.x * (2 * kHexTets) + 2 * k] * net.norm->out_scale[2 * k];
                    float const yh = out[threadIdx.x * (2 * kHexTets) + 2 * k + 1] * net.norm->out_scale[2 * k + 1];
                    L[k].x += yd / (B[k].w * inv_dt2);
                    L[k].y += yh / (C[k].x * inv_dt2);
                }
            }
        }
    }
    v3 D[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) D[k] = mk(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) {
        int const s0 = kHexSlots[k][0], s1 = kHexSlots[k][1], s2 = kHexSlots[k][2], s3 = kHexSlots[k][3];
        pos4 const Q[4] = {P[s0], P[s1], P[s2], P[s3]};
        v3 d[4];
        neohookean_displacement(Q, A[k], B[k], C[k], inv_dt2, expl != 0, scale, L[k], d);
        D[s0] = D[s0] + d[0];
        D[s1] = D[s1] + d[1];
        D[s2] = D[s2] + d[2];
        D[s3] = D[s3] + d[3];
    }
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) reinterpret_cast<float2*>(g.lambda)[tbase + uint32_t(k) * count + i] = L[k];
#pragma unroll
    for (int k = 0; k < 8; ++k) disp[std::size_t(k) * nrows + r] = make_float4(D[k].x, D[k].y, D[k].z, 0.0f);
}

template <int MODE>
__global__ void __launch_bounds__(64) k_hex_warm_disp(tet_group_dev g, cluster_dev cl, const pos4* __restrict__ xi,
                                                      const uint4* __restrict__ cinfo, uint32_t nrows, float inv_dt2,
                                                      int expl, float2 scale, float4* __restrict__ disp,
                                                      warm_net_dev net)
{
    hex_warm_disp<MODE>(g, cl, xi, cinfo, nrows, inv_dt2, expl, scale, disp, net);
}

// The learned variant, held to at most 256 registers a lane (2 waves per SIMD); its positions are
// read through a pointer without restrict so that the reads after the network are real loads.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_hex_warm_learned(tet_group_dev g, cluster_dev cl, const pos4* xi, const uint4* __restrict__ cinfo, uint32_t nrows,
                   float inv_dt2, int expl, float2 scale, float4* __restrict__ disp, warm_net_dev net)
{
    hex_warm_disp<kWarmLearned>(g, cl, xi, cinfo, nrows, inv_dt2, expl, scale, disp, net);
}

// Training rows of a hex group's captured substep: for j < rows, cluster ci = offset + j * stride,
// its row rowof[ci] of raw (the features at the substep start) and, as targets, the change of
// alpha~ lambda of its 12 multipliers over the substep (the final multipliers against feature
// values 0..11); a row holding a non-finite value is stored as zeros.
__global__ void k_hex_capture(tet_group_dev g, const uint4* __restrict__ cinfo, const uint32_t* __restrict__ rowof,
                              const float* __restrict__ raw, uint32_t stride, uint32_t offset, float inv_dt2,
                              float* __restrict__ ds_x, float* __restrict__ ds_t, uint64_t base, uint32_t rows)
{
    uint32_t const j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= rows) return;
    uint32_t const row = rowof[offset + j * stride];
    uint4 const info = cinfo[2 * (row / 64)];
    uint32_t const i = info.x + row % 64, count = info.y, tbase = info.z;
    const float* f = raw + std::size_t(row) * mlp::kMaxInputs;
    float t[2 * kHexTets];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) {
        uint32_t const tk = tbase + uint32_t(k) * count + i;
        float2 const lf = reinterpret_cast<const float2*>(g.lambda)[tk];
        t[2 * k] = g.m1[tk].w * inv_dt2 * lf.x - f[2 * k];
        t[2 * k + 1] = nh_m2(g, tk).x * inv_dt2 * lf.y - f[2 * k + 1];
        ok = ok && isfinite(t[2 * k]) && isfinite(t[2 * k + 1]);
    }
#pragma unroll
    for (int m = 0; m < kFeatures; ++m) ok = ok && isfinite(f[m]);
    std::size_t const o = base + j;
#pragma unroll
    for (int m = 0; m < mlp::kMaxInputs; ++m) ds_x[o * mlp::kMaxInputs + m] = ok ? f[m] : 0.0f;
#pragma unroll
    for (int m = 0; m < mlp::kMaxOutputs; ++m) ds_t[o * mlp::kMaxOutputs + m] = ok && m < 2 * kHexTets ? t[m] : 0.0f;
}

// x_p += the displacements of p's cluster slots, in FP64, in the order of the list (execution
// order of the clusters: the order of the colored pass).
__global__ void k_gather_disp(const uint32_t* __restrict__ off, const uint32_t* __restrict__ ent,
                              const float4* __restrict__ disp, pos4* __restrict__ xi, uint32_t n)
{
    uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t const b = off[p], e = off[p + 1];
    if (b == e) return;
    pos4 x = xi[p];
    for (uint32_t j = b; j < e; ++j) {
        float4 const d = disp[ent[j]];
        x.x += double(d.x);
        x.y += double(d.y);
        x.z += double(d.z);
    }
    xi[p] = x;
}

// The same with the list as a dense row of 8 slot indices per particle (unused: 0xffffffff), when
// no particle is in more than 8 clusters (hex lattices): the 8 loads of a row go out together and
// the unused ones add an exact 0.
// With surface_of (particle -> surface vertex or -1), the first iteration's contact projections of
// the surface vertices follow in the same thread, as k_contact(first) would run them next.
// With prow, thread r handles particle prow[r] (rows [0, n); padding rows hold 0xffffffff).
__global__ void k_gather_disp8(const uint4* __restrict__ ent, const float4* __restrict__ disp, pos4* __restrict__ xi,
                               uint32_t n, const int* __restrict__ surface_of, contact_dev c, float c_at,
                               const uint32_t* __restrict__ prow)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (prow) {
        p = prow[p];
        if (p == 0xffffffffu) return;
    }
    uint4 const a = ent[2 * std::size_t(p)], b = ent[2 * std::size_t(p) + 1];
    uint32_t const e[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    int const sv = surface_of ? surface_of[p] : -1;
    if (e[0] == 0xffffffffu && sv < 0) return;
    float4 d[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) d[m] = disp[e[m] != 0xffffffffu ? e[m] : 0u];
    pos4 x = xi[p];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        bool const ok = e[m] != 0xffffffffu;
        x.x += ok ? double(d[m].x) : 0.0;
        x.y += ok ? double(d[m].y) : 0.0;
        x.z += ok ? double(d[m].z) : 0.0;
    }
    if (sv >= 0) contact_vertex_at(c, x, uint32_t(sv), c_at, true);
    xi[p] = x;
}

// FIRST: the first iteration of a substep, a compile-time switch for the reason given at
// k_neohookean_hex.
template <int BLOCK, bool FIRST>
__global__ void __launch_bounds__(BLOCK) k_green_hex(tet_group_dev g, cluster_dev cl, pos4* xi,
                                                     const pos4* __restrict__ xn, uint32_t begin, uint32_t count,
                                                     uint32_t tbase, xcd_split sp, float alpha_tilde, float gamma)
{
    uint32_t const x = blockIdx.x % kXcds, local = (blockIdx.x / kXcds) * BLOCK + threadIdx.x;
    if (local >= sp.off[x + 1] - sp.off[x]) return;
    uint32_t const i = sp.off[x] + local;
    uint32_t const c = begin + i;
    int4 const va = cl.v0[c], vb = cl.v1[c];
    int const ids[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
    float4 A[kHexTets], B[kHexTets], Cm[kHexTets];
    float L[kHexTets];
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) {
        uint32_t const t = tbase + uint32_t(k) * count + i;
        A[k] = g.m0[t];
        B[k] = g.m1[t];
        Cm[k] = g.m2[t];
        L[k] = FIRST ? 0.0f : g.lambda[t];
    }
    pos4 P[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) P[k] = xi[ids[k]];
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) {
        int const s0 = kHexSlots[k][0], s1 = kHexSlots[k][1], s2 = kHexSlots[k][2], s3 = kHexSlots[k][3];
        pos4 Q[4] = {P[s0], P[s1], P[s2], P[s3]};
        int const id[4] = {ids[s0], ids[s1], ids[s2], ids[s3]};
        auto xn_of = [&](int q) { return xn[id[q]]; };
        // on the early-out the multiplier stays as loaded (0 on the first iteration)
        green_project(Q, xn_of, A[k], B[k], Cm[k], alpha_tilde, gamma, L[k]);
        P[s0] = Q[0];
        P[s1] = Q[1];
        P[s2] = Q[2];
        P[s3] = Q[3];
    }
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) g.lambda[tbase + uint32_t(k) * count + i] = L[k];
#pragma unroll
    for (int k = 0; k < 8; ++k) xi[ids[k]] = P[k];
}

// ---------------------------------------------------------------- distance constraint

// distance_constraint.cpp:24-53 (no guard against zero length, as there).
__global__ void __launch_bounds__(256) k_distance(distance_group_dev g, pos4* xi, const pos4* __restrict__ xn,
                                                  uint32_t begin, uint32_t count, float alpha_tilde, float gamma,
                                                  int first)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t const t = begin + i;
    uint2 const id = g.idx[t];
    float const d = g.rest[t];
    float const lm = first ? 0.0f : g.lambda[t];
    pos4 P1 = xi[id.x], P2 = xi[id.y];
    v3 const diff = sub(P1, P2);
    float const length = sqrtf(dot(diff, diff));
    v3 const n = (1.0f / length) * diff;
    float const C = length - d;
    float gdd = 0.0f;
    if (gamma != 0.0f) gdd = dot(n, sub(P1, xn[id.x])) - dot(n, sub(P2, xn[id.y]));
    float const w1 = wof(P1), w2 = wof(P2);
    float const dl = (-(C + alpha_tilde * lm) + gamma * gdd) / ((1.0f + gamma) * (w1 + w2) + alpha_tilde);
    g.lambda[t] = lm + dl;
    add_to(P1, (w1 * dl) * n);
    add_to(P2, -(w2 * dl) * n);
    xi[id.x] = P1;
    xi[id.y] = P2;
}

// ---------------------------------------------------------------- contacts

// One thread per surface vertex: contacts against every SDF, in SDF order, from the positions at
// the start of the frame (bvh_model.cpp:72-95, contact_handler.cpp:39-52). A vertex keeps at most
// c.slots contacts; more are counted in c.overflow.
__global__ void k_detect(contact_dev c, const pos4* __restrict__ x, const sdf_dev* __restrict__ sdfs, uint32_t nsdf)
{
    uint32_t const s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c.S) return;
    pos4 const p = x[c.surf[s]];
    uint32_t cnt = 0;
    for (uint32_t k = 0; k < nsdf; ++k) {
        double q[3], n[3];
        if (!sdf_contact(sdfs[k], p, q, n)) continue;
        if (cnt == c.slots) {
            atomicAdd(c.overflow, 1u);
            continue;
        }
        store_contact(c, s, cnt++, k, q, n);
    }
    c.count[s] = cnt;
}

// collision_constraint.cpp:21-41, all contacts of one particle in SDF order by one thread.
__global__ void k_snapshot_surface(const uint32_t* __restrict__ surf, const pos4* __restrict__ xi,
                                   pos4* __restrict__ snap, uint32_t S)
{
    uint32_t const s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) snap[s] = xi[surf[s]];
}

__global__ void k_contact(contact_dev c, pos4* xi, float alpha_tilde, int first)
{
    uint32_t const s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c.S) return;
    if (c.count[s] == 0) return;
    uint32_t const vi = c.surf[s];
    pos4 P = xi[vi];
    contact_vertex_at(c, P, s, alpha_tilde, first != 0);
    xi[vi] = P;
}

// ---------------------------------------------------------------- reductions (FP64)

template <int BLOCK>
__device__ __forceinline__ double block_sum(double v)
{
    static_assert(BLOCK % 64 == 0, "wave64 blocks");
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __shared__ double part[BLOCK / 64];
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < BLOCK / 64; ++i) r += part[i];
    __syncthreads();
    return r;
}

constexpr int kRedBlock = 256;

__global__ void __launch_bounds__(kRedBlock) k_green_residual(tet_group_dev g, const pos4* __restrict__ x,
                                                              uint32_t begin, uint32_t count, float alpha_tilde,
                                                              double* partial)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    double h2 = 0.0;
    if (i < count) {
        uint32_t const t = begin + i;
        int4 const id = g.idx[t];
        float4 const c = g.m2[t];
        pos4 const p3 = x[id.w];
        v3 f[4];
        float const C = green_core(sub(x[id.x], p3), sub(x[id.y], p3), sub(x[id.z], p3), load_rows(g.m0[t], g.m1[t], c),
                                   c.y, c.z, c.w, f);
        double const h = double(C) + double(alpha_tilde) * double(g.lambda[t]);
        h2 = h * h;
    }
    double const s = block_sum<kRedBlock>(h2);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ||F||_F^2 and det F - gamma0 of neo-Hookean tet id at positions x, in FP64 from the FP64 edge
// vectors and the FP32 rest data (the forms at nh_rest).
__device__ __forceinline__ void nh_constraints_f64(pos4 p0, pos4 p1, pos4 p2, pos4 p3, nh_rest const& q, float gamma0,
                                                   double& f2, double& det_minus)
{
    double const e[3][3] = {{p0.x - p3.x, p0.y - p3.y, p0.z - p3.z},
                            {p1.x - p3.x, p1.y - p3.y, p1.z - p3.z},
                            {p2.x - p3.x, p2.y - p3.y, p2.z - p3.z}};
    auto d = [&](int m, int n) { return e[m][0] * e[n][0] + e[m][1] * e[n][1] + e[m][2] * e[n][2]; };
    f2 = double(q.q00) * d(0, 0) + double(q.q11) * d(1, 1) + double(q.q22) * d(2, 2) +
         2.0 * (double(q.q01) * d(0, 1) + double(q.q02) * d(0, 2) + double(q.q12) * d(1, 2));
    f2 = f2 > 0.0 ? f2 : 0.0;
    double const vol = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) +
                       e[0][1] * (e[1][2] * e[2][0] - e[1][0] * e[2][2]) +
                       e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
    det_minus = double(q.detr) * vol - double(gamma0);
}

__device__ __forceinline__ void nh_constraints_f64(const pos4* __restrict__ x, int4 id, nh_rest const& q, float gamma0,
                                                   double& f2, double& det_minus)
{
    nh_constraints_f64(x[id.x], x[id.y], x[id.z], x[id.w], q, gamma0, f2, det_minus);
}

// h^2 of both constraints of neo-Hookean tet t at positions P0..P3 (FP64).
__device__ __forceinline__ double nh_residual2(tet_group_dev const& g, uint32_t t, pos4 p0, pos4 p1, pos4 p2, pos4 p3,
                                               float inv_dt2)
{
    float4 const b = g.m1[t];
    float2 const c = nh_m2(g, t);
    double cd, ch;
    nh_constraints_f64(p0, p1, p2, p3, load_nh(g.m0[t], b), c.y, cd, ch);
    float2 const lm = reinterpret_cast<const float2*>(g.lambda)[t];
    cd = sqrt(cd);
    double const hd = cd + double(b.w) * double(inv_dt2) * double(lm.x);
    double const hh = ch + double(c.x) * double(inv_dt2) * double(lm.y);
    return hd * hd + hh * hh;
}

// The same over a hex-cluster group, one thread per row of the all-color rows (a cluster of any
// color, or padding): the cluster's eight particles are read once for its six tets. One 64-row
// workgroup per record pair, so that workgroup b runs on the XCD of its rows' region (b mod 8).
constexpr int kHexRedBlock = 64;
__global__ void __launch_bounds__(kHexRedBlock) k_neohookean_hex_residual(tet_group_dev g, cluster_dev cl,
                                                                       const pos4* __restrict__ x,
                                                                       const uint4* __restrict__ cinfo,
                                                                       uint32_t nrows, float inv_dt2,
                                                                       double* partial)
{
    // the workgroup's run of clusters, as at k_hex_warm_disp
    uint4 const info = cinfo[2 * blockIdx.x];
    uint32_t const used = cinfo[2 * blockIdx.x + 1].x;
    double h2 = 0.0;
    if (threadIdx.x < used) {
        uint32_t const ci = info.w + threadIdx.x, i = info.x + threadIdx.x;
        int4 const va = cl.v0[ci], vb = cl.v1[ci];
        pos4 const P[8] = {x[va.x], x[va.y], x[va.z], x[va.w], x[vb.x], x[vb.y], x[vb.z], x[vb.w]};
#pragma unroll
        for (int k = 0; k < kHexTets; ++k)
            h2 += nh_residual2(g, info.z + uint32_t(k) * info.y + i, P[kHexSlots[k][0]], P[kHexSlots[k][1]],
                               P[kHexSlots[k][2]], P[kHexSlots[k][3]], inv_dt2);
    }
    double const s = block_sum<kHexRedBlock>(h2);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRedBlock) k_neohookean_residual(tet_group_dev g, const pos4* __restrict__ x,
                                                                   uint32_t begin, uint32_t count, float inv_dt2,
                                                                   double* partial)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    double h2 = 0.0;
    if (i < count) {
        uint32_t const t = begin + i;
        int4 const id = g.idx[t];
        h2 = nh_residual2(g, t, x[id.x], x[id.y], x[id.z], x[id.w], inv_dt2);
    }
    double const s = block_sum<kRedBlock>(h2);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRedBlock) k_distance_residual(distance_group_dev g, const pos4* __restrict__ x,
                                                                 uint32_t begin, uint32_t count, float alpha_tilde,
                                                                 double* partial)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    double h2 = 0.0;
    if (i < count) {
        uint32_t const t = begin + i;
        uint2 const id = g.idx[t];
        pos4 const a = x[id.x], b = x[id.y];
        double const dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        double const h =
            (sqrt(dx * dx + dy * dy + dz * dz) - double(g.rest[t])) + double(alpha_tilde) * double(g.lambda[t]);
        h2 = h * h;
    }
    double const s = block_sum<kRedBlock>(h2);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRedBlock) k_particle_energy(const pos4* __restrict__ x, const float4* __restrict__ v,
                                                               const float* __restrict__ mass, uint32_t n, float gx,
                                                               float gy, float gz, int force, double* partial)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0, pe = 0, px = 0, py = 0, pz = 0;
    if (i < n) {
        double const m = mass[i];
        pos4 const p = x[i];
        float4 const u = v[i];
        ke = 0.5 * m * (double(u.x) * u.x + double(u.y) * u.y + double(u.z) * u.z);
        pe = -(force ? 1.0 : m) * (double(gx) * p.x + double(gy) * p.y + double(gz) * p.z);
        px = m * u.x;
        py = m * u.y;
        pz = m * u.z;
    }
    double r[5] = {block_sum<kRedBlock>(ke), block_sum<kRedBlock>(pe), block_sum<kRedBlock>(px),
                   block_sum<kRedBlock>(py), block_sum<kRedBlock>(pz)};
    if (threadIdx.x == 0) {
        double* o = partial + 6 * blockIdx.x;
        o[0] = r[0];
        o[1] = r[1];
        o[2] = 0.0;
        o[3] = r[2];
        o[4] = r[3];
        o[5] = r[4];
    }
}

__global__ void __launch_bounds__(kRedBlock) k_green_energy(tet_group_dev g, const pos4* __restrict__ x, uint32_t count,
                                                            double* partial)
{
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (t < count) {
        int4 const id = g.idx[t];
        float4 const c = g.m2[t];
        pos4 const p3 = x[id.w];
        v3 f[4];
        e = green_core(sub(x[id.x], p3), sub(x[id.y], p3), sub(x[id.z], p3), load_rows(g.m0[t], g.m1[t], c), c.y, c.z,
                       c.w, f);
    }
    double const s = block_sum<kRedBlock>(e);
    if (threadIdx.x == 0) {
        double* o = partial + 6 * blockIdx.x;
        o[0] = o[1] = o[3] = o[4] = o[5] = 0.0;
        o[2] = s;
    }
}

__global__ void __launch_bounds__(kRedBlock) k_neohookean_energy(tet_group_dev g, const pos4* __restrict__ x,
                                                                 uint32_t count, double* partial)
{
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (t < count) {
        int4 const id = g.idx[t];
        float4 const b = g.m1[t];
        float2 const c = nh_m2(g, t);
        double ic, j;
        nh_constraints_f64(x, id, load_nh(g.m0[t], b), c.y, ic, j);
        // mu |V0| = 1 / m1.w, lambda |V0| = 1 / m2.x
        e = 0.5 / double(b.w) * (ic - 3.0) + 0.5 / double(c.x) * j * j;
    }
    double const s = block_sum<kRedBlock>(e);
    if (threadIdx.x == 0) {
        double* o = partial + 6 * blockIdx.x;
        o[0] = o[1] = o[3] = o[4] = o[5] = 0.0;
        o[2] = s;
    }
}

// ---------------------------------------------------------------- Jones-Plassmann coloring

__device__ __forceinline__ uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Priority of element t: (previous color, hash32(t)) -- higher first; ties are impossible because
// hash32 is a bijection on 32-bit indices.
__global__ void k_jp_keys(uint32_t n, const uint32_t* __restrict__ prev, uint64_t* __restrict__ key)
{
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    key[t] = (uint64_t(prev ? prev[t] : 0u) << 32) | hash32(t);
}

// An element whose key exceeds the keys of all its uncolored neighbours takes the smallest color
// unused by its colored neighbours. Over all rounds this equals sequential greedy coloring in
// descending key order: when an element is colored, exactly its higher-key neighbours are.
__global__ void k_jp_round(uint32_t n, uint32_t arity, const uint32_t* __restrict__ elem_off,
                           const uint32_t* __restrict__ verts, const uint32_t* __restrict__ inc_off,
                           const uint32_t* __restrict__ inc, const uint64_t* __restrict__ key,
                           const uint32_t* __restrict__ color_in, uint32_t* __restrict__ color_out, uint32_t* remaining,
                           uint32_t* overflow)
{
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t const mine = color_in[t];
    if (mine != kUncolored) {
        color_out[t] = mine;
        return;
    }
    uint64_t const kt = key[t];
    uint64_t used[kMaxColors / 64] = {0, 0, 0, 0};
    size_t const a0 = elem_off ? elem_off[t] : size_t(t) * arity;
    size_t const a1 = elem_off ? elem_off[t + 1] : a0 + arity;
    for (size_t a = a0; a < a1; ++a) {
        uint32_t const v = verts[a];
        uint32_t const e1 = inc_off[v + 1];
        for (uint32_t e = inc_off[v]; e < e1; ++e) {
            uint32_t const u = inc[e];
            if (u == t) continue;
            uint32_t const cu = color_in[u];
            if (cu == kUncolored) {
                if (key[u] > kt) {
                    color_out[t] = kUncolored;
                    atomicAdd(remaining, 1u);
                    return;
                }
            } else if (cu < kMaxColors) {
                used[cu >> 6] |= (uint64_t(1) << (cu & 63));
            }
        }
    }
    uint32_t c = kMaxColors;
#pragma unroll
    for (int w = 0; w < int(kMaxColors / 64); ++w) {
        if (c == kMaxColors && ~used[w] != 0) c = uint32_t(w) * 64 + uint32_t(__builtin_ctzll(~used[w]));
    }
    if (c >= kMaxColors) {
        atomicAdd(overflow, 1u);
        c = kMaxColors - 1;
    }
    color_out[t] = c;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t value, uint32_t n)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

}  // namespace

// ---------------------------------------------------------------- launchers

void launch_predict(hipStream_t s, pos4* xi, const pos4* xn, float4* v, uint32_t n, float gx, float gy, float gz,
                    int force, float dt, int block)
{
    if (n) k_predict<<<grid_for(n, block), block, 0, s>>>(xi, xn, v, n, gx, gy, gz, force, dt);
}

void launch_commit(hipStream_t s, const pos4* xi, pos4* xn, float4* v, uint32_t n, float dt, int block)
{
    if (n) k_commit<<<grid_for(n, block), block, 0, s>>>(xi, xn, v, n, dt);
}

void launch_commit_predict(hipStream_t s, pos4* xi, pos4* xn, float4* v, uint32_t n, float gx, float gy, float gz,
                           int force, float dt, int block, const uint32_t* prow, uint32_t nprows)
{
    if (prow && nprows)
        k_commit_predict<<<grid_for(nprows, kParticleRowBlock), kParticleRowBlock, 0, s>>>(xi, xn, v, nprows, gx, gy,
                                                                                          gz, force, dt, prow);
    else if (n)
        k_commit_predict<<<grid_for(n, block), block, 0, s>>>(xi, xn, v, n, gx, gy, gz, force, dt, nullptr);
}

void launch_green(hipStream_t s, tet_group_dev g, pos4* xi, const pos4* xn, uint32_t begin, uint32_t count,
                  float alpha_tilde, float gamma, int first, int block)
{
    if (count) k_green<<<grid_for(count, block), block, 0, s>>>(g, xi, xn, begin, count, alpha_tilde, gamma, first);
}

void launch_neohookean(hipStream_t s, tet_group_dev g, pos4* xi, const pos4* xn, uint32_t begin, uint32_t count,
                       float inv_dt2, float beta_dt, int mode, int block, float hyd_scale)
{
    if (count)
        k_neohookean<<<grid_for(count, block), block, 0, s>>>(g, xi, xn, begin, count, inv_dt2, beta_dt, mode,
                                                              hyd_scale);
}

void launch_distance(hipStream_t s, distance_group_dev g, pos4* xi, const pos4* xn, uint32_t begin, uint32_t count,
                     float alpha_tilde, float gamma, int first, int block)
{
    if (count)
        k_distance<<<grid_for(count, block), block, 0, s>>>(g, xi, xn, begin, count, alpha_tilde, gamma, first);
}

void launch_green_clusters(hipStream_t s, tet_group_dev g, cluster_dev cl, pos4* xi, const pos4* xn, uint32_t begin,
                           uint32_t count, float alpha_tilde, float gamma, int first, int block)
{
    if (!count) return;
    if (block <= 64)
        k_green_clusters<64><<<grid_for(count, 64), 64, 0, s>>>(g, cl, xi, xn, begin, count, alpha_tilde, gamma, first);
    else
        k_green_clusters<128><<<grid_for(count, 128), 128, 0, s>>>(g, cl, xi, xn, begin, count, alpha_tilde, gamma, first);
}

void launch_neohookean_clusters(hipStream_t s, tet_group_dev g, cluster_dev cl, pos4* xi, const pos4* xn,
                                uint32_t begin, uint32_t count, float inv_dt2, float beta_dt, int mode, int block,
                                float hyd_scale)
{
    if (!count) return;
    if (mode >= kApplyWarmStart) {
        k_neohookean_clusters_apply<<<grid_for(count, 128), 128, 0, s>>>(g, cl, xi, xn, begin, count, inv_dt2,
                                                                         mode == kApplyExplicit,
                                                                         make_float2(beta_dt, hyd_scale));
        return;
    }
    if (block <= 64)
        k_neohookean_clusters<64><<<grid_for(count, 64), 64, 0, s>>>(g, cl, xi, xn, begin, count, inv_dt2, beta_dt, mode);
    else
        k_neohookean_clusters<128><<<grid_for(count, 128), 128, 0, s>>>(g, cl, xi, xn, begin, count, inv_dt2, beta_dt,
                                                                       mode);
}

void launch_neohookean_hex(hipStream_t s, tet_group_dev g, cluster_dev cl, pos4* xi, const pos4* xn, uint32_t begin,
                           uint32_t count, uint32_t tbase, xcd_split const& sp, float inv_dt2, float beta_dt, int mode,
                           int block, float hyd_scale)
{
    if (!count) return;
    if (mode >= kApplyWarmStart) {
        k_neohookean_hex_apply<64><<<grid_for(count, 64), 64, 0, s>>>(g, cl, xi, xn, begin, count, tbase, inv_dt2,
                                                                       mode == kApplyExplicit,
                                                                       make_float2(beta_dt, hyd_scale));
        return;
    }
    // block 64 / 128: material prefetched; kHexStreamed: 64 threads, material loaded per tet, at
    // least 3 waves per SIMD (2 with damping, which would spill at 3). Damping (beta != 0) takes the
    // variants that carry the xn differences.
    bool const damped = beta_dt != 0.0f, first = mode == kSolveFirst;
    auto go = [&](auto kernel, int b) {
        uint32_t most = 0;  // workgroups per XCD
        for (int x = 0; x < kXcds; ++x) most = std::max(most, uint32_t(grid_for(sp.off[x + 1] - sp.off[x], b)));
        if (most) kernel<<<most * kXcds, b, 0, s>>>(g, cl, xi, xn, begin, count, tbase, sp, inv_dt2, beta_dt);
    };
    if (block == kHexStreamed) {
        if (damped)
            first ? go(k_neohookean_hex<64, false, 2, true, true>, 64) : go(k_neohookean_hex<64, false, 2, true, false>, 64);
        else
            first ? go(k_neohookean_hex<64, false, 3, false, true>, 64)
                  : go(k_neohookean_hex<64, false, 3, false, false>, 64);
    } else if (block <= 64) {
        if (damped)
            first ? go(k_neohookean_hex<64, true, 1, true, true>, 64) : go(k_neohookean_hex<64, true, 1, true, false>, 64);
        else
            first ? go(k_neohookean_hex<64, true, 1, false, true>, 64) : go(k_neohookean_hex<64, true, 1, false, false>, 64);
    } else {
        if (damped)
            first ? go(k_neohookean_hex<128, true, 1, true, true>, 128)
                  : go(k_neohookean_hex<128, true, 1, true, false>, 128);
        else
            first ? go(k_neohookean_hex<128, true, 1, false, true>, 128)
                  : go(k_neohookean_hex<128, true, 1, false, false>, 128);
    }
}

void launch_green_hex(hipStream_t s, tet_group_dev g, cluster_dev cl, pos4* xi, const pos4* xn, uint32_t begin,
                      uint32_t count, uint32_t tbase, xcd_split const& sp, float alpha_tilde, float gamma, int first,
                      int block)
{
    if (!count) return;
    int const b = block <= 64 ? 64 : 128;
    uint32_t most = 0;  // workgroups per XCD
    for (int x = 0; x < kXcds; ++x) most = std::max(most, uint32_t(grid_for(sp.off[x + 1] - sp.off[x], b)));
    if (!most) return;
    auto go = [&](auto kernel) { kernel<<<most * kXcds, b, 0, s>>>(g, cl, xi, xn, begin, count, tbase, sp, alpha_tilde, gamma); };
    if (b == 64)
        first ? go(k_green_hex<64, true>) : go(k_green_hex<64, false>);
    else
        first ? go(k_green_hex<128, true>) : go(k_green_hex<128, false>);
}

void launch_detect(hipStream_t s, contact_dev c, const pos4* x, const sdf_dev* sdfs, uint32_t nsdf, int block)
{
    if (c.S) k_detect<<<grid_for(c.S, block), block, 0, s>>>(c, x, sdfs, nsdf);
}

void launch_snapshot_surface(hipStream_t s, const uint32_t* surf, const pos4* xi, pos4* snap, uint32_t S)
{
    if (S) k_snapshot_surface<<<grid_for(S, 256), 256, 0, s>>>(surf, xi, snap, S);
}

void launch_contact(hipStream_t s, contact_dev c, pos4* xi, float alpha_tilde, int first, int block)
{
    if (c.S) k_contact<<<grid_for(c.S, block), block, 0, s>>>(c, xi, alpha_tilde, first);
}

void launch_green_residual(hipStream_t s, tet_group_dev g, const pos4* x, uint32_t begin, uint32_t count,
                           float alpha_tilde, double* partial, int)
{
    if (count) k_green_residual<<<grid_for(count, kRedBlock), kRedBlock, 0, s>>>(g, x, begin, count, alpha_tilde, partial);
}

void launch_neohookean_residual(hipStream_t s, tet_group_dev g, const pos4* x, uint32_t begin, uint32_t count,
                                float inv_dt2, double* partial, int)
{
    if (count)
        k_neohookean_residual<<<grid_for(count, kRedBlock), kRedBlock, 0, s>>>(g, x, begin, count, inv_dt2, partial);
}

__global__ void __launch_bounds__(256) k_copy_spans(copy_spans sp)
{
    uint64_t const tid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x, stride = uint64_t(gridDim.x) * blockDim.x;
    for (int k = 0; k < sp.n; ++k) {
        uint64_t const words = sp.bytes[k] / 16;
        uint4* d = static_cast<uint4*>(sp.dst[k]);
        const uint4* s = static_cast<const uint4*>(sp.src[k]);
        for (uint64_t i = tid; i < words; i += stride) d[i] = s[i];
        uint64_t const tail = (sp.bytes[k] - words * 16) / 4;  // the last < 16 bytes, 4 at a time
        if (tid < tail)
            static_cast<uint32_t*>(sp.dst[k])[words * 4 + tid] = static_cast<const uint32_t*>(sp.src[k])[words * 4 + tid];
    }
}

void launch_neohookean_hex_residual(hipStream_t s, tet_group_dev g, cluster_dev cl, const pos4* x,
                                    const uint4* cinfo, uint32_t nrows, float inv_dt2, double* partial)
{
    if (nrows)
        k_neohookean_hex_residual<<<grid_for(nrows, kHexRedBlock), kHexRedBlock, 0, s>>>(g, cl, x, cinfo, nrows,
                                                                                           inv_dt2, partial);
}

void launch_copy_spans(hipStream_t s, copy_spans const& spans)
{
    if (spans.n > 0) k_copy_spans<<<1024, 256, 0, s>>>(spans);
}

__global__ void __launch_bounds__(256) k_sum_partials(const double* __restrict__ partial, uint32_t n, double* out)
{
    double v = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) v += partial[i];
    double const t = block_sum<256>(v);
    if (threadIdx.x == 0) out[0] = t;
}

void launch_sum_partials(hipStream_t s, const double* partial, uint32_t n, double* out)
{
    k_sum_partials<<<1, 256, 0, s>>>(partial, n, out);
}

__global__ void k_publish_u32(const uint32_t* __restrict__ src, uint32_t* dst)
{
    if (threadIdx.x == 0) *reinterpret_cast<volatile uint32_t*>(dst) = *src;
}

void launch_publish_u32(hipStream_t s, const uint32_t* src, uint32_t* dst)
{
    k_publish_u32<<<1, 64, 0, s>>>(src, dst);
}

void launch_distance_residual(hipStream_t s, distance_group_dev g, const pos4* x, uint32_t begin, uint32_t count,
                              float alpha_tilde, double* partial, int)
{
    if (count)
        k_distance_residual<<<grid_for(count, kRedBlock), kRedBlock, 0, s>>>(g, x, begin, count, alpha_tilde, partial);
}

void launch_particle_energy(hipStream_t s, const pos4* x, const float4* v, const float* mass, uint32_t n, float gx,
                            float gy, float gz, int force, double* partial, int)
{
    if (n) k_particle_energy<<<grid_for(n, kRedBlock), kRedBlock, 0, s>>>(x, v, mass, n, gx, gy, gz, force, partial);
}

void launch_green_energy(hipStream_t s, tet_group_dev g, const pos4* x, uint32_t count, double* partial, int)
{
    if (count) k_green_energy<<<grid_for(count, kRedBlock), kRedBlock, 0, s>>>(g, x, count, partial);
}

void launch_neohookean_energy(hipStream_t s, tet_group_dev g, const pos4* x, uint32_t count, double* partial, int)
{
    if (count) k_neohookean_energy<<<grid_for(count, kRedBlock), kRedBlock, 0, s>>>(g, x, count, partial);
}

void launch_jp_round(hipStream_t s, uint32_t n, uint32_t arity, const uint32_t* elem_off, const uint32_t* verts,
                     const uint32_t* inc_off, const uint32_t* inc, const uint64_t* key, const uint32_t* color_in,
                     uint32_t* color_out, uint32_t* remaining, uint32_t* overflow, int block)
{
    if (n)
        k_jp_round<<<grid_for(n, block), block, 0, s>>>(n, arity, elem_off, verts, inc_off, inc, key, color_in,
                                                        color_out, remaining, overflow);
}

void launch_jp_keys(hipStream_t s, uint32_t n, const uint32_t* prev, uint64_t* key, int block)
{
    if (n) k_jp_keys<<<grid_for(n, block), block, 0, s>>>(n, prev, key);
}

void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t value, uint32_t n)
{
    if (n) k_fill_u32<<<grid_for(n, 256), 256, 0, s>>>(p, value, n);
}


void launch_hex_warm_start(hipStream_t s, tet_group_dev g, cluster_dev cl, const pos4* xt, pos4* xi,
                           const uint4* cinfo, uint32_t nrows, float inv_dt2, int expl, float2 scale,
                           float4* disp, const uint32_t* csr_off, const uint32_t* csr, uint32_t n,
                           const int* surface_of, contact_dev const& c, float c_at, const uint32_t* prow,
                           uint32_t nprows, int net_mode, warm_net_dev const& net)
{
    if (!nrows) return;
    if (net_mode == kWarmLearned)
        k_hex_warm_learned<<<grid_for(nrows, 64), 64, 0, s>>>(g, cl, xt, cinfo, nrows, inv_dt2, expl, scale,
                                                                         disp, net);
    else if (net_mode == kWarmFeatures)
        k_hex_warm_disp<kWarmFeatures><<<grid_for(nrows, 64), 64, 0, s>>>(g, cl, xt, cinfo, nrows, inv_dt2, expl,
                                                                          scale, disp, net);
    else
        k_hex_warm_disp<kWarmPlain><<<grid_for(nrows, 64), 64, 0, s>>>(g, cl, xt, cinfo, nrows, inv_dt2, expl, scale,
                                                                       disp, net);
    if (csr_off)
        k_gather_disp<<<grid_for(n, 256), 256, 0, s>>>(csr_off, csr, disp, xi, n);
    else if (prow && nprows)
        k_gather_disp8<<<grid_for(nprows, kParticleRowBlock), kParticleRowBlock, 0, s>>>(
            reinterpret_cast<const uint4*>(csr), disp, xi, nprows, surface_of, c, c_at, prow);
    else
        k_gather_disp8<<<grid_for(n, 256), 256, 0, s>>>(reinterpret_cast<const uint4*>(csr), disp, xi, n, surface_of,
                                                        c, c_at, nullptr);
}

void launch_hex_capture(hipStream_t s, tet_group_dev g, const uint4* cinfo, const uint32_t* rowof, const float* raw,
                        uint32_t stride, uint32_t offset, float inv_dt2, float* ds_x, float* ds_t, uint64_t base,
                        uint32_t rows)
{
    if (!rows) return;
    k_hex_capture<<<grid_for(rows, 256), 256, 0, s>>>(g, cinfo, rowof, raw, stride, offset, inv_dt2, ds_x, ds_t, base,
                                                      rows);
}

void launch_copy_pos(hipStream_t s, const pos4* a, pos4* b, uint32_t n)
{
    if (n) k_copy_pos<<<grid_for(n, 256), 256, 0, s>>>(a, b, n);
}



}  // namespace sbs
