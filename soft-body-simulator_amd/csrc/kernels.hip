// gfx950 kernels of the XPBD time step: prediction, per-color constraint projection (green strain,
// stable neo-Hookean, distance), SDF contact detection and projection, velocity update, FP64
// residual / energy reductions and Jones-Plassmann graph coloring.
//
// Particle state lives in HBM as float4 SoA streams: xi (working positions, w = inverse mass),
// xn (substep start positions = committed positions, w = inverse mass) and v (velocities).
// Constraint data of a group is stored in color order so that one color is one contiguous,
// coalesced range; the four particle gathers of a tet are 16-byte float4 loads.
#include "device_math.h"
#include "kernels.h"

namespace sbs {

namespace {

__device__ __forceinline__ float4 pack(v3 p, float w) { return make_float4(p.x, p.y, p.z, w); }

inline unsigned grid_for(uint32_t n, int block) { return (n + uint32_t(block) - 1) / uint32_t(block); }

// ---------------------------------------------------------------- prediction / commit

// timestep.cpp:35-43: v += a dt, xi = x + v dt. With gravity_is_force the acceleration is
// g * w (the reference's f = g, a = f w); otherwise g for every particle with w > 0.
__device__ __forceinline__ void predict_one(float4& xi, float4 const& x, float4& v, float gx, float gy, float gz,
                                            int force, float dt)
{
    float const w = x.w;
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
    xi = make_float4(x.x + v.x * dt, x.y + v.y * dt, x.z + v.z * dt, w);
}

__global__ void k_predict(float4* __restrict__ xi, const float4* __restrict__ xn, float4* __restrict__ v,
                          uint32_t n, float gx, float gy, float gz, int force, float dt)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 x = xn[i], vel = v[i], p;
    predict_one(p, x, vel, gx, gy, gz, force, dt);
    v[i] = vel;
    xi[i] = p;
}

// timestep.cpp:48-57: x = xi, v = (x - xn)/dt, xn = x.
__global__ void k_commit(const float4* __restrict__ xi, float4* __restrict__ xn, float4* __restrict__ v,
                         uint32_t n, float dt)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 const p = xi[i], q = xn[i];
    v[i] = make_float4((p.x - q.x) / dt, (p.y - q.y) / dt, (p.z - q.z) / dt, 0.0f);
    xn[i] = p;
}

// Commit of one substep fused with the prediction of the next (same per-particle order).
__global__ void k_commit_predict(float4* __restrict__ xi, float4* __restrict__ xn, float4* __restrict__ v,
                                 uint32_t n, float gx, float gy, float gz, int force, float dt)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 const p = xi[i], q = xn[i];
    float4 vel = make_float4((p.x - q.x) / dt, (p.y - q.y) / dt, (p.z - q.z) / dt, 0.0f);
    float4 np;
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

// green_constraint.cpp:61-133. Computes C = |V0| psi and f1..f4 = -grad C, with the reference's
// inversion flip of the smallest singular value (:92-96), the 0.577 clamp (:99-102), the StVK
// first Piola stress of the clamped F^ mapped back with U, V (:104-112), and psi built from
// E = U E^ V^T: |E|_F^2 = sum e_i^2 and tr(E) = sum e_i (u_i . v_i) (:108-110).
__device__ __forceinline__ float green_core(v3 p0, v3 p1, v3 p2, v3 p3, tet_rows const& R, float V0, float mu,
                                            float la, v3 f[4])
{
    v3 const e0 = p0 - p3, e1 = p1 - p3, e2 = p2 - p3;
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

__global__ void __launch_bounds__(256) k_green(tet_group_dev g, float4* xi, const float4* __restrict__ xn,
                                               uint32_t begin, uint32_t count, float alpha_tilde, float gamma,
                                               int first)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t const t = begin + i;
    int4 const id = g.idx[t];
    float4 const a = g.m0[t], b = g.m1[t], c = g.m2[t];
    float const lm = first ? 0.0f : g.lambda[t];
    float4 P[4] = {xi[id.x], xi[id.y], xi[id.z], xi[id.w]};
    tet_rows const R = load_rows(a, b, c);
    v3 f[4];
    float const C = green_core(xyz(P[0]), xyz(P[1]), xyz(P[2]), xyz(P[3]), R, c.y, c.z, c.w, f);
    float const wsum =
        P[0].w * dot(f[0], f[0]) + P[1].w * dot(f[1], f[1]) + P[2].w * dot(f[2], f[2]) + P[3].w * dot(f[3], f[3]);
    if (wsum < 1e-20f) {  // green_constraint.cpp:130-131
        if (first) g.lambda[t] = 0.0f;
        return;
    }
    float gdd = 0.0f;
    if (gamma != 0.0f) {  // damping term (:140-144)
        int const ids[4] = {id.x, id.y, id.z, id.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) gdd += dot(f[k], xyz(P[k]) - xyz(xn[ids[k]]));
    }
    float const dl = hw_div(-(C + alpha_tilde * lm) + gamma * gdd, (1.0f + gamma) * wsum + alpha_tilde);
    g.lambda[t] = lm + dl;
    // xi += w (-f) dl (:154-157)
    xi[id.x] = pack(xyz(P[0]) - (P[0].w * dl) * f[0], P[0].w);
    xi[id.y] = pack(xyz(P[1]) - (P[1].w * dl) * f[1], P[1].w);
    xi[id.z] = pack(xyz(P[2]) - (P[2].w * dl) * f[2], P[2].w);
    xi[id.w] = pack(xyz(P[3]) - (P[3].w * dl) * f[3], P[3].w);
}

// ---------------------------------------------------------------- stable neo-Hookean constraint

// One XPBD update of a tet constraint with gradients g_k = G r_k (k < 3), g_3 = -sum.
__device__ __forceinline__ void tet_update(float4 P[4], const float4* __restrict__ xn, int4 id, m3 const& G,
                                           tet_rows const& R, float C, float alpha_tilde, float gamma, float& lm)
{
    v3 gk[4];
    gk[0] = mul(G, R.r0);
    gk[1] = mul(G, R.r1);
    gk[2] = mul(G, R.r2);
    gk[3] = -(gk[0] + gk[1] + gk[2]);
    float const wsum = P[0].w * dot(gk[0], gk[0]) + P[1].w * dot(gk[1], gk[1]) + P[2].w * dot(gk[2], gk[2]) +
                       P[3].w * dot(gk[3], gk[3]);
    if (wsum < 1e-20f) return;
    float gdd = 0.0f;
    if (gamma != 0.0f) {
        int const ids[4] = {id.x, id.y, id.z, id.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) gdd += dot(gk[k], xyz(P[k]) - xyz(xn[ids[k]]));
    }
    float const dl = hw_div(-(C + alpha_tilde * lm) + gamma * gdd, (1.0f + gamma) * wsum + alpha_tilde);
    lm += dl;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v3 const p = xyz(P[k]) + (P[k].w * dl) * gk[k];
        P[k] = pack(p, P[k].w);
    }
}

__device__ __forceinline__ m3 cofactor(m3 const& F)
{
    return {cross(F.c1, F.c2), cross(F.c2, F.c0), cross(F.c0, F.c1)};
}

__global__ void __launch_bounds__(256) k_neohookean(tet_group_dev g, float4* xi, const float4* __restrict__ xn,
                                                    uint32_t begin, uint32_t count, float inv_dt2, float beta_dt,
                                                    int first)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t const t = begin + i;
    int4 const id = g.idx[t];
    float4 const a = g.m0[t], b = g.m1[t], c = g.m2[t];
    float2 lm = first ? make_float2(0.0f, 0.0f) : reinterpret_cast<const float2*>(g.lambda)[t];
    float4 P[4] = {xi[id.x], xi[id.y], xi[id.z], xi[id.w]};
    tet_rows const R = load_rows(a, b, c);
    {  // deviatoric: C = ||F||_F, dC/dF = F / C, compliance 1/(mu |V0|)
        m3 const F = deformation_gradient(xyz(P[0]) - xyz(P[3]), xyz(P[1]) - xyz(P[3]), xyz(P[2]) - xyz(P[3]),
                                          R.r0, R.r1, R.r2);
        float const C = hw_sqrt(dot(F.c0, F.c0) + dot(F.c1, F.c1) + dot(F.c2, F.c2));
        float const inv = C > 0.0f ? hw_rcp(C) : 0.0f;
        m3 const G = {inv * F.c0, inv * F.c1, inv * F.c2};
        float const at = c.y * inv_dt2;
        tet_update(P, xn, id, G, R, C, at, at * beta_dt, lm.x);
    }
    {  // hydrostatic: C = det F - (1 + mu/lambda), dC/dF = cof F, compliance 1/(lambda |V0|)
        m3 const F = deformation_gradient(xyz(P[0]) - xyz(P[3]), xyz(P[1]) - xyz(P[3]), xyz(P[2]) - xyz(P[3]),
                                          R.r0, R.r1, R.r2);
        m3 const G = cofactor(F);
        float const C = dot(F.c0, G.c0) - c.w;
        float const at = c.z * inv_dt2;
        tet_update(P, xn, id, G, R, C, at, at * beta_dt, lm.y);
    }
    reinterpret_cast<float2*>(g.lambda)[t] = lm;
    xi[id.x] = P[0];
    xi[id.y] = P[1];
    xi[id.z] = P[2];
    xi[id.w] = P[3];
}

// ---------------------------------------------------------------- clustered Gauss-Seidel

// One thread per cluster: the cluster's particles are staged in LDS (slot-major, so that the
// threads of a wave hit consecutive 16-byte words), its constraints run in order against the
// staged copies, and the particles are written back once. Clusters of one color share no
// particle, so this equals the sequential sweep over the color's clusters.
template <int BLOCK>
struct cluster_stage {
    float4 p[kClusterSlots][BLOCK];
};

template <int BLOCK, class Project>
__device__ __forceinline__ void run_cluster(tet_group_dev const& g, cluster_dev const& cl, float4* xi,
                                            uint32_t begin, uint32_t count, Project&& project)
{
    __shared__ cluster_stage<BLOCK> st;
    uint32_t const i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= count) return;
    uint32_t const c = begin + i;
    int4 const va = cl.v0[c], vb = cl.v1[c];
    int const ids[kClusterSlots] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
    float4 (*P)[BLOCK] = st.p;
#pragma unroll
    for (int k = 0; k < kClusterSlots; ++k)
        if (ids[k] >= 0) P[k][threadIdx.x] = xi[ids[k]];
    uint32_t const t0 = cl.off[c], t1 = cl.off[c + 1];
    for (uint32_t t = t0; t < t1; ++t) {
        uint32_t const loc = cl.loc[t];
        uint32_t const s0 = loc & 0xffu, s1 = (loc >> 8) & 0xffu, s2 = (loc >> 16) & 0xffu, s3 = loc >> 24;
        float4 Q[4] = {P[s0][threadIdx.x], P[s1][threadIdx.x], P[s2][threadIdx.x], P[s3][threadIdx.x]};
        project(t, Q, s0, s1, s2, s3, va, vb);
        P[s0][threadIdx.x] = Q[0];
        P[s1][threadIdx.x] = Q[1];
        P[s2][threadIdx.x] = Q[2];
        P[s3][threadIdx.x] = Q[3];
    }
#pragma unroll
    for (int k = 0; k < kClusterSlots; ++k)
        if (ids[k] >= 0) xi[ids[k]] = P[k][threadIdx.x];
}

__device__ __forceinline__ int slot_id(int4 va, int4 vb, uint32_t s)
{
    int const lo = s == 0 ? va.x : (s == 1 ? va.y : (s == 2 ? va.z : va.w));
    int const hi = s == 4 ? vb.x : (s == 5 ? vb.y : (s == 6 ? vb.z : vb.w));
    return s < 4 ? lo : hi;
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_green_clusters(tet_group_dev g, cluster_dev cl, float4* xi,
                                                          const float4* __restrict__ xn, uint32_t begin,
                                                          uint32_t count, float alpha_tilde, float gamma, int first)
{
    run_cluster<BLOCK>(g, cl, xi, begin, count,
                       [&](uint32_t t, float4 P[4], uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3, int4 va,
                           int4 vb) {
        float4 const a = g.m0[t], b = g.m1[t], c = g.m2[t];
        float const lm = first ? 0.0f : g.lambda[t];
        tet_rows const R = load_rows(a, b, c);
        v3 f[4];
        float const C = green_core(xyz(P[0]), xyz(P[1]), xyz(P[2]), xyz(P[3]), R, c.y, c.z, c.w, f);
        float const wsum = P[0].w * dot(f[0], f[0]) + P[1].w * dot(f[1], f[1]) + P[2].w * dot(f[2], f[2]) +
                           P[3].w * dot(f[3], f[3]);
        if (wsum < 1e-20f) {
            if (first) g.lambda[t] = 0.0f;
            return;
        }
        float gdd = 0.0f;
        if (gamma != 0.0f) {
            uint32_t const sl[4] = {s0, s1, s2, s3};
#pragma unroll
            for (int k = 0; k < 4; ++k) gdd += dot(f[k], xyz(P[k]) - xyz(xn[slot_id(va, vb, sl[k])]));
        }
        float const dl = hw_div(-(C + alpha_tilde * lm) + gamma * gdd, (1.0f + gamma) * wsum + alpha_tilde);
        g.lambda[t] = lm + dl;
#pragma unroll
        for (int k = 0; k < 4; ++k) P[k] = pack(xyz(P[k]) - (P[k].w * dl) * f[k], P[k].w);
    });
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_neohookean_clusters(tet_group_dev g, cluster_dev cl, float4* xi,
                                                               const float4* __restrict__ xn, uint32_t begin,
                                                               uint32_t count, float inv_dt2, float beta_dt,
                                                               int first)
{
    run_cluster<BLOCK>(g, cl, xi, begin, count,
                       [&](uint32_t t, float4 P[4], uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3, int4 va,
                           int4 vb) {
        float4 const a = g.m0[t], b = g.m1[t], c = g.m2[t];
        float2 lm = first ? make_float2(0.0f, 0.0f) : reinterpret_cast<const float2*>(g.lambda)[t];
        tet_rows const R = load_rows(a, b, c);
        int4 const id = make_int4(slot_id(va, vb, s0), slot_id(va, vb, s1), slot_id(va, vb, s2), slot_id(va, vb, s3));
        {
            m3 const F = deformation_gradient(xyz(P[0]) - xyz(P[3]), xyz(P[1]) - xyz(P[3]), xyz(P[2]) - xyz(P[3]),
                                              R.r0, R.r1, R.r2);
            float const C = hw_sqrt(dot(F.c0, F.c0) + dot(F.c1, F.c1) + dot(F.c2, F.c2));
            float const inv = C > 0.0f ? hw_rcp(C) : 0.0f;
            m3 const G = {inv * F.c0, inv * F.c1, inv * F.c2};
            float const at = c.y * inv_dt2;
            tet_update(P, xn, id, G, R, C, at, at * beta_dt, lm.x);
        }
        {
            m3 const F = deformation_gradient(xyz(P[0]) - xyz(P[3]), xyz(P[1]) - xyz(P[3]), xyz(P[2]) - xyz(P[3]),
                                              R.r0, R.r1, R.r2);
            m3 const G = cofactor(F);
            float const C = dot(F.c0, G.c0) - c.w;
            float const at = c.z * inv_dt2;
            tet_update(P, xn, id, G, R, C, at, at * beta_dt, lm.y);
        }
        reinterpret_cast<float2*>(g.lambda)[t] = lm;
    });
}

// ---------------------------------------------------------------- hex-cell clusters

// Slot pattern shared by every cluster made of the 6 tets of one hex cell of the box split (slots
// numbered by first appearance). With the pattern known at compile time the 8 particles live in
// registers and every constraint's material and multipliers are loaded before the first one runs.
constexpr int kHexTets = 6;
constexpr int kHexSlots[kHexTets][4] = {{0, 1, 2, 3}, {0, 4, 1, 3}, {0, 2, 5, 3}, {0, 5, 6, 3}, {0, 7, 4, 3}, {0, 6, 7, 3}};

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_neohookean_hex(tet_group_dev g, cluster_dev cl, float4* xi,
                                                          const float4* __restrict__ xn, uint32_t begin,
                                                          uint32_t count, uint32_t tbase, float inv_dt2, float beta_dt,
                                                          int first)
{
    uint32_t const i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= count) return;
    uint32_t const c = begin + i;
    int4 const va = cl.v0[c], vb = cl.v1[c];
    int const ids[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
    float4 A[kHexTets], B[kHexTets], Cm[kHexTets];
    float2 L[kHexTets];
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) {
        uint32_t const t = tbase + uint32_t(k) * count + i;
        A[k] = g.m0[t];
        B[k] = g.m1[t];
        Cm[k] = g.m2[t];
        L[k] = first ? make_float2(0.0f, 0.0f) : reinterpret_cast<const float2*>(g.lambda)[t];
    }
    float4 P[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) P[k] = xi[ids[k]];
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) {
        int const s0 = kHexSlots[k][0], s1 = kHexSlots[k][1], s2 = kHexSlots[k][2], s3 = kHexSlots[k][3];
        float4 Q[4] = {P[s0], P[s1], P[s2], P[s3]};
        int4 const id = make_int4(ids[s0], ids[s1], ids[s2], ids[s3]);
        tet_rows const R = load_rows(A[k], B[k], Cm[k]);
        {
            m3 const F = deformation_gradient(xyz(Q[0]) - xyz(Q[3]), xyz(Q[1]) - xyz(Q[3]), xyz(Q[2]) - xyz(Q[3]),
                                              R.r0, R.r1, R.r2);
            float const C = hw_sqrt(dot(F.c0, F.c0) + dot(F.c1, F.c1) + dot(F.c2, F.c2));
            float const inv = C > 0.0f ? hw_rcp(C) : 0.0f;
            m3 const G = {inv * F.c0, inv * F.c1, inv * F.c2};
            float const at = Cm[k].y * inv_dt2;
            tet_update(Q, xn, id, G, R, C, at, at * beta_dt, L[k].x);
        }
        {
            m3 const F = deformation_gradient(xyz(Q[0]) - xyz(Q[3]), xyz(Q[1]) - xyz(Q[3]), xyz(Q[2]) - xyz(Q[3]),
                                              R.r0, R.r1, R.r2);
            m3 const G = cofactor(F);
            float const C = dot(F.c0, G.c0) - Cm[k].w;
            float const at = Cm[k].z * inv_dt2;
            tet_update(Q, xn, id, G, R, C, at, at * beta_dt, L[k].y);
        }
        P[s0] = Q[0];
        P[s1] = Q[1];
        P[s2] = Q[2];
        P[s3] = Q[3];
    }
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) reinterpret_cast<float2*>(g.lambda)[tbase + uint32_t(k) * count + i] = L[k];
#pragma unroll
    for (int k = 0; k < 8; ++k) xi[ids[k]] = P[k];
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_green_hex(tet_group_dev g, cluster_dev cl, float4* xi,
                                                     const float4* __restrict__ xn, uint32_t begin, uint32_t count,
                                                     uint32_t tbase, float alpha_tilde, float gamma, int first)
{
    uint32_t const i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= count) return;
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
        L[k] = first ? 0.0f : g.lambda[t];
    }
    float4 P[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) P[k] = xi[ids[k]];
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) {
        int const s[4] = {kHexSlots[k][0], kHexSlots[k][1], kHexSlots[k][2], kHexSlots[k][3]};
        v3 f[4];
        float const C = green_core(xyz(P[s[0]]), xyz(P[s[1]]), xyz(P[s[2]]), xyz(P[s[3]]), load_rows(A[k], B[k], Cm[k]),
                                   Cm[k].y, Cm[k].z, Cm[k].w, f);
        float wsum = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q) wsum += P[s[q]].w * dot(f[q], f[q]);
        if (wsum < 1e-20f) continue;  // lambda stays as loaded (0 on the first iteration)
        float gdd = 0.0f;
        if (gamma != 0.0f) {
#pragma unroll
            for (int q = 0; q < 4; ++q) gdd += dot(f[q], xyz(P[s[q]]) - xyz(xn[ids[s[q]]]));
        }
        float const dl = hw_div(-(C + alpha_tilde * L[k]) + gamma * gdd, (1.0f + gamma) * wsum + alpha_tilde);
        L[k] += dl;
#pragma unroll
        for (int q = 0; q < 4; ++q) P[s[q]] = pack(xyz(P[s[q]]) - (P[s[q]].w * dl) * f[q], P[s[q]].w);
    }
#pragma unroll
    for (int k = 0; k < kHexTets; ++k) g.lambda[tbase + uint32_t(k) * count + i] = L[k];
#pragma unroll
    for (int k = 0; k < 8; ++k) xi[ids[k]] = P[k];
}

// ---------------------------------------------------------------- distance constraint

// distance_constraint.cpp:24-53 (no guard against zero length, as there).
__global__ void __launch_bounds__(256) k_distance(distance_group_dev g, float4* xi, const float4* __restrict__ xn,
                                                  uint32_t begin, uint32_t count, float alpha_tilde, float gamma,
                                                  int first)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t const t = begin + i;
    uint2 const id = g.idx[t];
    float const d = g.rest[t];
    float const lm = first ? 0.0f : g.lambda[t];
    float4 const P1 = xi[id.x], P2 = xi[id.y];
    v3 const diff = xyz(P1) - xyz(P2);
    float const length = sqrtf(dot(diff, diff));
    v3 const n = (1.0f / length) * diff;
    float const C = length - d;
    float gdd = 0.0f;
    if (gamma != 0.0f) gdd = dot(n, xyz(P1) - xyz(xn[id.x])) - dot(n, xyz(P2) - xyz(xn[id.y]));
    float const dl = (-(C + alpha_tilde * lm) + gamma * gdd) / ((1.0f + gamma) * (P1.w + P2.w) + alpha_tilde);
    g.lambda[t] = lm + dl;
    xi[id.x] = pack(xyz(P1) + (P1.w * dl) * n, P1.w);
    xi[id.y] = pack(xyz(P2) - (P2.w * dl) * n, P2.w);
}

// ---------------------------------------------------------------- contacts

__device__ __forceinline__ void sdf_eval(sdf_dev const& f, v3 p, float& sd, v3& grad)
{
    if (f.kind == 0) {
        grad = mk(f.a[0], f.a[1], f.a[2]);
        sd = dot(grad, p) + f.a[3];
    } else {
        v3 const d = p - mk(f.a[0], f.a[1], f.a[2]);
        float const len = sqrtf(dot(d, d));
        sd = len - f.a[3];
        grad = len > 0.0f ? (1.0f / len) * d : mk(0.0f, 1.0f, 0.0f);
    }
}

// One thread per surface vertex: contacts against every SDF, in SDF order, from the positions at
// the start of the frame (bvh_model.cpp:72-95, contact_handler.cpp:39-52).
__global__ void k_detect(contact_dev c, const float4* __restrict__ x, const sdf_dev* __restrict__ sdfs,
                         uint32_t nsdf)
{
    uint32_t const s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c.S) return;
    v3 const p = xyz(x[c.surf[s]]);
    uint32_t cnt = 0;
    for (uint32_t k = 0; k < nsdf; ++k) {
        sdf_dev const f = sdfs[k];
        float sd;
        v3 grad;
        sdf_eval(f, p, sd, grad);
        if (!(sd < 0.0f)) continue;
        if (!(p.x >= f.vmin[0] && p.x <= f.vmax[0] && p.y >= f.vmin[1] && p.y <= f.vmax[1] && p.z >= f.vmin[2] &&
              p.z <= f.vmax[2]))
            continue;
        float const gl = sqrtf(dot(grad, grad));
        v3 const nn = (1.0f / gl) * grad;
        v3 const q = p + fabsf(sd) * nn;
        uint32_t const slot = cnt * c.S + s;
        c.q[slot] = make_float4(q.x, q.y, q.z, float(k));
        c.n[slot] = make_float4(nn.x, nn.y, nn.z, 0.0f);
        ++cnt;
    }
    c.count[s] = cnt;
}

// collision_constraint.cpp:21-41, all contacts of one particle in SDF order by one thread.
__global__ void k_contact(contact_dev c, float4* xi, float alpha_tilde, int first)
{
    uint32_t const s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c.S) return;
    uint32_t const cnt = c.count[s];
    if (cnt == 0) return;
    uint32_t const vi = c.surf[s];
    float4 P = xi[vi];
    v3 p = xyz(P);
    float const w = P.w;
    for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t const slot = k * c.S + s;
        v3 const q = xyz(c.q[slot]);
        v3 const n = xyz(c.n[slot]);
        float const lm = first ? 0.0f : c.lambda[slot];
        float const C = dot(p - q, n);
        if (C >= 0.0f) {
            if (first) c.lambda[slot] = 0.0f;
            continue;
        }
        float const dl = -(C + alpha_tilde * lm) / (w + alpha_tilde);
        c.lambda[slot] = lm + dl;
        p = p + (w * dl) * n;
    }
    xi[vi] = pack(p, w);
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

__global__ void __launch_bounds__(kRedBlock) k_green_residual(tet_group_dev g, const float4* __restrict__ x,
                                                              uint32_t begin, uint32_t count, float alpha_tilde,
                                                              double* partial)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    double h2 = 0.0;
    if (i < count) {
        uint32_t const t = begin + i;
        int4 const id = g.idx[t];
        float4 const c = g.m2[t];
        v3 f[4];
        float const C = green_core(xyz(x[id.x]), xyz(x[id.y]), xyz(x[id.z]), xyz(x[id.w]),
                                   load_rows(g.m0[t], g.m1[t], c), c.y, c.z, c.w, f);
        double const h = double(C) + double(alpha_tilde) * double(g.lambda[t]);
        h2 = h * h;
    }
    double const s = block_sum<kRedBlock>(h2);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRedBlock) k_neohookean_residual(tet_group_dev g, const float4* __restrict__ x,
                                                                   uint32_t begin, uint32_t count, float inv_dt2,
                                                                   double* partial)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    double h2 = 0.0;
    if (i < count) {
        uint32_t const t = begin + i;
        int4 const id = g.idx[t];
        float4 const c = g.m2[t];
        tet_rows const R = load_rows(g.m0[t], g.m1[t], c);
        v3 const p3 = xyz(x[id.w]);
        m3 const F = deformation_gradient(xyz(x[id.x]) - p3, xyz(x[id.y]) - p3, xyz(x[id.z]) - p3, R.r0, R.r1, R.r2);
        float2 const lm = reinterpret_cast<const float2*>(g.lambda)[t];
        double const cd = sqrt(double(dot(F.c0, F.c0)) + double(dot(F.c1, F.c1)) + double(dot(F.c2, F.c2)));
        double const ch = double(dot(F.c0, cross(F.c1, F.c2))) - double(c.w);
        double const hd = cd + double(c.y) * double(inv_dt2) * double(lm.x);
        double const hh = ch + double(c.z) * double(inv_dt2) * double(lm.y);
        h2 = hd * hd + hh * hh;
    }
    double const s = block_sum<kRedBlock>(h2);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRedBlock) k_distance_residual(distance_group_dev g, const float4* __restrict__ x,
                                                                 uint32_t begin, uint32_t count, float alpha_tilde,
                                                                 double* partial)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    double h2 = 0.0;
    if (i < count) {
        uint32_t const t = begin + i;
        uint2 const id = g.idx[t];
        v3 const d = xyz(x[id.x]) - xyz(x[id.y]);
        double const h = (sqrt(double(dot(d, d))) - double(g.rest[t])) + double(alpha_tilde) * double(g.lambda[t]);
        h2 = h * h;
    }
    double const s = block_sum<kRedBlock>(h2);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRedBlock) k_particle_energy(const float4* __restrict__ x,
                                                               const float4* __restrict__ v,
                                                               const float* __restrict__ mass, uint32_t n, float gx,
                                                               float gy, float gz, int force, double* partial)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0, pe = 0, px = 0, py = 0, pz = 0;
    if (i < n) {
        double const m = mass[i];
        float4 const p = x[i], u = v[i];
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

__global__ void __launch_bounds__(kRedBlock) k_green_energy(tet_group_dev g, const float4* __restrict__ x,
                                                            uint32_t count, double* partial)
{
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (t < count) {
        int4 const id = g.idx[t];
        float4 const c = g.m2[t];
        v3 f[4];
        e = green_core(xyz(x[id.x]), xyz(x[id.y]), xyz(x[id.z]), xyz(x[id.w]), load_rows(g.m0[t], g.m1[t], c), c.y,
                       c.z, c.w, f);
    }
    double const s = block_sum<kRedBlock>(e);
    if (threadIdx.x == 0) {
        double* o = partial + 6 * blockIdx.x;
        o[0] = o[1] = o[3] = o[4] = o[5] = 0.0;
        o[2] = s;
    }
}

__global__ void __launch_bounds__(kRedBlock) k_neohookean_energy(tet_group_dev g, const float4* __restrict__ x,
                                                                 uint32_t count, double* partial)
{
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (t < count) {
        int4 const id = g.idx[t];
        float4 const c = g.m2[t];
        tet_rows const R = load_rows(g.m0[t], g.m1[t], c);
        v3 const p3 = xyz(x[id.w]);
        m3 const F = deformation_gradient(xyz(x[id.x]) - p3, xyz(x[id.y]) - p3, xyz(x[id.z]) - p3, R.r0, R.r1, R.r2);
        double const ic = double(dot(F.c0, F.c0)) + double(dot(F.c1, F.c1)) + double(dot(F.c2, F.c2));
        double const j = double(dot(F.c0, cross(F.c1, F.c2))) - double(c.w);
        // mu |V0| = 1 / m2.y, lambda |V0| = 1 / m2.z
        e = 0.5 / double(c.y) * (ic - 3.0) + 0.5 / double(c.z) * j * j;
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

// Priority of constraint t: (previous color, hash32(t)) -- higher first; ties are impossible
// because hash32 is a bijection on 32-bit indices.
__global__ void k_jp_keys(uint32_t n, const uint32_t* __restrict__ prev, uint64_t* __restrict__ key)
{
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    key[t] = (uint64_t(prev ? prev[t] : 0u) << 32) | hash32(t);
}

// A constraint whose key exceeds the keys of all its uncolored neighbours takes the smallest color
// unused by its colored neighbours. Over all rounds this equals sequential greedy coloring in
// descending key order: when a constraint is colored, exactly its higher-key neighbours are.
__global__ void k_jp_round(uint32_t n, uint32_t arity, const uint32_t* __restrict__ elem_off,
                           const uint32_t* __restrict__ verts,
                           const uint32_t* __restrict__ inc_off, const uint32_t* __restrict__ inc,
                           const uint64_t* __restrict__ key, const uint32_t* __restrict__ color_in,
                           uint32_t* __restrict__ color_out, uint32_t* remaining, uint32_t* overflow)
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

void launch_predict(hipStream_t s, float4* xi, const float4* xn, float4* v, uint32_t n, float gx, float gy,
                    float gz, int force, float dt, int block)
{
    if (n) k_predict<<<grid_for(n, block), block, 0, s>>>(xi, xn, v, n, gx, gy, gz, force, dt);
}

void launch_commit(hipStream_t s, const float4* xi, float4* xn, float4* v, uint32_t n, float dt, int block)
{
    if (n) k_commit<<<grid_for(n, block), block, 0, s>>>(xi, xn, v, n, dt);
}

void launch_commit_predict(hipStream_t s, float4* xi, float4* xn, float4* v, uint32_t n, float gx, float gy,
                           float gz, int force, float dt, int block)
{
    if (n) k_commit_predict<<<grid_for(n, block), block, 0, s>>>(xi, xn, v, n, gx, gy, gz, force, dt);
}

void launch_green(hipStream_t s, tet_group_dev g, float4* xi, const float4* xn, uint32_t begin, uint32_t count,
                  float alpha_tilde, float gamma, int first, int block)
{
    if (count) k_green<<<grid_for(count, block), block, 0, s>>>(g, xi, xn, begin, count, alpha_tilde, gamma, first);
}

void launch_neohookean(hipStream_t s, tet_group_dev g, float4* xi, const float4* xn, uint32_t begin,
                       uint32_t count, float inv_dt2, float beta_dt, int first, int block)
{
    if (count)
        k_neohookean<<<grid_for(count, block), block, 0, s>>>(g, xi, xn, begin, count, inv_dt2, beta_dt, first);
}

void launch_distance(hipStream_t s, distance_group_dev g, float4* xi, const float4* xn, uint32_t begin,
                     uint32_t count, float alpha_tilde, float gamma, int first, int block)
{
    if (count)
        k_distance<<<grid_for(count, block), block, 0, s>>>(g, xi, xn, begin, count, alpha_tilde, gamma, first);
}

void launch_detect(hipStream_t s, contact_dev c, const float4* x, const sdf_dev* sdfs, uint32_t nsdf, int block)
{
    if (c.S) k_detect<<<grid_for(c.S, block), block, 0, s>>>(c, x, sdfs, nsdf);
}

void launch_contact(hipStream_t s, contact_dev c, float4* xi, float alpha_tilde, int first, int block)
{
    if (c.S) k_contact<<<grid_for(c.S, block), block, 0, s>>>(c, xi, alpha_tilde, first);
}

void launch_green_residual(hipStream_t s, tet_group_dev g, const float4* x, uint32_t begin, uint32_t count,
                           float alpha_tilde, double* partial, int)
{
    if (count) k_green_residual<<<grid_for(count, kRedBlock), kRedBlock, 0, s>>>(g, x, begin, count, alpha_tilde, partial);
}

void launch_neohookean_residual(hipStream_t s, tet_group_dev g, const float4* x, uint32_t begin, uint32_t count,
                                float inv_dt2, double* partial, int)
{
    if (count)
        k_neohookean_residual<<<grid_for(count, kRedBlock), kRedBlock, 0, s>>>(g, x, begin, count, inv_dt2, partial);
}

void launch_distance_residual(hipStream_t s, distance_group_dev g, const float4* x, uint32_t begin,
                              uint32_t count, float alpha_tilde, double* partial, int)
{
    if (count)
        k_distance_residual<<<grid_for(count, kRedBlock), kRedBlock, 0, s>>>(g, x, begin, count, alpha_tilde, partial);
}

void launch_particle_energy(hipStream_t s, const float4* x, const float4* v, const float* mass, uint32_t n, float gx,
                            float gy, float gz, int force, double* partial, int)
{
    if (n) k_particle_energy<<<grid_for(n, kRedBlock), kRedBlock, 0, s>>>(x, v, mass, n, gx, gy, gz, force, partial);
}

void launch_green_energy(hipStream_t s, tet_group_dev g, const float4* x, uint32_t count, double* partial, int)
{
    if (count) k_green_energy<<<grid_for(count, kRedBlock), kRedBlock, 0, s>>>(g, x, count, partial);
}

void launch_neohookean_energy(hipStream_t s, tet_group_dev g, const float4* x, uint32_t count, double* partial, int)
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

void launch_green_clusters(hipStream_t s, tet_group_dev g, cluster_dev cl, float4* xi, const float4* xn,
                           uint32_t begin, uint32_t count, float alpha_tilde, float gamma, int first, int block)
{
    if (!count) return;
    if (block <= 64)
        k_green_clusters<64><<<grid_for(count, 64), 64, 0, s>>>(g, cl, xi, xn, begin, count, alpha_tilde, gamma, first);
    else
        k_green_clusters<128><<<grid_for(count, 128), 128, 0, s>>>(g, cl, xi, xn, begin, count, alpha_tilde, gamma, first);
}

void launch_neohookean_hex(hipStream_t s, tet_group_dev g, cluster_dev cl, float4* xi, const float4* xn,
                           uint32_t begin, uint32_t count, uint32_t tbase, float inv_dt2, float beta_dt, int first,
                           int block)
{
    if (!count) return;
    if (block <= 64)
        k_neohookean_hex<64><<<grid_for(count, 64), 64, 0, s>>>(g, cl, xi, xn, begin, count, tbase, inv_dt2, beta_dt,
                                                                first);
    else
        k_neohookean_hex<128><<<grid_for(count, 128), 128, 0, s>>>(g, cl, xi, xn, begin, count, tbase, inv_dt2,
                                                                   beta_dt, first);
}

void launch_green_hex(hipStream_t s, tet_group_dev g, cluster_dev cl, float4* xi, const float4* xn, uint32_t begin,
                      uint32_t count, uint32_t tbase, float alpha_tilde, float gamma, int first, int block)
{
    if (!count) return;
    if (block <= 64)
        k_green_hex<64><<<grid_for(count, 64), 64, 0, s>>>(g, cl, xi, xn, begin, count, tbase, alpha_tilde, gamma,
                                                           first);
    else
        k_green_hex<128><<<grid_for(count, 128), 128, 0, s>>>(g, cl, xi, xn, begin, count, tbase, alpha_tilde, gamma,
                                                              first);
}

void launch_neohookean_clusters(hipStream_t s, tet_group_dev g, cluster_dev cl, float4* xi, const float4* xn,
                                uint32_t begin, uint32_t count, float inv_dt2, float beta_dt, int first, int block)
{
    if (!count) return;
    if (block <= 64)
        k_neohookean_clusters<64><<<grid_for(count, 64), 64, 0, s>>>(g, cl, xi, xn, begin, count, inv_dt2, beta_dt, first);
    else
        k_neohookean_clusters<128><<<grid_for(count, 128), 128, 0, s>>>(g, cl, xi, xn, begin, count, inv_dt2, beta_dt,
                                                                       first);
}

void launch_jp_keys(hipStream_t s, uint32_t n, const uint32_t* prev, uint64_t* key, int block)
{
    if (n) k_jp_keys<<<grid_for(n, block), block, 0, s>>>(n, prev, key);
}

void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t value, uint32_t n)
{
    if (n) k_fill_u32<<<grid_for(n, 256), 256, 0, s>>>(p, value, n);
}

}  // namespace sbs
