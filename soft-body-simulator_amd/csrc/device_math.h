// FP32 vector / 3x3 helpers and the 3x3 SVD used by the tet projection kernels on gfx950.
#pragma once
#include <hip/hip_runtime.h>

namespace sbs {

struct v3 {
    float x, y, z;
};

__device__ __forceinline__ v3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ v3 xyz(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ v3 operator*(float s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Hardware square root / reciprocal / reciprocal square root (v_sqrt_f32, v_rcp_f32, v_rsq_f32,
// about 1 ulp) instead of the multi-instruction correctly rounded sequences; the projection math
// is latency bound, one instruction per 4 cycles per wave.
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hw_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float hw_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Columns of a 3x3 matrix.
struct m3 {
    v3 c0, c1, c2;
};

// M * v with M given by columns.
__device__ __forceinline__ v3 mul(m3 const& M, v3 v)
{
    return {M.c0.x * v.x + M.c1.x * v.y + M.c2.x * v.z, M.c0.y * v.x + M.c1.y * v.y + M.c2.y * v.z,
            M.c0.z * v.x + M.c1.z * v.y + M.c2.z * v.z};
}

// F = Ds * DmInv with Ds = [e0 e1 e2] (columns) and DmInv given by rows r0, r1, r2:
// F = e0 r0^T + e1 r1^T + e2 r2^T, column j of F = e0 r0[j] + e1 r1[j] + e2 r2[j].
__device__ __forceinline__ m3 deformation_gradient(v3 e0, v3 e1, v3 e2, v3 r0, v3 r1, v3 r2)
{
    m3 F;
    F.c0 = r0.x * e0 + r1.x * e1 + r2.x * e2;
    F.c1 = r0.y * e0 + r1.y * e1 + r2.y * e2;
    F.c2 = r0.z * e0 + r1.z * e1 + r2.z * e2;
    return F;
}

// One Jacobi rotation of the symmetric matrix S (entries spp, sqq, spq, and the two entries
// srp, srq coupling the third index r) in the (p, q) plane, accumulated into columns vp, vq.
// Rotation angle from the classic formulation (Golub & Van Loan 8.5.2): theta = (sqq-spp)/(2spq),
// t = sgn(theta)/(|theta| + sqrt(1+theta^2)), c = 1/sqrt(1+t^2), s = t c.
__device__ __forceinline__ void jacobi_rotate(float& spp, float& sqq, float& spq, float& srp, float& srq, v3& vp,
                                              v3& vq)
{
    if (fabsf(spq) < 1e-30f) return;
    float const theta = hw_div(sqq - spp, 2.0f * spq);
    float const t = copysignf(hw_rcp(fabsf(theta) + hw_sqrt(fmaf(theta, theta, 1.0f))), theta);
    float const c = hw_rsq(fmaf(t, t, 1.0f));
    float const s = t * c;
    spp = spp - t * spq;
    sqq = sqq + t * spq;
    spq = 0.0f;
    float const rp = srp, rq = srq;
    srp = c * rp - s * rq;
    srq = s * rp + c * rq;
    v3 const a = vp, b = vq;
    vp = c * a - s * b;
    vq = s * a + c * b;
}

// Signed SVD of a 3x3 F: F = U diag(s0, s1, s2) V^T with U, V rotations, s0 >= s1 >= |s2| and
// s2 carrying the sign of det F. Method: cyclic Jacobi eigen-decomposition of F^T F (fixed
// number of sweeps, eigenvalues sorted descending), B = F V, then Gram-Schmidt of B's columns
// (which are orthogonal up to rounding) gives U and the singular values (the scheme of McAdams,
// Selle, Tamstorf, Teran, Sifakis 2011, "Computing the SVD of 3x3 matrices with minimal
// branching", with exact rather than approximate Givens angles).
__device__ __forceinline__ void svd3(m3 const& F, m3& U, float s[3], m3& V)
{
    float s00 = dot(F.c0, F.c0), s11 = dot(F.c1, F.c1), s22 = dot(F.c2, F.c2);
    float s01 = dot(F.c0, F.c1), s02 = dot(F.c0, F.c2), s12 = dot(F.c1, F.c2);
    v3 v0 = mk(1, 0, 0), v1 = mk(0, 1, 0), v2 = mk(0, 0, 1);
#pragma unroll
    for (int sweep = 0; sweep < 5; ++sweep) {
        jacobi_rotate(s00, s11, s01, s02, s12, v0, v1);  // (p,q,r) = (0,1,2)
        jacobi_rotate(s00, s22, s02, s01, s12, v0, v2);  // (0,2,1): srp = s10, srq = s12
        jacobi_rotate(s11, s22, s12, s01, s02, v1, v2);  // (1,2,0): srp = s01, srq = s02
    }
    // sort eigenvalues descending; every swap negates one column so det(V) stays +1
    float e0 = s00, e1 = s11, e2 = s22;
    if (e0 < e1) {
        float t = e0; e0 = e1; e1 = t;
        v3 tv = v0; v0 = v1; v1 = -tv;
    }
    if (e0 < e2) {
        float t = e0; e0 = e2; e2 = t;
        v3 tv = v0; v0 = v2; v2 = -tv;
    }
    if (e1 < e2) {
        float t = e1; e1 = e2; e2 = t;
        v3 tv = v1; v1 = v2; v2 = -tv;
    }
    V.c0 = v0;
    V.c1 = v1;
    V.c2 = v2;
    v3 const b0 = mul(F, v0), b1 = mul(F, v1), b2 = mul(F, v2);
    float const n0 = hw_sqrt(dot(b0, b0));
    v3 u0;
    if (n0 > 1e-30f) {
        u0 = hw_rcp(n0) * b0;
    } else {
        u0 = mk(1, 0, 0);
    }
    v3 w1 = b1 - dot(u0, b1) * u0;
    float n1 = hw_sqrt(dot(w1, w1));
    v3 u1;
    if (n1 > 1e-30f) {
        u1 = hw_rcp(n1) * w1;
    } else {
        // any unit vector orthogonal to u0
        v3 const a = fabsf(u0.x) < 0.57f ? mk(1, 0, 0) : (fabsf(u0.y) < 0.57f ? mk(0, 1, 0) : mk(0, 0, 1));
        v3 const c = cross(u0, a);
        u1 = hw_rsq(dot(c, c)) * c;
        n1 = 0.0f;
    }
    v3 const u2 = cross(u0, u1);
    U.c0 = u0;
    U.c1 = u1;
    U.c2 = u2;
    s[0] = n0;
    s[1] = dot(u1, b1);
    s[2] = dot(u2, b2);
}

}  // namespace sbs
