// FP64 3x3 linear algebra for the oracle.
//
// TEST INFRASTRUCTURE ONLY: the oracle is the CPU checker of the GPU path. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// The reference uses Eigen 3 for these operations (green_constraint.cpp:43-44 inverse/determinant,
// :81 JacobiSVD; sdf_model.cpp:56-61 Hyperplane). Eigen is not part of the reference tree, so this
// file restates the published algorithms:
//   * determinant / inverse of a 3x3 by cofactor expansion (what Eigen does for fixed 3x3),
//   * a two-sided Jacobi SVD (Forsythe-Henrici / Brent-Luk 2x2 real SVD sweeps, the algorithm
//     class of Eigen::JacobiSVD): singular values made non-negative and sorted descending.
// Outputs the reference consumes (U diag(s) V^T products, singular values) are independent of the
// sign convention of U/V columns except in the degenerate case of repeated singular values.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>

namespace oracle {

struct vec3 {
    double x = 0, y = 0, z = 0;
    vec3() = default;
    vec3(double a, double b, double c) : x(a), y(b), z(c) {}
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(double s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3 operator*(vec3 a, double s) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3 operator/(vec3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec3& operator+=(vec3& a, vec3 b) { a = a + b; return a; }
inline vec3& operator-=(vec3& a, vec3 b) { a = a - b; return a; }
inline double dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm2(vec3 a) { return dot(a, a); }
inline double norm(vec3 a) { return std::sqrt(norm2(a)); }
inline vec3 cross(vec3 a, vec3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Row-major 3x3: m[r][c].
struct mat3 {
    double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    static mat3 identity()
    {
        mat3 r;
        r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0;
        return r;
    }
    static mat3 from_cols(vec3 a, vec3 b, vec3 c)
    {
        mat3 r;
        for (int i = 0; i < 3; ++i) {
            r.m[i][0] = a[i];
            r.m[i][1] = b[i];
            r.m[i][2] = c[i];
        }
        return r;
    }
    vec3 col(int j) const { return {m[0][j], m[1][j], m[2][j]}; }
    vec3 row(int i) const { return {m[i][0], m[i][1], m[i][2]}; }
    void set_col(int j, vec3 v)
    {
        m[0][j] = v.x;
        m[1][j] = v.y;
        m[2][j] = v.z;
    }
};

inline mat3 operator*(mat3 const& a, mat3 const& b)
{
    mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}
inline vec3 operator*(mat3 const& a, vec3 v)
{
    return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
            a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
            a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
inline mat3 operator*(double s, mat3 a)
{
    for (auto& r : a.m)
        for (double& x : r) x *= s;
    return a;
}
inline mat3 operator+(mat3 a, mat3 const& b)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a.m[i][j] += b.m[i][j];
    return a;
}
inline mat3 operator-(mat3 a, mat3 const& b)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a.m[i][j] -= b.m[i][j];
    return a;
}
inline mat3 transpose(mat3 const& a)
{
    mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
inline double trace(mat3 const& a) { return a.m[0][0] + a.m[1][1] + a.m[2][2]; }
inline double frob2(mat3 const& a)
{
    double s = 0;
    for (auto& r : a.m)
        for (double x : r) s += x * x;
    return s;
}

// Cofactor expansion along the first row.
inline double determinant(mat3 const& a)
{
    auto const& m = a.m;
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
           m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

// Adjugate / determinant.
inline mat3 inverse(mat3 const& a)
{
    auto const& m = a.m;
    mat3 c;  // cofactor matrix transposed (adjugate)
    c.m[0][0] = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    c.m[0][1] = m[0][2] * m[2][1] - m[0][1] * m[2][2];
    c.m[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
    c.m[1][0] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    c.m[1][1] = m[0][0] * m[2][2] - m[0][2] * m[2][0];
    c.m[1][2] = m[0][2] * m[1][0] - m[0][0] * m[1][2];
    c.m[2][0] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    c.m[2][1] = m[0][1] * m[2][0] - m[0][0] * m[2][1];
    c.m[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
    double const det = m[0][0] * c.m[0][0] + m[0][1] * c.m[1][0] + m[0][2] * c.m[2][0];
    return (1.0 / det) * c;
}

// Two-sided Jacobi SVD: A = U diag(s) V^T, s[0] >= s[1] >= s[2] >= 0, U and V orthogonal.
// Each step applies a 2x2 real SVD to rows/cols (p,q): a left rotation that symmetrizes the
// 2x2 block followed by the symmetric Jacobi rotation that diagonalizes it. Sweeps repeat until
// every off-diagonal pair is below 2*eps relative to the larger diagonal magnitude.
inline void jacobi_svd3(mat3 const& a, mat3& U, double s[3], mat3& V)
{
    double scale = 0;
    for (auto& r : a.m)
        for (double x : r) scale = std::max(scale, std::fabs(x));
    if (scale == 0) scale = 1;
    mat3 w = (1.0 / scale) * a;
    U = mat3::identity();
    V = mat3::identity();
    double const precision = 2.0 * 2.220446049250313e-16;
    double const tiny = 2.2250738585072014e-308;

    for (int sweep = 0; sweep < 64; ++sweep) {
        bool finished = true;
        for (int p = 1; p < 3; ++p) {
            for (int q = 0; q < p; ++q) {
                double const thr =
                    std::max(tiny, precision * std::max(std::fabs(w.m[p][p]), std::fabs(w.m[q][q])));
                if (std::max(std::fabs(w.m[p][q]), std::fabs(w.m[q][p])) <= thr) continue;
                finished = false;
                // 2x2 block in index order (q, p): [[a00, a01], [a10, a11]]
                double const a00 = w.m[q][q], a01 = w.m[q][p], a10 = w.m[p][q], a11 = w.m[p][p];
                // left rotation G = [[c, s], [-s, c]] with G*M symmetric:
                // tan(theta) = (a10 - a01) / (a00 + a11)
                double const t = a00 + a11, d = a10 - a01;
                double c1 = 1, s1 = 0;
                double const r = std::hypot(t, d);
                if (r > 0) {
                    c1 = t / r;
                    s1 = d / r;
                }
                double const b00 = c1 * a00 + s1 * a10;
                double const b01 = c1 * a01 + s1 * a11;
                double const b11 = -s1 * a01 + c1 * a11;
                // symmetric Jacobi J = [[c, s], [-s, c]] with J^T B J diagonal
                double c2 = 1, s2 = 0;
                if (b01 != 0) {
                    double const tau = (b11 - b00) / (2 * b01);
                    double const tt = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1 + tau * tau));
                    c2 = 1 / std::sqrt(1 + tt * tt);
                    s2 = tt * c2;
                }
                // left total L = J^T G, right total R = J
                double const l00 = c2 * c1 + s2 * s1;
                double const l01 = c2 * s1 - s2 * c1;
                double const l10 = s2 * c1 - c2 * s1;
                double const l11 = s2 * s1 + c2 * c1;
                // apply L on the left to rows q, p of w; U <- U L^T
                for (int j = 0; j < 3; ++j) {
                    double const wq = w.m[q][j], wp = w.m[p][j];
                    w.m[q][j] = l00 * wq + l01 * wp;
                    w.m[p][j] = l10 * wq + l11 * wp;
                }
                for (int i = 0; i < 3; ++i) {
                    double const uq = U.m[i][q], up = U.m[i][p];
                    U.m[i][q] = uq * l00 + up * l01;
                    U.m[i][p] = uq * l10 + up * l11;
                }
                // apply R = [[c2, s2], [-s2, c2]] on the right to cols q, p of w and V
                for (int i = 0; i < 3; ++i) {
                    double const wq = w.m[i][q], wp = w.m[i][p];
                    w.m[i][q] = wq * c2 - wp * s2;
                    w.m[i][p] = wq * s2 + wp * c2;
                    double const vq = V.m[i][q], vp = V.m[i][p];
                    V.m[i][q] = vq * c2 - vp * s2;
                    V.m[i][p] = vq * s2 + vp * c2;
                }
            }
        }
        if (finished) break;
    }
    for (int i = 0; i < 3; ++i) {
        double const d = w.m[i][i];
        s[i] = std::fabs(d) * scale;
        if (d < 0)
            for (int r = 0; r < 3; ++r) U.m[r][i] = -U.m[r][i];
    }
    // sort descending (selection), permuting U and V columns alongside
    for (int i = 0; i < 2; ++i) {
        int k = i;
        for (int j = i + 1; j < 3; ++j)
            if (s[j] > s[k]) k = j;
        if (k != i) {
            std::swap(s[i], s[k]);
            for (int r = 0; r < 3; ++r) {
                std::swap(U.m[r][i], U.m[r][k]);
                std::swap(V.m[r][i], V.m[r][k]);
            }
        }
    }
}

}  // namespace oracle
