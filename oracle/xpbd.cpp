// TEST INFRASTRUCTURE ONLY (see xpbd.h).
#include "xpbd.h"

#include <algorithm>
#include <array>
#include <map>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace oracle {

namespace {

mat3 rest_shape(simulation_t const& sim, uint32_t body, uint32_t const v[4])
{
    auto const& P = sim.particles.at(body);
    vec3 const p4 = P.at(v[3]).x0;
    return mat3::from_cols(P.at(v[0]).x0 - p4, P.at(v[1]).x0 - p4, P.at(v[2]).x0 - p4);
}

void lame(double young, double poisson, double* mu, double* lam)
{
    // green_constraint.cpp:45-46
    *mu = young / (2. * (1 + poisson));
    *lam = (young * poisson) / ((1 + poisson) * (1 - 2 * poisson));
}

// Green-strain constraint core (green_constraint.cpp:61-131): computes C = |V0| psi and the
// forces f1..f4 (= -grad C) at positions p. Returns the weighted sum of squared gradients.
struct green_eval {
    uint8_t branches = 0;  // bit 0: inverted, bits 1-3: singular value i clamped
    double C, wsum;
    vec3 f[4];
};

green_eval green_core(mat3 const& DmInv, double V0, double mu, double lam, vec3 const p[4], double const w[4])
{
    // signed_volume (:160-173)
    mat3 const Dsv = mat3::from_cols(p[0] - p[3], p[1] - p[3], p[2] - p[3]);
    double const Vsigned = (1. / 6.) * determinant(Dsv);
    bool const is_V_positive = Vsigned >= 0.;
    bool const is_V0_positive = V0 >= 0.;
    bool const is_tet_inverted = (is_V_positive && !is_V0_positive) || (!is_V_positive && is_V0_positive);

    mat3 const F = Dsv * DmInv;  // (:69-74)
    mat3 U, Vm;
    double s[3];
    jacobi_svd3(F, U, s, Vm);  // (:81)
    mat3 Fhat;
    Fhat.m[0][0] = s[0];
    Fhat.m[1][1] = s[1];
    Fhat.m[2][2] = s[2];
    if (is_tet_inverted) {  // (:92-96)
        Fhat.m[2][2] = -Fhat.m[2][2];
        U.set_col(2, -U.col(2));
    }
    double const min_singular_value = 0.577;  // (:99-102)
    uint8_t branches = is_tet_inverted ? 1 : 0;
    for (int i = 0; i < 3; ++i) {
        if (Fhat.m[i][i] < min_singular_value) branches |= uint8_t(2 << i);
        Fhat.m[i][i] = std::max(Fhat.m[i][i], min_singular_value);
    }

    mat3 const I = mat3::identity();
    mat3 const Ehat = 0.5 * (transpose(Fhat) * Fhat - I);  // (:104-106)
    double const EhatTrace = trace(Ehat);
    mat3 const Piolahat = Fhat * ((2. * mu * Ehat) + (lam * EhatTrace * I));

    mat3 const E = U * Ehat * transpose(Vm);  // (:108-110)
    double const Etrace = trace(E);
    double const psi = mu * frob2(E) + 0.5 * lam * Etrace * Etrace;

    mat3 const Piola = U * Piolahat * transpose(Vm);  // (:112)
    double const V0abs = std::abs(V0);               // (:115-120)
    mat3 const H = (-V0abs) * Piola * transpose(DmInv);
    green_eval r;
    r.branches = branches;
    r.f[0] = H.col(0);
    r.f[1] = H.col(1);
    r.f[2] = H.col(2);
    r.f[3] = -(r.f[0] + r.f[1] + r.f[2]);
    r.wsum = 0;
    if (w)
        r.wsum = w[0] * norm2(r.f[0]) + w[1] * norm2(r.f[1]) + w[2] * norm2(r.f[2]) + w[3] * norm2(r.f[3]);
    r.C = V0abs * psi;  // (:133)
    return r;
}

// One XPBD multiplier update with the reference's damping term (green_constraint.cpp:133-151,
// distance_constraint.cpp:37-48). grad[k] is grad C at vertex k.
double xpbd_delta(double C, double lagrange, double alpha, double beta, double dt, double wsum,
                  double grad_dot_disp)
{
    double const dt2 = dt * dt;
    double const alpha_tilde = alpha / dt2;
    double const beta_tilde = beta * dt2;
    double const gamma = alpha_tilde * beta_tilde / dt;
    double const num = -(C + alpha_tilde * lagrange) + gamma * grad_dot_disp;
    double const den = (1. + gamma) * wsum + alpha_tilde;
    return num / den;
}

}  // namespace

green_constraint_t::green_constraint_t(double alpha, double beta, simulation_t const& sim, uint32_t body,
                                       uint32_t const vs[4], double young, double poisson)
    : constraint_t(alpha, beta), bi(body)
{
    std::copy(vs, vs + 4, v);
    mat3 const Dm = rest_shape(sim, body, v);  // green_constraint.cpp:38-44
    DmInv = inverse(Dm);
    V0 = (1. / 6.) * determinant(Dm);
    lame(young, poisson, &mu, &lam);
}

double green_constraint_t::evaluate(vec3 const p[4], bool* early_out, double const w[4]) const
{
    green_eval const g = green_core(DmInv, V0, mu, lam, p, w);
    if (early_out) *early_out = g.wsum < 1e-20;
    return g.C;
}

double green_constraint_t::evaluate_forces(vec3 const p[4], vec3 f[4]) const
{
    green_eval const g = green_core(DmInv, V0, mu, lam, p, nullptr);
    for (int k = 0; k < 4; ++k) f[k] = g.f[k];
    return g.C;
}

void green_constraint_t::project_positions(simulation_t& sim, double dt)
{
    auto& P = sim.particles[bi];
    particle_t* pp[4] = {&P[v[0]], &P[v[1]], &P[v[2]], &P[v[3]]};
    vec3 const p[4] = {pp[0]->xi, pp[1]->xi, pp[2]->xi, pp[3]->xi};
    double const w[4] = {pp[0]->invmass(), pp[1]->invmass(), pp[2]->invmass(), pp[3]->invmass()};
    green_eval const g = green_core(DmInv, V0, mu, lam, p, w);
    double constexpr epsilon = 1e-20;  // (:67, :130-131)
    if (sim.green_trace) sim.green_trace->push_back(uint8_t(g.branches | (g.wsum < epsilon ? 16 : 0)));
    if (g.wsum < epsilon) return;
    double gdd = 0;
    for (int k = 0; k < 4; ++k) gdd += dot(g.f[k], pp[k]->xi - pp[k]->xn);  // (:140-144)
    double const dl = xpbd_delta(g.C, lagrange, alpha, beta, dt, g.wsum, gdd);
    lagrange += dl;  // (:152-157)
    for (int k = 0; k < 4; ++k) pp[k]->xi += w[k] * (-g.f[k]) * dl;
}

void green_constraint_t::residual_sq(simulation_t const& sim, double dt, double* acc) const
{
    auto const& P = sim.particles[bi];
    vec3 const p[4] = {P[v[0]].x, P[v[1]].x, P[v[2]].x, P[v[3]].x};
    double const C = evaluate(p);
    double const h = C + (alpha / (dt * dt)) * lagrange;
    *acc += h * h;
}

double green_constraint_t::energy(simulation_t const& sim) const
{
    auto const& P = sim.particles[bi];
    vec3 const p[4] = {P[v[0]].x, P[v[1]].x, P[v[2]].x, P[v[3]].x};
    return evaluate(p);
}

neohookean_constraint_t::neohookean_constraint_t(double beta, simulation_t const& sim, uint32_t body,
                                                 uint32_t const vs[4], double young, double poisson)
    : constraint_t(0.0, beta), bi(body)
{
    std::copy(vs, vs + 4, v);
    mat3 const Dm = rest_shape(sim, body, v);
    DmInv = inverse(Dm);
    V0 = (1. / 6.) * determinant(Dm);
    lame(young, poisson, &mu, &lam);
}

namespace {

// grad_x C for vertices 1..3 = (dC/dF) * (row k of DmInv)^T, vertex 4 = -sum.
void tet_gradients(mat3 const& dCdF, mat3 const& DmInv, vec3 g[4])
{
    for (int k = 0; k < 3; ++k) g[k] = dCdF * DmInv.row(k);
    g[3] = -(g[0] + g[1] + g[2]);
}

mat3 cofactor(mat3 const& F)
{
    vec3 const c0 = F.col(0), c1 = F.col(1), c2 = F.col(2);
    return mat3::from_cols(cross(c1, c2), cross(c2, c0), cross(c0, c1));
}

}  // namespace

void neohookean_constraint_t::project_positions(simulation_t& sim, double dt)
{
    auto& P = sim.particles[bi];
    particle_t* pp[4] = {&P[v[0]], &P[v[1]], &P[v[2]], &P[v[3]]};
    double const w[4] = {pp[0]->invmass(), pp[1]->invmass(), pp[2]->invmass(), pp[3]->invmass()};
    double const V0abs = std::abs(V0);
    double constexpr epsilon = 1e-20;
    for (int part = 0; part < 2; ++part) {
        mat3 const Ds = mat3::from_cols(pp[0]->xi - pp[3]->xi, pp[1]->xi - pp[3]->xi, pp[2]->xi - pp[3]->xi);
        mat3 const F = Ds * DmInv;
        double C, compliance;
        mat3 dCdF;
        if (part == 0) {
            double const r = std::sqrt(frob2(F));
            C = r;
            compliance = 1.0 / (mu * V0abs);
            dCdF = (r > 0 ? 1.0 / r : 0.0) * F;
        } else {
            C = determinant(F) - (1.0 + mu / lam);
            compliance = 1.0 / (lam * V0abs);
            dCdF = cofactor(F);
        }
        vec3 g[4];
        tet_gradients(dCdF, DmInv, g);
        double wsum = 0, gdd = 0;
        for (int k = 0; k < 4; ++k) {
            wsum += w[k] * norm2(g[k]);
            gdd += dot(g[k], pp[k]->xi - pp[k]->xn);
        }
        if (wsum < epsilon) continue;
        double& lm = (part == 0) ? lagrange : lagrange_h;
        double const dl = xpbd_delta(C, lm, compliance, beta, dt, wsum, gdd);
        lm += dl;
        for (int k = 0; k < 4; ++k) pp[k]->xi += w[k] * g[k] * dl;
    }
}

void neohookean_constraint_t::multiplier_displacement(simulation_t const& sim, vec3 d[4])
{
    auto const& P = sim.particles[bi];
    particle_t const* pp[4] = {&P[v[0]], &P[v[1]], &P[v[2]], &P[v[3]]};
    double const w[4] = {pp[0]->invmass(), pp[1]->invmass(), pp[2]->invmass(), pp[3]->invmass()};
    mat3 const Ds = mat3::from_cols(pp[0]->xi - pp[3]->xi, pp[1]->xi - pp[3]->xi, pp[2]->xi - pp[3]->xi);
    mat3 const F = Ds * DmInv;
    double const r = std::sqrt(frob2(F));
    vec3 gd[4], gh[4];
    tet_gradients((r > 0 ? 1.0 / r : 0.0) * F, DmInv, gd);
    tet_gradients(cofactor(F), DmInv, gh);
    double wd = 0, wh = 0;
    for (int k = 0; k < 4; ++k) {
        wd += w[k] * norm2(gd[k]);
        wh += w[k] * norm2(gh[k]);
    }
    if (sim.warm_start == 3) {  // the multipliers that zero the XPBD residual at the predicted positions
        double const dt2 = sim.warm_dt * sim.warm_dt, V0abs = std::abs(V0);
        lagrange = -r / (1.0 / (mu * V0abs) / dt2);
        lagrange_h = -(determinant(F) - (1.0 + mu / lam)) / (1.0 / (lam * V0abs) / dt2);
    }
    lagrange *= sim.warm_scale;
    lagrange_h *= sim.warm_scale_h;
    double const ld = wd < 1e-20 ? 0.0 : lagrange, lh = wh < 1e-20 ? 0.0 : lagrange_h;
    for (int k = 0; k < 4; ++k) d[k] = w[k] * (gd[k] * ld + gh[k] * lh);
}

void neohookean_constraint_t::residual_sq(simulation_t const& sim, double dt, double* acc) const
{
    auto const& P = sim.particles[bi];
    mat3 const Ds = mat3::from_cols(P[v[0]].x - P[v[3]].x, P[v[1]].x - P[v[3]].x, P[v[2]].x - P[v[3]].x);
    mat3 const F = Ds * DmInv;
    double const V0abs = std::abs(V0);
    double const dt2 = dt * dt;
    double const hd = std::sqrt(frob2(F)) + (1.0 / (mu * V0abs)) / dt2 * lagrange;
    double const hh = (determinant(F) - (1.0 + mu / lam)) + (1.0 / (lam * V0abs)) / dt2 * lagrange_h;
    *acc += hd * hd + hh * hh;
}

double neohookean_constraint_t::energy(simulation_t const& sim) const
{
    auto const& P = sim.particles[bi];
    mat3 const Ds = mat3::from_cols(P[v[0]].x - P[v[3]].x, P[v[1]].x - P[v[3]].x, P[v[2]].x - P[v[3]].x);
    mat3 const F = Ds * DmInv;
    double const V0abs = std::abs(V0);
    double const j = determinant(F) - (1.0 + mu / lam);
    return V0abs * (0.5 * mu * (frob2(F) - 3.0) + 0.5 * lam * j * j);
}

distance_constraint_t::distance_constraint_t(double alpha, double beta, simulation_t const& sim, uint32_t b1_,
                                             uint32_t v1_, uint32_t b2_, uint32_t v2_)
    : constraint_t(alpha, beta), b1(b1_), b2(b2_), v1(v1_), v2(v2_)
{
    d = norm(sim.particles.at(b1).at(v1).x0 - sim.particles.at(b2).at(v2).x0);  // (:18-21)
}

void distance_constraint_t::project_positions(simulation_t& sim, double dt)
{
    // distance_constraint.cpp:24-53 (zero length gives NaN there as well: no guard)
    particle_t& p1 = sim.particles[b1][v1];
    particle_t& p2 = sim.particles[b2][v2];
    double const w1 = p1.invmass(), w2 = p2.invmass();
    vec3 const diff = p1.xi - p2.xi;
    double const length = norm(diff);
    vec3 const n = diff / length;
    double const C = length - d;
    double const gdd = dot(n, p1.xi - p1.xn) - dot(n, p2.xi - p2.xn);
    double const dl = xpbd_delta(C, lagrange, alpha, beta, dt, w1 + w2, gdd);
    lagrange += dl;
    p1.xi += w1 * n * dl;
    p2.xi += w2 * (-n) * dl;
}

void distance_constraint_t::residual_sq(simulation_t const& sim, double dt, double* acc) const
{
    vec3 const diff = sim.particles[b1][v1].x - sim.particles[b2][v2].x;
    double const h = (norm(diff) - d) + (alpha / (dt * dt)) * lagrange;
    *acc += h * h;
}

void collision_constraint_t::project_positions(simulation_t& sim, double dt)
{
    // collision_constraint.cpp:21-41; beta is accepted but unused, as there
    particle_t& p = sim.particles[bi][vi];
    double const w = p.invmass();
    double const C = evaluate(p.xi);
    if (C >= 0.) return;
    double const alpha_tilde = alpha / (dt * dt);
    double const dl = -(C + alpha_tilde * lagrange) / (w + alpha_tilde);
    lagrange += dl;
    p.xi += w * n * dl;
}

void collision_constraint_t::residual_sq(simulation_t const& sim, double dt, double* acc) const
{
    double const C = evaluate(sim.particles[bi][vi].x);
    double const h = std::min(C, 0.0) + (alpha / (dt * dt)) * lagrange;
    *acc += h * h;
}

vec3 body_collision_constraint_t::plane_point(simulation_t const& sim) const
{
    vec3 q = w[0] * sim.contact_snap[cb[0]][cv[0]];
    q = q + w[1] * sim.contact_snap[cb[1]][cv[1]];
    q = q + w[2] * sim.contact_snap[cb[2]][cv[2]];
    return q + h * n;
}

void body_collision_constraint_t::project_positions(simulation_t& sim, double dt)
{
    particle_t& p = sim.particles[bi][vi];
    double const wi = p.invmass();
    double const C = dot(p.xi - plane_point(sim), n);
    if (C >= 0.) return;
    // the triangle point's inverse mass: the vertex takes its share of the two-sided projection
    double wt = 0.0;
    for (int k = 0; k < 3; ++k) wt += w[k] * w[k] * sim.particles[cb[k]][cv[k]].invmass();
    double const alpha_tilde = alpha / (dt * dt);
    double const dl = -(C + alpha_tilde * lagrange) / (wi + wt + alpha_tilde);
    lagrange += dl;
    p.xi += wi * n * dl;
}

void body_collision_constraint_t::residual_sq(simulation_t const& sim, double dt, double* acc) const
{
    double const C = dot(sim.particles[bi][vi].x - plane_point(sim), n);
    double const hh = std::min(C, 0.0) + (alpha / (dt * dt)) * lagrange;
    *acc += hh * hh;
}

void sdf_t::evaluate(vec3 p, double* sd, vec3* grad) const
{
    if (kind == 0) {
        *sd = dot(n, p) + offset;
        *grad = n;
    } else if (kind == 2) {
        *sd = grid->interpolate(p, grad);
    } else {
        vec3 const d = p - c;
        double const len = norm(d);
        *sd = len - r;
        *grad = len > 0 ? d / len : vec3{0.0, 1.0, 0.0};
    }
}

// Contact generation for every (tet body, SDF) pair in order, surface vertices in surface order.
// A surface vertex at x (the previous frame's position, bvh_model.cpp:75 reads the surface copy
// refreshed at timestep.cpp:63) produces a contact when sd(x) < 0 (bvh_model.cpp:76-80) and x lies
// in the SDF's englobing volume (the KD-tree traversal reaches every such vertex; vertices outside
// the volume depend on the tree's pruning and are not produced). Normal = grad/|grad|, contact
// point q = x + |sd| n (:82-84). The handler maps the surface vertex to its tet vertex and adds a
// collision constraint with the collision compliance/damping (contact_handler.cpp:39-52).
void simulation_t::detect_contacts()
{
    contacts.clear();
    collision_constraints.clear();
    have_body_contacts = false;
    for (uint32_t b = 0; b < particles.size(); ++b) {
        if (b >= surface.size()) continue;
        for (uint32_t k = 0; k < sdfs.size(); ++k) {
            for (uint32_t s = 0; s < surface[b].size(); ++s) {
                uint32_t const vi = surface[b][s];
                vec3 const x = particles[b][vi].x;
                double sd;
                vec3 grad;
                sdfs[k].evaluate(x, &sd, &grad);
                if (!(sd < 0.)) continue;
                if (!sdfs[k].in_volume(x)) continue;
                // Eigen's normalized(): a zero vector is returned unchanged
                double const gl2 = norm2(grad);
                vec3 const nn = gl2 > 0.0 ? grad / std::sqrt(gl2) : grad;
                vec3 const q = x + std::abs(sd) * nn;
                contacts.push_back({b, s, vi, k, q, nn});
                collision_constraints.push_back(std::make_unique<collision_constraint_t>(
                    collision_compliance, collision_damping, b, vi, q, nn));
            }
        }
    }
    if (body_contacts) detect_body_contacts();
}

void simulation_t::detect_body_contacts()
{
    // global surface vertices (body-major) and triangles, outward
    std::vector<std::pair<uint32_t, uint32_t>> sv;  // (body, surface vertex of the body)
    std::vector<uint32_t> base(particles.size() + 1, 0);
    for (uint32_t b = 0; b < particles.size(); ++b) {
        uint32_t const ns = b < surface.size() ? uint32_t(surface[b].size()) : 0u;
        for (uint32_t s = 0; s < ns; ++s) sv.emplace_back(b, s);
        base[b + 1] = base[b] + ns;
    }
    struct tri_t {
        uint32_t v[3], body;
    };
    std::vector<tri_t> tr;
    for (uint32_t b = 0; b < particles.size() && b < triangles.size(); ++b)
        for (std::size_t t = 0; t + 2 < triangles[b].size(); t += 3)
            tr.push_back({{base[b] + triangles[b][t], base[b] + triangles[b][t + 1], base[b] + triangles[b][t + 2]}, b});
    if (tr.empty() || sv.empty()) return;
    auto rest = [&](uint32_t s) { return particles[sv[s].first][surface[sv[s].first][sv[s].second]].x0; };
    auto cur = [&](uint32_t s) { return particles[sv[s].first][surface[sv[s].first][sv[s].second]].x; };
    std::vector<double> vol(particles.size(), 0.0);
    double edge_sum = 0.0;
    for (auto const& t : tr) {
        vec3 const a = rest(t.v[0]), b = rest(t.v[1]), c = rest(t.v[2]);
        vol[t.body] += a.x * (b.y * c.z - b.z * c.y) + a.y * (b.z * c.x - b.x * c.z) + a.z * (b.x * c.y - b.y * c.x);
        edge_sum += norm(a - b) + norm(b - c) + norm(c - a);
    }
    for (auto& t : tr)
        if (vol[t.body] < 0.0) std::swap(t.v[1], t.v[2]);
    double const mean_edge = edge_sum / (3.0 * double(tr.size()));
    // unit normals of the triangles at the current positions, the edge neighbours of each triangle
    // (edge k runs from corner k to corner k+1) and the angle-weighted pseudo-normals of the surface
    // vertices (Baerentzen and Aanaes, "Signed distance computation using the angle weighted
    // pseudonormal", IEEE TVCG 11(3), 2005): the sign of (p - c).N at the closest point c, with N the
    // pseudo-normal of the closest feature, tells whether p lies inside the closed surface
    std::vector<vec3> tn(tr.size());
    for (uint32_t t = 0; t < tr.size(); ++t) {
        vec3 const a = cur(tr[t].v[0]);
        vec3 const n = cross(cur(tr[t].v[1]) - a, cur(tr[t].v[2]) - a);
        double const l2 = norm2(n);
        double const l = std::sqrt(l2);
        tn[t] = l2 > 0.0 ? vec3{n.x / l, n.y / l, n.z / l} : vec3{0.0, 0.0, 0.0};
    }
    std::vector<std::array<uint32_t, 3>> nb(tr.size(), {UINT32_MAX, UINT32_MAX, UINT32_MAX});
    {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> open;  // directed edge -> (triangle, edge)
        for (uint32_t t = 0; t < tr.size(); ++t)
            for (int k = 0; k < 3; ++k) open[{tr[t].v[k], tr[t].v[(k + 1) % 3]}] = 3 * t + uint32_t(k);
        for (uint32_t t = 0; t < tr.size(); ++t)
            for (int k = 0; k < 3; ++k) {
                auto it = open.find({tr[t].v[(k + 1) % 3], tr[t].v[k]});
                if (it != open.end()) nb[t][k] = it->second / 3;
            }
    }
    std::vector<vec3> vn(sv.size(), vec3{0.0, 0.0, 0.0});
    for (uint32_t t = 0; t < tr.size(); ++t)
        for (int k = 0; k < 3; ++k) {
            vec3 const o = cur(tr[t].v[k]);
            vec3 const e1 = cur(tr[t].v[(k + 1) % 3]) - o, e2 = cur(tr[t].v[(k + 2) % 3]) - o;
            double const ang = std::atan2(std::sqrt(norm2(cross(e1, e2))), dot(e1, e2));
            vn[tr[t].v[k]] = vn[tr[t].v[k]] + ang * tn[t];
        }
    auto pseudo_normal = [&](uint32_t t, int feature) {
        switch (feature) {
        case 0: return vn[tr[t].v[0]];
        case 1: return vn[tr[t].v[1]];
        case 2: return vn[tr[t].v[2]];
        case 3: return nb[t][0] == UINT32_MAX ? tn[t] : tn[t] + tn[nb[t][0]];
        case 4: return nb[t][1] == UINT32_MAX ? tn[t] : tn[t] + tn[nb[t][1]];
        case 5: return nb[t][2] == UINT32_MAX ? tn[t] : tn[t] + tn[nb[t][2]];
        default: return tn[t];
        }
    };
    double const H = body_cell_request > 0.0 ? body_cell_request : 2.0 * mean_edge;
    double const h = body_thickness_request > 0.0 ? body_thickness_request : 0.1 * mean_edge;
    double const r = body_radius_request > 0.0 ? body_radius_request : mean_edge;
    double const inv_H = 1.0 / H;
    auto cell = [&](double v) { return (long long)std::floor(v * inv_H); };
    std::map<std::array<long long, 3>, std::vector<uint32_t>> grid;
    if (!body_brute_force)
        for (uint32_t t = 0; t < tr.size(); ++t) {
            vec3 const a = cur(tr[t].v[0]), b = cur(tr[t].v[1]), c = cur(tr[t].v[2]);
            long long lo[3], hi[3];
            double const mn[3] = {std::min({a.x, b.x, c.x}), std::min({a.y, b.y, c.y}), std::min({a.z, b.z, c.z})};
            double const mx[3] = {std::max({a.x, b.x, c.x}), std::max({a.y, b.y, c.y}), std::max({a.z, b.z, c.z})};
            for (int d = 0; d < 3; ++d) {
                lo[d] = cell(mn[d]);
                hi[d] = cell(mx[d]);
            }
            for (long long k = lo[2]; k <= hi[2]; ++k)
                for (long long j = lo[1]; j <= hi[1]; ++j)
                    for (long long i = lo[0]; i <= hi[0]; ++i) grid[{i, j, k}].push_back(t);
        }
    std::vector<uint32_t> all;
    if (body_brute_force) {
        all.resize(tr.size());
        for (uint32_t t = 0; t < tr.size(); ++t) all[t] = t;
    }
    uint32_t const nsdf = uint32_t(sdfs.size());
    for (uint32_t s = 0; s < sv.size(); ++s) {
        vec3 const p = cur(s);
        std::vector<uint32_t> near;  // the triangles of the 27 cells around p's cell
        if (!body_brute_force) {
            long long const ci = cell(p.x), cj = cell(p.y), ck = cell(p.z);
            for (long long k = ck - 1; k <= ck + 1; ++k)
                for (long long j = cj - 1; j <= cj + 1; ++j)
                    for (long long i = ci - 1; i <= ci + 1; ++i) {
                        auto it = grid.find({i, j, k});
                        if (it != grid.end()) near.insert(near.end(), it->second.begin(), it->second.end());
                    }
        }
        std::vector<uint32_t> const* cand = body_brute_force ? &all : &near;
        for (int own = 0; own < 2; ++own) {  // another body's nearest triangle, then its own surface's
            double best = r * r;
            uint32_t bt = UINT32_MAX;
            int bf = 6;
            vec3 bq{};
            double bw[3] = {0.0, 0.0, 0.0};
            for (uint32_t t : *cand) {
                if ((tr[t].body == sv[s].first) != (own != 0)) continue;
                if (tr[t].v[0] == s || tr[t].v[1] == s || tr[t].v[2] == s) continue;
                int feature;
                double wts[3];
                vec3 const a = cur(tr[t].v[0]), b = cur(tr[t].v[1]), c = cur(tr[t].v[2]);
                vec3 const q = closest_point_on_triangle(p, a, b, c, &feature, wts);
                double const d2 = norm2(p - q);
                if (!(d2 < best || (d2 == best && bt != UINT32_MAX && t < bt))) continue;
                if (own) {  // its own surface: more than h / 10 outside along the feature's pseudo-normal
                    vec3 const N = pseudo_normal(t, feature);
                    if (!(dot(p - q, N) > 0.1 * h * std::sqrt(norm2(N)))) continue;
                }
                best = d2;
                bt = t;
                bf = feature;
                bq = q;
                bw[0] = wts[0], bw[1] = wts[1], bw[2] = wts[2];
            }
            if (bt == UINT32_MAX) continue;
            // along the line from the closest point to the vertex, out of the surface (towards the
            // closest point when the vertex lies inside another body, by the sign of (p - c).N); the
            // triangle's normal when the vertex lies on the triangle
            vec3 n = tn[bt];
            vec3 const d = p - bq;
            double const dd = norm2(d);
            if (dd > 0.0) {
                double const l = dot(d, pseudo_normal(bt, bf)) < 0.0 ? -std::sqrt(dd) : std::sqrt(dd);
                n = vec3{d.x / l, d.y / l, d.z / l};
            }
            if (!(norm2(n) > 0.0)) continue;  // a degenerate triangle has no plane
            uint32_t const b = sv[s].first, vi = surface[b][sv[s].second];
            uint32_t cb[3], cvv[3];
            for (int k = 0; k < 3; ++k) {
                cb[k] = sv[tr[bt].v[k]].first;
                cvv[k] = surface[cb[k]][sv[tr[bt].v[k]].second];
            }
            // the contact record reports the barycentric weights in place of the contact point
            contacts.push_back({b, sv[s].second, vi, nsdf + bt, vec3{bw[0], bw[1], bw[2]}, n});
            collision_constraints.push_back(std::make_unique<body_collision_constraint_t>(
                collision_compliance, collision_damping, b, vi, cb, cvv, bw, h, n));
            have_body_contacts = true;
        }
    }
}

void simulation_t::build_order()
{
    order.clear();
    stage_off.assign(1, 0);
    stage_elems.clear();
    if (!colored_order) {
        for (uint32_t i = 0; i < constraints.size(); ++i) order.push_back(i);
    } else {
        for (auto const& g : groups) {
            std::size_t const ne = g.cluster_off.empty() ? g.count : g.cluster_off.size() - 1;
            // elements of each color in index order (one pass over the elements)
            std::vector<std::vector<uint32_t>> by_color(g.num_colors);
            for (std::size_t e = 0; e < ne; ++e) {
                std::size_t const b = g.cluster_off.empty() ? e : g.cluster_off[e];
                by_color[g.color[b]].push_back(uint32_t(e));
            }
            for (uint32_t c = 0; c < g.num_colors; ++c) {
                for (uint32_t e : by_color[c]) {
                    std::size_t const b = g.cluster_off.empty() ? e : g.cluster_off[e];
                    std::size_t const en = g.cluster_off.empty() ? e + 1 : g.cluster_off[e + 1];
                    uint32_t const at = uint32_t(order.size());
                    for (std::size_t i = b; i < en; ++i) order.push_back(uint32_t(g.first + i));
                    stage_elems.emplace_back(at, uint32_t(order.size()));
                }
                stage_off.push_back(stage_elems.size());
            }
        }
    }
    order_dirty = false;
}

// gauss_seidel_solver.cpp:8-37, with the elastic constraints visited in `order`.
// x = x~ + M^-1 grad C(x~)^T lambda over every neo-Hookean constraint, all gradients at the
// predicted positions (a sum, so the visiting order does not matter).
void simulation_t::apply_warm_start()
{
    std::vector<std::vector<vec3>> d(particles.size());
    for (std::size_t b = 0; b < particles.size(); ++b) d[b].assign(particles[b].size(), vec3{});
    // every constraint's displacement (independent: gradients at the predicted positions, each its
    // own multipliers), then the sums per particle in constraint order
    std::vector<neohookean_constraint_t*> nh;
    for (auto const& c : constraints)
        if (auto* p = dynamic_cast<neohookean_constraint_t*>(c.get())) nh.push_back(p);
    std::vector<vec3> dd(nh.size() * 4);
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (long i = 0; i < long(nh.size()); ++i) nh[i]->multiplier_displacement(*this, &dd[4 * std::size_t(i)]);
    for (std::size_t i = 0; i < nh.size(); ++i)
        for (int k = 0; k < 4; ++k) d[nh[i]->bi][nh[i]->v[k]] += dd[4 * i + k];
    for (std::size_t b = 0; b < particles.size(); ++b)
        for (std::size_t i = 0; i < particles[b].size(); ++i) particles[b][i].xi += d[b][i];
}

double bf16_round(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u = ((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16) << 16;
    float r;
    std::memcpy(&r, &u, 4);
    return double(r);
}

void warm_net_t::set(const float* p, const float* mu, const float* is, const float* os, int ni, int no)
{
    nin = ni;
    nout = no;
    auto take = [&](std::size_t& at, std::size_t n, bool round) {
        std::vector<double> v(n);
        for (std::size_t i = 0; i < n; ++i) v[i] = round ? bf16_round(p[at + i]) : double(p[at + i]);
        at += n;
        return v;
    };
    std::size_t at = 0;
    w1 = take(at, std::size_t(IN) * H, true);
    b1 = take(at, H, false);
    w2 = take(at, std::size_t(H) * H, true);
    b2 = take(at, H, false);
    w3 = take(at, std::size_t(H) * H, true);
    b3 = take(at, H, false);
    w4 = take(at, std::size_t(H) * OUT, true);
    b4 = take(at, OUT, false);
    mean.assign(mu, mu + IN);
    inv_std.assign(is, is + IN);
    out_scale.assign(os, os + no);
}

void warm_net_t::evaluate(const double* x, double* y) const
{
    double a0[IN], a1[H], a2[H], a3[H];
    for (int k = 0; k < IN; ++k) {
        float const z = k < nin ? (float(x[k]) - mean[k]) * inv_std[k] : 0.0f;
        a0[k] = bf16_round(z);
    }
    auto layer = [](const double* a, int n, const std::vector<double>& w, const std::vector<double>& b, int m,
                    double* out) {
        for (int o = 0; o < m; ++o) {
            double z = b[o];
            for (int i = 0; i < n; ++i) z += a[i] * w[std::size_t(i) * m + o];
            out[o] = z;
        }
    };
    layer(a0, IN, w1, b1, H, a1);
    for (int o = 0; o < H; ++o) a1[o] = bf16_round(float(std::max(a1[o], 0.0)));
    layer(a1, H, w2, b2, H, a2);
    for (int o = 0; o < H; ++o) a2[o] = bf16_round(float(std::max(a2[o], 0.0)));
    layer(a2, H, w3, b3, H, a3);
    for (int o = 0; o < H; ++o) a3[o] = bf16_round(float(std::max(a3[o], 0.0)));
    double out[OUT];
    layer(a3, H, w4, b4, OUT, out);
    for (int o = 0; o < nout; ++o) y[o] = double(out_scale[o]) * out[o];
}

std::vector<double> simulation_t::tets_per_body(group_t const& g) const
{
    std::vector<double> n(particles.size(), 0.0);
    for (std::size_t i = 0; i < g.count; ++i)
        n[static_cast<neohookean_constraint_t const*>(constraints[g.first + i].get())->bi] += 1.0;
    return n;
}

bool simulation_t::cluster_features(group_t const& g, std::size_t c, double dt, std::vector<double> const& per_body,
                                    double f[39], double atd[6], double ath[6]) const
{
    std::size_t const t0 = g.first + g.cluster_off[c], t1 = g.first + g.cluster_off[c + 1];
    if (t1 - t0 != 6) return false;  // hex cells only
    double const dt2 = dt * dt;
    for (int k = 0; k < 6; ++k) {
        auto const* T = static_cast<neohookean_constraint_t const*>(constraints[t0 + k].get());
        auto const& P = particles[T->bi];
        mat3 const Ds = mat3::from_cols(P[T->v[0]].xi - P[T->v[3]].xi, P[T->v[1]].xi - P[T->v[3]].xi,
                                        P[T->v[2]].xi - P[T->v[3]].xi);
        mat3 const F = Ds * T->DmInv;
        double const V0abs = std::abs(T->V0);
        atd[k] = (1.0 / (T->mu * V0abs)) / dt2;
        ath[k] = (1.0 / (T->lam * V0abs)) / dt2;
        f[2 * k] = atd[k] * T->lagrange;
        f[2 * k + 1] = ath[k] * T->lagrange_h;
        f[12 + 2 * k] = atd[k] * T->hist;
        f[13 + 2 * k] = ath[k] * T->hist_h;
        f[24 + 2 * k] = std::sqrt(frob2(F));
        f[25 + 2 * k] = determinant(F) - (1.0 + T->mu / T->lam);
    }
    f[36] = std::log(atd[0]);
    f[37] = std::log(ath[0]);
    f[38] = std::log(per_body[static_cast<neohookean_constraint_t const*>(constraints[t0].get())->bi]);
    return true;
}

void simulation_t::apply_learned_multipliers()
{
    for (auto const& g : groups) {
        if (g.kind != 1 || g.cluster_off.empty()) continue;
        long const nc = long(g.cluster_off.size()) - 1;
        std::vector<double> const per_body = tets_per_body(g);
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
        for (long c = 0; c < nc; ++c) {
            double f[39], atd[6], ath[6], y[12];
            if (!cluster_features(g, std::size_t(c), warm_dt, per_body, f, atd, ath)) continue;
            // the network corrects carried-over multipliers: a cluster whose multipliers or their
            // history are all zero (the first two substeps after a reset) keeps the previous start
            bool carried = false, history = false;
            for (int m = 0; m < 12; ++m) {
                carried = carried || float(f[m]) != 0.0f;
                history = history || float(f[12 + m]) != 0.0f;
            }
            if (carried && history) net->evaluate(f, y);
            std::size_t const t0 = g.first + g.cluster_off[c];
            for (int k = 0; k < 6; ++k) {
                auto* T = static_cast<neohookean_constraint_t*>(constraints[t0 + k].get());
                T->hist = T->lagrange;
                T->hist_h = T->lagrange_h;
                if (!(carried && history)) continue;
                T->lagrange += y[2 * k] / atd[k];
                T->lagrange_h += y[2 * k + 1] / ath[k];
            }
        }
    }
}

void simulation_t::solve(double dt, std::size_t iterations)
{
    if (order_dirty) build_order();
    for (auto& c : collision_constraints) c->prepare_for_projection();
    for (auto& c : constraints)
        if (!(warm_start && dynamic_cast<neohookean_constraint_t*>(c.get()))) c->prepare_for_projection();
    warm_dt = dt;
    if (warm_start == 2) apply_learned_multipliers();
    if (warm_start) apply_warm_start();
    bool const parallel = threads > 1 && colored_order;
    for (std::size_t k = 0; k < iterations; ++k) {
        if (have_body_contacts) {  // the positions this iteration's triangle contacts read
            contact_snap.resize(particles.size());
            for (std::size_t b = 0; b < particles.size(); ++b) {
                contact_snap[b].resize(particles[b].size());
                for (std::size_t i = 0; i < particles[b].size(); ++i) contact_snap[b][i] = particles[b][i].xi;
            }
        }
        for (auto& c : collision_constraints) c->project_positions(*this, dt);
        if (!parallel) {
            for (uint32_t i : order) constraints[i]->project_positions(*this, dt);
            continue;
        }
        for (std::size_t st = 0; st + 1 < stage_off.size(); ++st) {
            long const a = long(stage_off[st]), b = long(stage_off[st + 1]);
#pragma omp parallel for schedule(static) num_threads(threads)
            for (long e = a; e < b; ++e)
                for (uint32_t j = stage_elems[e].first; j < stage_elems[e].second; ++j)
                    constraints[order[j]]->project_positions(*this, dt);
        }
    }
}

// timestep.cpp:20-70. Gravity: with gravity_is_force the reference's f += g; a = f w
// (:39-40, particle.cpp:50-53); otherwise g is an acceleration applied to every particle with
// w > 0 (lumped-mass scenes).
void simulation_t::step(double dt_frame, std::size_t iterations, std::size_t substeps)
{
    double const dt = dt_frame / static_cast<double>(substeps);
    detect_contacts();  // (:29-30)
    for (std::size_t s = 0; s < substeps; ++s) {
        for (auto& body : particles)
            for (particle_t& p : body) {
                if (gravity_is_force) {
                    p.f += gravity;
                    p.v = p.v + p.a() * dt;
                } else if (p.invmass() > 0.0) {
                    p.v = p.v + gravity * dt;
                }
                p.xi = p.x + p.v * dt;
            }
        solve(dt, iterations);  // (:45)
        for (auto& body : particles)
            for (particle_t& p : body) {  // (:48-57)
                p.x = p.xi;
                p.v = (p.x - p.xn) / dt;
                p.xn = p.x;
                p.f = vec3{};
            }
    }
    collision_constraints.clear();  // (:68)
}

uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

uint32_t greedy_color_csr(uint32_t n, const uint32_t* eoff, const uint64_t* verts, uint32_t passes, bool index_order,
                          uint32_t* color)
{
    std::size_t const nv = eoff[n];
    // vertex keys -> dense ids, then vertex -> elements incidence
    std::vector<uint64_t> uniq(verts, verts + nv);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    std::vector<uint32_t> vid(nv);
    for (std::size_t i = 0; i < nv; ++i)
        vid[i] = uint32_t(std::lower_bound(uniq.begin(), uniq.end(), verts[i]) - uniq.begin());
    std::vector<uint32_t> off(uniq.size() + 1, 0), inc(nv);
    for (uint32_t x : vid) off[x + 1]++;
    for (std::size_t i = 0; i < uniq.size(); ++i) off[i + 1] += off[i];
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (uint32_t t = 0; t < n; ++t)
        for (uint32_t a = eoff[t]; a < eoff[t + 1]; ++a) inc[fill[vid[a]]++] = t;

    std::vector<uint32_t> prev(n, 0);
    uint32_t ncolors = 0;
    uint32_t const npass = index_order ? 1 : passes + 1;
    for (uint32_t pass = 0; pass < npass; ++pass) {
        std::vector<uint32_t> ord(n);
        for (uint32_t t = 0; t < n; ++t) ord[t] = t;
        if (!index_order) {
            std::vector<uint64_t> key(n);
            for (uint32_t t = 0; t < n; ++t) key[t] = (uint64_t(pass == 0 ? 0 : prev[t]) << 32) | hash32(t);
            std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
                return key[a] != key[b] ? key[a] > key[b] : a < b;
            });
        }
        std::vector<int64_t> col(n, -1);
        ncolors = 0;
        for (uint32_t t : ord) {
            std::vector<bool> used;
            for (uint32_t a = eoff[t]; a < eoff[t + 1]; ++a) {
                uint32_t const x = vid[a];
                for (uint32_t e = off[x]; e < off[x + 1]; ++e) {
                    uint32_t const u = inc[e];
                    if (u == t || col[u] < 0) continue;
                    if (used.size() <= std::size_t(col[u])) used.resize(col[u] + 1, false);
                    used[col[u]] = true;
                }
            }
            uint32_t c = 0;
            while (c < used.size() && used[c]) ++c;
            col[t] = c;
            ncolors = std::max(ncolors, c + 1);
        }
        for (uint32_t t = 0; t < n; ++t) prev[t] = uint32_t(col[t]);
    }
    std::copy(prev.begin(), prev.end(), color);
    return ncolors;
}

uint32_t greedy_color(uint32_t n, uint32_t arity, const uint64_t* verts, uint32_t passes, uint32_t* color)
{
    std::vector<uint32_t> eoff(std::size_t(n) + 1);
    for (uint32_t t = 0; t <= n; ++t) eoff[t] = t * arity;
    return greedy_color_csr(n, eoff.data(), verts, passes, false, color);
}

}  // namespace oracle
