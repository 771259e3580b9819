// C entry points of the FP64 oracle, loaded with ctypes by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg.
//
// TEST INFRASTRUCTURE ONLY: the oracle is the checker, never the thing measured as the product or
// shipped in it.
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "mesh.h"
#include "xpbd.h"

using namespace oracle;

extern "C" {

void* orc_create() { return new simulation_t(); }
void orc_destroy(void* h) { delete static_cast<simulation_t*>(h); }

void orc_set_params(void* h, double gx, double gy, double gz, int gravity_is_force, double collision_compliance,
                    double collision_damping)
{
    auto* s = static_cast<simulation_t*>(h);
    s->gravity = {gx, gy, gz};
    s->gravity_is_force = gravity_is_force != 0;
    s->collision_compliance = collision_compliance;
    s->collision_damping = collision_damping;
}

uint32_t orc_add_body(void* h, const double* x0, const double* mass, uint32_t n)
{
    auto* s = static_cast<simulation_t*>(h);
    std::vector<particle_t> ps;
    ps.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        particle_t p({x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]});
        if (mass) p.m = mass[i];
        ps.push_back(p);
    }
    s->particles.push_back(std::move(ps));
    s->surface.emplace_back();
    return uint32_t(s->particles.size() - 1);
}

// Overwrites x (and xi, xn, which equal x between frames) and/or v of a body.
void orc_set_body_state(void* h, uint32_t body, const double* x, const double* v)
{
    auto* s = static_cast<simulation_t*>(h);
    auto& P = s->particles.at(body);
    for (std::size_t i = 0; i < P.size(); ++i) {
        if (x) P[i].x = P[i].xi = P[i].xn = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
        if (v) P[i].v = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    }
}

void orc_set_surface(void* h, uint32_t body, const uint32_t* map, uint32_t n)
{
    auto* s = static_cast<simulation_t*>(h);
    s->surface.at(body).assign(map, map + n);
}

void orc_set_surface_triangles(void* h, uint32_t body, const uint32_t* tris, uint32_t n)
{
    auto* s = static_cast<simulation_t*>(h);
    if (s->triangles.size() < s->particles.size()) s->triangles.resize(s->particles.size());
    s->triangles.at(body).assign(tris, tris + 3 * std::size_t(n));
}

void orc_set_body_contacts(void* h, int enable, double cell, double thickness, double radius, int brute_force)
{
    auto* s = static_cast<simulation_t*>(h);
    s->body_contacts = enable != 0;
    s->body_cell_request = cell;
    s->body_thickness_request = thickness;
    s->body_radius_request = radius;
    s->body_brute_force = brute_force != 0;
}

static int add_group(simulation_t* s, int kind, std::size_t first)
{
    group_t g;
    g.kind = kind;
    g.first = first;
    g.count = s->constraints.size() - first;
    s->groups.push_back(g);
    s->order_dirty = true;
    return int(s->groups.size() - 1);
}

int orc_add_green(void* h, uint32_t body, const uint32_t* tets, uint32_t n, double alpha, double beta,
                  double young, double poisson)
{
    auto* s = static_cast<simulation_t*>(h);
    std::size_t const first = s->constraints.size();
    for (uint32_t t = 0; t < n; ++t)
        s->constraints.push_back(
            std::make_unique<green_constraint_t>(alpha, beta, *s, body, tets + 4 * t, young, poisson));
    return add_group(s, 0, first);
}

int orc_add_neohookean(void* h, uint32_t body, const uint32_t* tets, uint32_t n, double beta, double young,
                       double poisson)
{
    auto* s = static_cast<simulation_t*>(h);
    std::size_t const first = s->constraints.size();
    for (uint32_t t = 0; t < n; ++t)
        s->constraints.push_back(
            std::make_unique<neohookean_constraint_t>(beta, *s, body, tets + 4 * t, young, poisson));
    return add_group(s, 1, first);
}

// One group of tets spanning several bodies, each tet with its own body and material.
int orc_add_green_batched(void* h, const uint32_t* bodies, const uint32_t* tets, uint32_t n, double alpha, double beta,
                          const double* young, const double* poisson)
{
    auto* s = static_cast<simulation_t*>(h);
    std::size_t const first = s->constraints.size();
    for (uint32_t t = 0; t < n; ++t)
        s->constraints.push_back(
            std::make_unique<green_constraint_t>(alpha, beta, *s, bodies[t], tets + 4 * t, young[t], poisson[t]));
    return add_group(s, 0, first);
}

int orc_add_neohookean_batched(void* h, const uint32_t* bodies, const uint32_t* tets, uint32_t n, double beta,
                               const double* young, const double* poisson)
{
    auto* s = static_cast<simulation_t*>(h);
    std::size_t const first = s->constraints.size();
    for (uint32_t t = 0; t < n; ++t)
        s->constraints.push_back(
            std::make_unique<neohookean_constraint_t>(beta, *s, bodies[t], tets + 4 * t, young[t], poisson[t]));
    return add_group(s, 1, first);
}

int orc_add_distance(void* h, const uint32_t* b1v1b2v2, uint32_t n, double alpha, double beta)
{
    auto* s = static_cast<simulation_t*>(h);
    std::size_t const first = s->constraints.size();
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t const* e = b1v1b2v2 + 4 * i;
        s->constraints.push_back(std::make_unique<distance_constraint_t>(alpha, beta, *s, e[0], e[1], e[2], e[3]));
    }
    return add_group(s, 2, first);
}

void orc_add_plane(void* h, const double n[3], const double point[3], const double vmin[3], const double vmax[3])
{
    auto* s = static_cast<simulation_t*>(h);
    sdf_t f;
    f.kind = 0;
    f.n = {n[0], n[1], n[2]};
    f.offset = -dot(f.n, vec3{point[0], point[1], point[2]});
    f.vmin = {vmin[0], vmin[1], vmin[2]};
    f.vmax = {vmax[0], vmax[1], vmax[2]};
    s->sdfs.push_back(f);
}

void orc_add_sphere(void* h, const double c[3], double r)
{
    auto* s = static_cast<simulation_t*>(h);
    sdf_t f;
    f.kind = 1;
    f.c = {c[0], c[1], c[2]};
    f.r = r;
    f.vmin = {c[0] - r, c[1] - r, c[2] - r};
    f.vmax = {c[0] + r, c[1] + r, c[2] + r};
    s->sdfs.push_back(f);
}

// Grid SDF obstacle: domain, resolution (cells per axis), node values, englobing volume.
void orc_add_grid_sdf(void* h, const double lo[3], const double hi[3], const uint32_t res[3], const double* nodes,
                      const double vmin[3], const double vmax[3])
{
    auto* s = static_cast<simulation_t*>(h);
    sdf_t f;
    f.kind = 2;
    f.grid = std::make_shared<grid_sdf_t>();
    f.grid->init(vec3{lo[0], lo[1], lo[2]}, vec3{hi[0], hi[1], hi[2]}, res);
    std::copy(nodes, nodes + f.grid->nodes.size(), f.grid->nodes.begin());
    f.vmin = {vmin[0], vmin[1], vmin[2]};
    f.vmax = {vmax[0], vmax[1], vmax[2]};
    s->sdfs.push_back(f);
}

// sdf k evaluated at n points: sd (n) and grad (3 n)
void orc_evaluate_sdf(void* h, uint32_t k, const double* pts, uint32_t n, double* sd, double* grad)
{
    auto* s = static_cast<simulation_t*>(h);
    for (uint32_t i = 0; i < n; ++i) {
        vec3 g;
        s->sdfs[k].evaluate(vec3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, &sd[i], &g);
        grad[3 * i] = g.x;
        grad[3 * i + 1] = g.y;
        grad[3 * i + 2] = g.z;
    }
}

uint64_t orc_grid_node_count(const uint32_t res[3])
{
    grid_sdf_t g;
    for (int d = 0; d < 3; ++d) g.res[d] = res[d];
    return g.node_count();
}

void orc_grid_node_positions(const double lo[3], const double hi[3], const uint32_t res[3], double* out)
{
    grid_sdf_t g;
    g.init(vec3{lo[0], lo[1], lo[2]}, vec3{hi[0], hi[1], hi[2]}, res);
    for (uint64_t l = 0; l < g.node_count(); ++l) {
        vec3 const x = g.node_position(l);
        out[3 * l] = x.x;
        out[3 * l + 1] = x.y;
        out[3 * l + 2] = x.z;
    }
}

// Node values of a triangle mesh's signed distance (MeshDistance at every grid node).
void orc_grid_from_mesh(const double* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, const double lo[3],
                        const double hi[3], const uint32_t res[3], double* nodes)
{
    grid_sdf_t g;
    g.init(vec3{lo[0], lo[1], lo[2]}, vec3{hi[0], hi[1], hi[2]}, res);
    mesh_distance_t const md(pos, nv, tris, nt);
    for (uint64_t l = 0; l < g.node_count(); ++l) nodes[l] = md.signed_distance(g.node_position(l));
}

void orc_mesh_signed_distance(const double* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, const double* pts,
                              uint32_t n, double* out)
{
    mesh_distance_t const md(pos, nv, tris, nt);
    for (uint32_t i = 0; i < n; ++i) out[i] = md.signed_distance(vec3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]});
}

void orc_environment_domain(const double* pos, uint32_t nv, const double lo[3], const double hi[3], double out_lo[3],
                            double out_hi[3])
{
    vec3 a, b;
    environment_domain(pos, nv, vec3{lo[0], lo[1], lo[2]}, vec3{hi[0], hi[1], hi[2]}, &a, &b);
    for (int d = 0; d < 3; ++d) {
        out_lo[d] = a[d];
        out_hi[d] = b[d];
    }
}

static void constraint_vertices(simulation_t const& s, group_t const& g, std::vector<uint64_t>& keys, uint32_t* arity)
{
    keys.clear();
    if (g.kind == 2) {
        *arity = 2;
        for (std::size_t i = 0; i < g.count; ++i) {
            auto const* d = static_cast<distance_constraint_t const*>(s.constraints[g.first + i].get());
            keys.push_back((uint64_t(d->b1) << 32) | d->v1);
            keys.push_back((uint64_t(d->b2) << 32) | d->v2);
        }
    } else {
        *arity = 4;
        for (std::size_t i = 0; i < g.count; ++i) {
            constraint_t const* c = s.constraints[g.first + i].get();
            uint32_t bi;
            uint32_t const* v;
            if (g.kind == 0) {
                auto const* t = static_cast<green_constraint_t const*>(c);
                bi = t->bi;
                v = t->v;
            } else {
                auto const* t = static_cast<neohookean_constraint_t const*>(c);
                bi = t->bi;
                v = t->v;
            }
            for (int k = 0; k < 4; ++k) keys.push_back((uint64_t(bi) << 32) | v[k]);
        }
    }
}

void orc_set_clusters(void* h, uint32_t group, const uint32_t* offsets, uint32_t n)
{
    auto* s = static_cast<simulation_t*>(h);
    auto& g = s->groups.at(group);
    if (n == 0)
        g.cluster_off.clear();
    else
        g.cluster_off.assign(offsets, offsets + n + 1);
    s->order_dirty = true;
}

void orc_set_coloring(void* h, int order) { static_cast<simulation_t*>(h)->coloring = order; }
// Threads of the colored order's stages (1: sequential); the result is the same bit for bit.
void orc_set_threads(void* h, int threads) { static_cast<simulation_t*>(h)->threads = threads < 1 ? 1 : threads; }
// The learned warm start's network: FP32 parameters in the device layout, input normalization (48
// values each) and output scale (nout values).
void orc_set_network(void* h, const float* params, const float* mean, const float* inv_std, const float* out_scale,
                     int nin, int nout)
{
    auto net = std::make_shared<warm_net_t>();
    net->set(params, mean, inv_std, out_scale, nin, nout);
    static_cast<simulation_t*>(h)->net = net;
}

// Test hooks: the network on n rows of 39 raw features (y: n x nout), and the features of the
// clusters of group `group` at the current positions (set with orc_set_body_state) for a substep of dt
// (out: clusters x 39; rows of clusters that are not 6-tet cells are zero).
void orc_network_forward(void* h, const double* x, uint32_t n, double* y)
{
    auto const& net = *static_cast<simulation_t*>(h)->net;
    for (uint32_t r = 0; r < n; ++r) net.evaluate(x + std::size_t(r) * 39, y + std::size_t(r) * net.nout);
}

// Green projection branch trace (test hook): on = 1 starts a new trace, 0 stops; read copies up
// to cap entries and returns the trace length.
void orc_green_trace(void* h, int on)
{
    auto* s = static_cast<simulation_t*>(h);
    s->green_trace = on ? std::make_unique<std::vector<uint8_t>>() : nullptr;
}

uint64_t orc_green_trace_read(void* h, uint8_t* out, uint64_t cap)
{
    auto* s = static_cast<simulation_t*>(h);
    if (!s->green_trace) return 0;
    uint64_t const n = s->green_trace->size();
    if (out) std::memcpy(out, s->green_trace->data(), std::min(n, cap));
    return n;
}

// (lambda, lambda_h, hist, hist_h) of the neo-Hookean constraints of a group (test hook).
void orc_multipliers(void* h, uint32_t group, double* out)
{
    auto* s = static_cast<simulation_t*>(h);
    auto const& g = s->groups.at(group);
    for (std::size_t i = 0; i < g.count; ++i) {
        auto const* T = dynamic_cast<neohookean_constraint_t const*>(s->constraints[g.first + i].get());
        double* o = out + 4 * i;
        o[0] = T ? T->lagrange : 0.0;
        o[1] = T ? T->lagrange_h : 0.0;
        o[2] = T ? T->hist : 0.0;
        o[3] = T ? T->hist_h : 0.0;
    }
}

uint32_t orc_cluster_features(void* h, uint32_t group, double dt, double* out)
{
    auto* s = static_cast<simulation_t*>(h);
    auto const& g = s->groups.at(group);
    if (g.cluster_off.empty()) return 0;
    uint32_t const nc = uint32_t(g.cluster_off.size() - 1);
    std::vector<double> const per_body = s->tets_per_body(g);
    for (uint32_t c = 0; c < nc; ++c) {
        double atd[6], ath[6];
        if (!s->cluster_features(g, c, dt, per_body, out + std::size_t(c) * 39, atd, ath))
            std::fill(out + std::size_t(c) * 39, out + std::size_t(c + 1) * 39, 0.0);
    }
    return nc;
}

void orc_set_warm_start(void* h, int mode, double scale, double scale_h)
{
    static_cast<simulation_t*>(h)->warm_start = mode;
    static_cast<simulation_t*>(h)->warm_scale = scale;
    static_cast<simulation_t*>(h)->warm_scale_h = scale_h;
}

// colored != 0 selects the colored visiting order (and computes the coloring of every group).
void orc_set_order(void* h, int colored, uint32_t recolor_passes)
{
    auto* s = static_cast<simulation_t*>(h);
    s->colored_order = colored != 0;
    s->recolor_passes = recolor_passes;
    if (s->colored_order) {
        std::vector<uint64_t> keys;
        for (auto& g : s->groups) {
            uint32_t arity;
            constraint_vertices(*s, g, keys, &arity);
            g.color.assign(g.count, 0);
            g.num_colors = 0;
            if (!g.count) continue;
            // elements: clusters, or single constraints; element vertex lists without duplicates
            std::size_t const ne = g.cluster_off.empty() ? g.count : g.cluster_off.size() - 1;
            std::vector<uint32_t> eoff(ne + 1, 0);
            std::vector<uint64_t> ev;
            for (std::size_t e = 0; e < ne; ++e) {
                std::size_t const b = g.cluster_off.empty() ? e : g.cluster_off[e];
                std::size_t const en = g.cluster_off.empty() ? e + 1 : g.cluster_off[e + 1];
                std::size_t const start = ev.size();
                for (std::size_t c = b; c < en; ++c)
                    for (uint32_t a = 0; a < arity; ++a) {
                        uint64_t const k = keys[c * arity + a];
                        if (std::find(ev.begin() + start, ev.end(), k) == ev.end()) ev.push_back(k);
                    }
                eoff[e + 1] = uint32_t(ev.size());
            }
            std::vector<uint32_t> ec(ne);
            g.num_colors = greedy_color_csr(uint32_t(ne), eoff.data(), ev.data(), recolor_passes, s->coloring == 1,
                                            ec.data());
            for (std::size_t e = 0; e < ne; ++e) {
                std::size_t const b = g.cluster_off.empty() ? e : g.cluster_off[e];
                std::size_t const en = g.cluster_off.empty() ? e + 1 : g.cluster_off[e + 1];
                for (std::size_t c = b; c < en; ++c) g.color[c] = ec[e];
            }
        }
    }
    s->order_dirty = true;
}

uint32_t orc_group_colors(void* h, uint32_t group, uint32_t* color_out)
{
    auto* s = static_cast<simulation_t*>(h);
    auto const& g = s->groups.at(group);
    if (color_out && !g.color.empty()) std::memcpy(color_out, g.color.data(), g.count * sizeof(uint32_t));
    return g.num_colors;
}

void orc_step(void* h, double dt, uint32_t iterations, uint32_t substeps)
{
    static_cast<simulation_t*>(h)->step(dt, iterations, substeps);
}

void orc_read_body(void* h, uint32_t body, double* x, double* v)
{
    auto* s = static_cast<simulation_t*>(h);
    auto const& P = s->particles.at(body);
    for (std::size_t i = 0; i < P.size(); ++i) {
        if (x) {
            x[3 * i] = P[i].x.x;
            x[3 * i + 1] = P[i].x.y;
            x[3 * i + 2] = P[i].x.z;
        }
        if (v) {
            v[3 * i] = P[i].v.x;
            v[3 * i + 1] = P[i].v.y;
            v[3 * i + 2] = P[i].v.z;
        }
    }
}

uint32_t orc_detect_contacts(void* h)
{
    auto* s = static_cast<simulation_t*>(h);
    s->detect_contacts();
    return uint32_t(s->contacts.size());
}

// ids: 4 per contact (body, surface vertex, body vertex, sdf); qn: 6 per contact (q, n).
uint32_t orc_read_contacts(void* h, uint32_t* ids, double* qn, uint32_t cap)
{
    auto* s = static_cast<simulation_t*>(h);
    uint32_t const n = uint32_t(std::min<std::size_t>(cap, s->contacts.size()));
    for (uint32_t i = 0; i < n; ++i) {
        auto const& c = s->contacts[i];
        if (ids) {
            ids[4 * i] = c.body;
            ids[4 * i + 1] = c.surface_vertex;
            ids[4 * i + 2] = c.vertex;
            ids[4 * i + 3] = c.sdf;
        }
        if (qn) {
            double const v[6] = {c.q.x, c.q.y, c.q.z, c.n.x, c.n.y, c.n.z};
            std::memcpy(qn + 6 * i, v, sizeof v);
        }
    }
    return uint32_t(s->contacts.size());
}

// Squared-residual sums of one group's constraints, per color (colored order) or in one bucket,
// evaluated at the current positions with the multipliers of the last substep.
void orc_residual(void* h, uint32_t group, double dt, double* per_color, uint32_t cap)
{
    auto* s = static_cast<simulation_t*>(h);
    auto const& g = s->groups.at(group);
    for (uint32_t c = 0; c < cap; ++c) per_color[c] = 0;
    for (std::size_t i = 0; i < g.count; ++i) {
        uint32_t const c = g.color.empty() ? 0 : g.color[i];
        if (c < cap) s->constraints[g.first + i]->residual_sq(*s, dt, &per_color[c]);
    }
}

// out: kinetic, gravitational potential, elastic, momentum x/y/z.
void orc_energy(void* h, double* out)
{
    auto* s = static_cast<simulation_t*>(h);
    double ke = 0, pe = 0, ee = 0, px = 0, py = 0, pz = 0;
    for (auto const& body : s->particles)
        for (auto const& p : body) {
            double const m = p.m;
            ke += 0.5 * m * norm2(p.v);
            pe -= (s->gravity_is_force ? 1.0 : m) * dot(s->gravity, p.x);
            px += m * p.v.x;
            py += m * p.v.y;
            pz += m * p.v.z;
        }
    for (auto const& c : s->constraints) ee += c->energy(*s);
    double const r[6] = {ke, pe, ee, px, py, pz};
    std::memcpy(out, r, sizeof r);
}

void orc_svd3(const double* F, double* U, double* sv, double* V)
{
    mat3 a, u, v;
    for (int i = 0; i < 9; ++i) a.m[i / 3][i % 3] = F[i];
    jacobi_svd3(a, u, sv, v);
    for (int i = 0; i < 9; ++i) {
        U[i] = u.m[i / 3][i % 3];
        V[i] = v.m[i / 3][i % 3];
    }
}

void orc_inverse3(const double* A, double* out, double* det)
{
    mat3 a;
    for (int i = 0; i < 9; ++i) a.m[i / 3][i % 3] = A[i];
    mat3 const r = inverse(a);
    for (int i = 0; i < 9; ++i) out[i] = r.m[i / 3][i % 3];
    *det = determinant(a);
}

// Green constraint value C = |V0| psi for a tet with rest positions x0 (12) at positions p (12).
double orc_green_value(const double* x0, const double* p, double young, double poisson)
{
    simulation_t s;
    uint32_t const vs[4] = {0, 1, 2, 3};
    std::vector<particle_t> ps;
    for (int i = 0; i < 4; ++i) ps.push_back(particle_t({x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]}));
    s.particles.push_back(ps);
    green_constraint_t g(0.0, 0.0, s, 0, vs, young, poisson);
    vec3 const q[4] = {{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}, {p[9], p[10], p[11]}};
    return g.evaluate(q);
}

// Green constraint value and forces f1..f4 (= -grad C, 12 doubles) for one tet.
double orc_green_eval(const double* x0, const double* p, double young, double poisson, double* f)
{
    simulation_t s;
    uint32_t const vs[4] = {0, 1, 2, 3};
    std::vector<particle_t> ps;
    for (int i = 0; i < 4; ++i) ps.push_back(particle_t({x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]}));
    s.particles.push_back(ps);
    green_constraint_t g(0.0, 0.0, s, 0, vs, young, poisson);
    vec3 const q[4] = {{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}, {p[9], p[10], p[11]}};
    vec3 fk[4];
    double const C = g.evaluate_forces(q, fk);
    for (int k = 0; k < 4; ++k) {
        f[3 * k] = fk[k].x;
        f[3 * k + 1] = fk[k].y;
        f[3 * k + 2] = fk[k].z;
    }
    return C;
}

uint32_t orc_bar5(uint32_t w, uint32_t hh, uint32_t d, double* pos, uint32_t* tets)
{
    tet_mesh const m = bar5(w, hh, d);
    if (pos) std::memcpy(pos, m.positions.data(), m.positions.size() * sizeof(double));
    if (tets) std::memcpy(tets, m.tets.data(), m.tets.size() * sizeof(uint32_t));
    return uint32_t(m.tets.size() / 4);
}

uint32_t orc_box6(uint32_t nx, uint32_t ny, uint32_t nz, double edge, const double* origin, double* pos, uint32_t* tets)
{
    tet_mesh const m = box6(nx, ny, nz, edge, origin);
    if (pos) std::memcpy(pos, m.positions.data(), m.positions.size() * sizeof(double));
    if (tets) std::memcpy(tets, m.tets.data(), m.tets.size() * sizeof(uint32_t));
    return uint32_t(m.tets.size() / 4);
}

// Boundary extraction; call once with null outputs to get the sizes.
void orc_boundary(uint32_t nv, const uint32_t* tets, uint32_t nt, uint32_t* s2t, uint32_t* tris, uint32_t* tri2face,
                  uint32_t* nsurf, uint32_t* ntris)
{
    boundary const b = extract_boundary(nv, tets, nt);
    *nsurf = uint32_t(b.surface_to_tet.size());
    *ntris = uint32_t(b.triangle_to_face.size());
    if (s2t) std::memcpy(s2t, b.surface_to_tet.data(), b.surface_to_tet.size() * sizeof(uint32_t));
    if (tris) std::memcpy(tris, b.triangles.data(), b.triangles.size() * sizeof(uint32_t));
    if (tri2face) std::memcpy(tri2face, b.triangle_to_face.data(), b.triangle_to_face.size() * sizeof(uint32_t));
}

uint32_t orc_greedy_color(uint32_t n, uint32_t arity, const uint64_t* verts, uint32_t passes, uint32_t* color)
{
    return greedy_color(n, arity, verts, passes, color);
}

uint32_t orc_hash32(uint32_t x) { return hash32(x); }

}  // extern "C"
