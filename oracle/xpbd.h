// FP64 restatement of the reference XPBD time step.
//
// TEST INFRASTRUCTURE ONLY: this is the oracle the GPU path is checked against and the CPU
// baseline bench.py times. Product code never links it.
//
// Structure deliberately follows the reference (AoS particles, heap-allocated constraints with a
// virtual project_positions, sequential Gauss-Seidel), so that its timing is representative of
// the reference's CPU solver:
//   particle_t                  src/physics/particle.cpp:6-53, include/sbs/physics/particle.h:46-52
//   constraint_t                src/physics/constraint.cpp:7-16
//   green_constraint_t          src/physics/xpbd/green_constraint.cpp:11-173
//   distance_constraint_t       src/physics/xpbd/distance_constraint.cpp:8-53
//   collision_constraint_t      src/physics/xpbd/collision_constraint.cpp:9-48
//   gauss_seidel_solver_t       src/physics/gauss_seidel_solver.cpp:8-37
//   timestep_t::step            src/physics/timestep.cpp:20-70
//   contact generation          src/physics/collision/bvh_model.cpp:66-95, sdf_model.cpp:52-75,
//                               src/physics/xpbd/contact_handler.cpp:14-54
// and adds the components the reference lacks (stable neo-Hookean constraint, colored visiting
// order, sphere SDFs, residual norms, energies), each documented where it is defined.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "grid_sdf.h"
#include "linalg.h"

namespace oracle {

struct particle_t {
    vec3 x0, xi, xn, x, v, f;
    double m = 1.0;  // particle.cpp:7 (unit mass by default)
    explicit particle_t(vec3 p) : x0(p), xi(p), xn(p), x(p) {}
    double invmass() const { return m > 0.0 ? 1.0 / m : 0.0; }  // particle.cpp:39-44
    vec3 a() const { return f * invmass(); }                    // particle.cpp:50-53
};

struct simulation_t;

struct constraint_t {
    double alpha, beta, lagrange = 0.0;
    constraint_t(double a, double b) : alpha(a), beta(b) {}
    virtual ~constraint_t() = default;
    void prepare_for_projection() { lagrange = 0.0; prepare_impl(); }  // constraint.cpp:12-16
    virtual void prepare_impl() {}
    virtual void project_positions(simulation_t& sim, double dt) = 0;
    // Adds the squared XPBD residual h = C(x) + alpha~ * lambda of this constraint, evaluated at
    // the particles' current x, to *acc. Returns nothing else.
    virtual void residual_sq(simulation_t const& sim, double dt, double* acc) const = 0;
    virtual double energy(simulation_t const& sim) const { return 0.0; }
};

struct green_constraint_t final : constraint_t {
    uint32_t bi, v[4];
    mat3 DmInv;
    double V0, mu, lam;
    green_constraint_t(double alpha, double beta, simulation_t const& sim, uint32_t body,
                       uint32_t const vs[4], double young, double poisson);
    void project_positions(simulation_t& sim, double dt) override;
    void residual_sq(simulation_t const& sim, double dt, double* acc) const override;
    double energy(simulation_t const& sim) const override;
    // C = |V0| psi at positions p (the constraint value used by project_positions).
    double evaluate(vec3 const p[4], bool* early_out = nullptr, double const w[4] = nullptr) const;
    double evaluate_forces(vec3 const p[4], vec3 f[4]) const;
};

// Stable neo-Hookean as two XPBD constraints per tet (Macklin & Mueller 2021, "A Constraint-based
// Formulation of Stable Neo-Hookean Materials"): deviatoric C_D = ||F||_F with compliance
// 1/(mu |V0|), then hydrostatic C_H = det F - (1 + mu/lambda) with compliance 1/(lambda |V0|),
// projected one after the other, each with its own multiplier. No reference counterpart.
struct neohookean_constraint_t final : constraint_t {
    uint32_t bi, v[4];
    mat3 DmInv;
    double V0, mu, lam;
    double lagrange_h = 0.0;
    double hist = 0.0, hist_h = 0.0;  // the multipliers one substep before (learned warm start)
    neohookean_constraint_t(double beta, simulation_t const& sim, uint32_t body, uint32_t const vs[4],
                            double young, double poisson);
    void prepare_impl() override { lagrange_h = 0.0; }
    void project_positions(simulation_t& sim, double dt) override;
    // Warm start: d[k] = M^-1 (grad_k C_dev lambda + grad_k C_hyd lambda_h) at the current positions
    // for the tet's vertices k.
    // With simulation_t::warm_start == 3 the multipliers are first set to -C / alpha~; then they are
    // multiplied by simulation_t::warm_scale (deviatoric) and warm_scale_h (hydrostatic).
    void multiplier_displacement(simulation_t const& sim, vec3 d[4]);
    void residual_sq(simulation_t const& sim, double dt, double* acc) const override;
    double energy(simulation_t const& sim) const override;
};

// The learned warm start's network (no reference counterpart): 48 inputs -> 64 -> 64 -> 64 (ReLU)
// -> 32 outputs, the parameter layout of the device network (W1[48][64] b1 W2[64][64] b2 W3[64][64]
// b3 W4[64][32] b4). Evaluated in FP64 with the device's BF16 rounding points: the normalized
// inputs (normalized in FP32), the weights, and the hidden activations after bias and ReLU (through
// FP32); the biases and the output scale stay FP32 values.
struct warm_net_t {
    static constexpr int IN = 48, H = 64, OUT = 32;
    std::vector<double> w1, b1, w2, b2, w3, b3, w4, b4;
    std::vector<float> mean, inv_std, out_scale;
    int nin = 0, nout = 0;
    void set(const float* params, const float* mean, const float* inv_std, const float* out_scale, int nin, int nout);
    // y[o] = out_scale[o] * net((x - mean) * inv_std)[o] for o < nout
    void evaluate(const double* x, double* y) const;
};
double bf16_round(float x);  // nearest BF16 (ties to even), as a double

struct distance_constraint_t final : constraint_t {
    uint32_t b1, b2, v1, v2;
    double d;
    distance_constraint_t(double alpha, double beta, simulation_t const& sim, uint32_t b1, uint32_t v1,
                          uint32_t b2, uint32_t v2);
    void project_positions(simulation_t& sim, double dt) override;
    void residual_sq(simulation_t const& sim, double dt, double* acc) const override;
};

struct collision_constraint_t final : constraint_t {
    uint32_t bi, vi;
    vec3 q, n;
    collision_constraint_t(double alpha, double beta, uint32_t b, uint32_t v, vec3 q_, vec3 n_)
        : constraint_t(alpha, beta), bi(b), vi(v), q(q_), n(n_) {}
    double evaluate(vec3 p) const { return dot(p - q, n); }  // collision_constraint.cpp:43-48
    void project_positions(simulation_t& sim, double dt) override;
    void residual_sq(simulation_t const& sim, double dt, double* acc) const override;
};

// A contact of a surface vertex with a surface triangle (no reference counterpart): the triangle's
// corners (body, vertex), the closest point's barycentric weights w and the normal n, fixed at
// detection; the plane point q = sum_k w_k c_k + h n is taken at the corners' positions copied at
// the start of the contact projections of each iteration (simulation_t::contact_snap); otherwise
// the SDF contact's update (C = (p - q).n, projected when negative) with the triangle point's inverse
// mass sum_k w_k^2 w(c_k) added to the denominator: the vertex moves by its share of the two-sided
// projection, the corners take theirs through their own contacts.
struct body_collision_constraint_t final : constraint_t {
    uint32_t bi, vi, cb[3], cv[3];
    double w[3], h;
    vec3 n;
    body_collision_constraint_t(double alpha, double beta, uint32_t b, uint32_t v, uint32_t const corner_b[3],
                                uint32_t const corner_v[3], double const wts[3], double thickness, vec3 n_)
        : constraint_t(alpha, beta), bi(b), vi(v), h(thickness), n(n_)
    {
        for (int k = 0; k < 3; ++k) {
            cb[k] = corner_b[k];
            cv[k] = corner_v[k];
            w[k] = wts[k];
        }
    }
    vec3 plane_point(simulation_t const& sim) const;
    void project_positions(simulation_t& sim, double dt) override;
    void residual_sq(simulation_t const& sim, double dt, double* acc) const override;
};

// Analytic SDF obstacles. Plane: Eigen::Hyperplane(n, point) semantics used by
// sdf_model_t::from_plane (sdf_model.cpp:52-64): sd = n.p - n.point, grad = n.
// Sphere: sd = |p - c| - r, grad = (p - c)/|p - c| (no reference counterpart).
// Grid: sdf_model_t(Discregrid::CubicLagrangeDiscreteGrid) (sdf_model.cpp:17, evaluate :70-74).
struct sdf_t {
    int kind = 0;  // 0 plane, 1 sphere, 2 grid
    vec3 n;
    double offset = 0;
    vec3 c;
    double r = 0;
    std::shared_ptr<grid_sdf_t> grid;
    vec3 vmin, vmax;  // englobing volume (bvh_model.cpp:50)
    void evaluate(vec3 p, double* sd, vec3* grad) const;
    bool in_volume(vec3 p) const
    {
        return p.x >= vmin.x && p.x <= vmax.x && p.y >= vmin.y && p.y <= vmax.y && p.z >= vmin.z &&
               p.z <= vmax.z;
    }
};

struct contact_t {
    uint32_t body, surface_vertex, vertex, sdf;
    vec3 q, n;
};

// One constraint group = one batch of constraints added together; the colored visiting order
// processes groups in insertion order and, inside a group, colors in increasing order.
// A group may be partitioned into clusters of consecutive constraints; the colored order then
// colors clusters and visits colors, clusters of a color in index order, and the constraints of a
// cluster in order.
struct group_t {
    int kind;                // 0 green, 1 neo-Hookean, 2 distance
    std::size_t first, count;  // range in simulation_t::constraints
    std::vector<uint32_t> cluster_off;  // empty: one cluster per constraint
    std::vector<uint32_t> color;  // per constraint of the group
    uint32_t num_colors = 0;
};

struct simulation_t {
    std::vector<std::vector<particle_t>> particles;
    std::vector<std::vector<uint32_t>> surface;  // per body: surface vertex -> body vertex
    std::vector<std::vector<uint32_t>> triangles;  // per body: boundary triangles, 3 surface vertices each
    std::vector<std::unique_ptr<constraint_t>> constraints;
    std::vector<std::unique_ptr<constraint_t>> collision_constraints;
    // the positions the triangle contacts of an iteration's contact projections read (copied at their
    // start; per body, per vertex)
    std::vector<std::vector<vec3>> contact_snap;
    bool have_body_contacts = false;
    std::vector<sdf_t> sdfs;
    std::vector<group_t> groups;
    std::vector<contact_t> contacts;  // contacts of the current frame, in generation order

    vec3 gravity{0.0, -9.81, 0.0};
    bool gravity_is_force = true;  // timestep.cpp:39-40 semantics
    double collision_compliance = 1e-8, collision_damping = 0.0;  // simulation_parameters.h:26-27
    bool colored_order = false;
    uint32_t recolor_passes = 2;
    int coloring = 0;  // 0: hash priorities (+ recolor passes), 1: index order
    // Warm start of the neo-Hookean multipliers (the device's SBS_WARM_START_* numbering): 1 carries
    // them over from the previous substep, 3 sets them to -C(x~)/alpha~; either way their
    // displacement M^-1 grad C(x~)^T lambda is added before the iterations.
    int warm_start = 0;
    double warm_dt = 0.0;
    double warm_scale = 1.0;    // factor on the warm-start multipliers (deviatoric)
    double warm_scale_h = 1.0;  // factor on the hydrostatic warm-start multipliers
    // warm_start == 2 (learned): per hex cluster of a neo-Hookean group (its 6 tets in order) the
    // network's 39 features at the predicted positions -- alpha~ lambda of the 6 (deviatoric,
    // hydrostatic) pairs, alpha~ times the pairs one substep before, the 12 constraint values, log
    // alpha~ of tet 0 (deviatoric, hydrostatic), log of the number of the group's tets in the
    // cluster's body -- give the change of alpha~ lambda over the substep; lambda_init = lambda +
    // change / alpha~ (then the scales), and the history takes lambda.
    std::shared_ptr<warm_net_t> net;
    void apply_learned_multipliers();
    // The 39 features of cluster c of group g at the current predicted positions (xi) for a substep
    // of dt (test hook of apply_learned_multipliers' features); false unless it is a 6-tet cluster.
    bool cluster_features(group_t const& g, std::size_t c, double dt, std::vector<double> const& per_body,
                          double f[39], double atd[6], double ath[6]) const;
    std::vector<double> tets_per_body(group_t const& g) const;  // the group's tets in each body

    // GS visiting order of the elastic constraints (indices into constraints)
    std::vector<uint32_t> order;
    // colored order only: the order split into stages (one per group color); a stage's elements
    // (stage_elems[stage_off[k] .. stage_off[k+1]), each a range [first, last) of `order`) share no
    // particle, so with threads > 1 they run in parallel and give the sequential order's bits
    std::vector<std::size_t> stage_off;
    std::vector<std::pair<uint32_t, uint32_t>> stage_elems;
    int threads = 1;
    // test hook: with threads == 1, every green projection appends its branches (bit 0 inverted,
    // bits 1-3 singular value i clamped to 0.577, bit 4 the epsilon early-out)
    std::unique_ptr<std::vector<uint8_t>> green_trace;
    bool order_dirty = true;
    void build_order();

    void detect_contacts();
    // Contacts between bodies (no reference counterpart): once per frame, after the SDF contacts,
    // each surface vertex p (global surface order) gets a contact with the nearest surface triangle
    // t (global order; of another body with p not more than r inside its plane, or of its own body
    // without p as a corner and with p more than h outside its plane over its face) whose closest
    // point c has |p - c|^2 < r^2 (the lowest triangle index on ties): plane point c + h n, n the
    // unit direction from c to p, negated when p lies behind t's outward normal (t's normal when p
    // lies on t or t is p's own body's; a body whose surface encloses a negative volume at rest has
    // its triangles flipped);
    // contact index: number of SDFs + t. Defaults (0): cell edge twice, h a
    // tenth of and r equal to the mean surface edge at rest. The broad phase is a uniform grid of the triangles' boxes,
    // queried in the 27 cells around p's cell (body_brute_force: every triangle), which holds every
    // triangle closer than the cell edge.
    bool body_contacts = false, body_brute_force = false;
    double body_cell_request = 0.0, body_thickness_request = 0.0, body_radius_request = 0.0;
    void detect_body_contacts();
    void solve(double dt, std::size_t iterations);  // gauss_seidel_solver_t::solve
    void apply_warm_start();
    void step(double dt, std::size_t iterations, std::size_t substeps);  // timestep_t::step
};

// Greedy coloring of elements (constraints or clusters) in priority order, restating the device
// Jones-Plassmann coloring: pass 0 visits elements by descending hash32(index); pass r>0 visits by
// descending (previous color, hash32(index)). With index_order the single pass visits elements in
// increasing index. Each visited element takes the smallest color not used by an already-colored
// element sharing a vertex with it. Element e touches verts[eoff[e] .. eoff[e+1]).
uint32_t hash32(uint32_t x);
uint32_t greedy_color_csr(uint32_t n, const uint32_t* eoff, const uint64_t* verts, uint32_t passes, bool index_order,
                          uint32_t* color);
uint32_t greedy_color(uint32_t n, uint32_t arity, const uint64_t* verts, uint32_t passes,
                      uint32_t* color);

}  // namespace oracle
