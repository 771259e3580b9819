// C++ host mirror of the reference's simulation / time-step API over the C ABI (sbs_gpu.h).
//
// Same names and argument meanings as the reference, so a caller of
//   sbs::physics::simulation_t       (include/sbs/physics/simulation.h:15-44)
//   sbs::physics::timestep_t         (include/sbs/physics/timestep.h:20-38)
//   xpbd::simulation_parameters_t    (include/sbs/physics/xpbd/simulation_parameters.h:10-28)
// switches by changing the namespace. Bodies, constraints and SDF obstacles are described on the
// host; the first step() commits the scene to the device (coloring and upload) and every step()
// advances one frame there. Errors from the device library are thrown as std::runtime_error (the
// reference's own calls return void).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "sbs_gpu.h"

namespace sbs_amd {

using index_type = std::uint32_t;  // include/sbs/aliases.h:8
using scalar_type = double;        // include/sbs/aliases.h:9
using vec3 = std::array<scalar_type, 3>;

inline void check(int code)
{
    if (code != SBS_OK) throw std::runtime_error(std::string("sbs_gpu: ") + sbs_gpu_last_error());
}

namespace physics {
namespace xpbd {

struct simulation_parameters_t {
    scalar_type compliance = 1e-4;
    scalar_type damping = 0.;
    scalar_type young_modulus = 1e6;
    scalar_type poisson_ratio = 0.30;
    scalar_type hooke_coefficient = 1.;
    scalar_type collision_compliance = 1e-8;
    scalar_type collision_damping = 0.;
};

}  // namespace xpbd

class simulation_t {
  public:
    explicit simulation_t(int device = 0) { check(sbs_gpu_create(device, &h_)); }
    ~simulation_t() { sbs_gpu_destroy(h_); }
    simulation_t(simulation_t const&) = delete;
    simulation_t& operator=(simulation_t const&) = delete;

    // tetrahedral_body_t(simulation, id, geometry): particles at x0, unit mass unless given
    index_type add_body(std::vector<vec3> const& x0, std::vector<scalar_type> const& mass = {})
    {
        index_type id = 0;
        check(sbs_gpu_add_body(h_, x0.empty() ? nullptr : x0[0].data(), mass.empty() ? nullptr : mass.data(),
                               index_type(x0.size()), &id));
        counts_.push_back(x0.size());
        return id;
    }
    // the surface -> tet vertex map the contact handler uses (from_surface_vertex)
    void set_collision_surface(index_type body, std::vector<index_type> const& surface_to_body)
    {
        check(sbs_gpu_set_collision_surface(h_, body, surface_to_body.data(), index_type(surface_to_body.size())));
    }
    // one green_constraint_t per tet, as main.cpp:37-54 adds them
    index_type add_green_constraints(index_type body, std::vector<std::array<index_type, 4>> const& tets,
                                     scalar_type alpha, scalar_type beta, scalar_type young, scalar_type poisson)
    {
        index_type g = 0;
        check(sbs_gpu_add_green_constraints(h_, body, tets.empty() ? nullptr : tets[0].data(), index_type(tets.size()),
                                            alpha, beta, young, poisson, &g));
        return g;
    }
    index_type add_neohookean_constraints(index_type body, std::vector<std::array<index_type, 4>> const& tets,
                                          scalar_type beta, scalar_type young, scalar_type poisson)
    {
        index_type g = 0;
        check(sbs_gpu_add_neohookean_constraints(h_, body, tets.empty() ? nullptr : tets[0].data(),
                                                 index_type(tets.size()), beta, young, poisson, &g));
        return g;
    }
    // distance_constraint_t(alpha, beta, simulation, b1, b2, v1, v2) per entry {b1, v1, b2, v2}
    index_type add_distance_constraints(std::vector<std::array<index_type, 4>> const& b1v1b2v2, scalar_type alpha,
                                        scalar_type beta)
    {
        index_type g = 0;
        check(sbs_gpu_add_distance_constraints(h_, b1v1b2v2.empty() ? nullptr : b1v1b2v2[0].data(),
                                               index_type(b1v1b2v2.size()), alpha, beta, &g));
        return g;
    }
    void set_group_clusters(index_type group, std::vector<index_type> const& offsets)
    {
        check(sbs_gpu_set_group_clusters(h_, group, offsets.data(), index_type(offsets.size() - 1)));
    }
    // sdf_model_t::from_plane(Hyperplane(normal, point), volume)
    void add_plane(vec3 const& normal, vec3 const& point, vec3 const& vmin, vec3 const& vmax)
    {
        check(sbs_gpu_add_plane(h_, normal.data(), point.data(), vmin.data(), vmax.data()));
    }
    void add_sphere(vec3 const& center, scalar_type radius) { check(sbs_gpu_add_sphere(h_, center.data(), radius)); }

    xpbd::simulation_parameters_t& simulation_parameters() { return params_; }
    vec3& gravity() { return gravity_; }
    bool& gravity_is_force() { return gravity_is_force_; }

    // particles()[body][i].x() / .v() after a step
    std::vector<vec3> positions(index_type body) const
    {
        std::vector<vec3> x(counts_.at(body));
        check(sbs_gpu_read_body(h_, body, x.empty() ? nullptr : x[0].data(), nullptr));
        return x;
    }
    std::vector<vec3> velocities(index_type body) const
    {
        std::vector<vec3> v(counts_.at(body));
        check(sbs_gpu_read_body(h_, body, nullptr, v.empty() ? nullptr : v[0].data()));
        return v;
    }

    sbs_gpu* handle() { return h_; }

    void commit_if_needed(std::uint32_t recolor_passes = 2)
    {
        sync_params();
        if (!committed_) {
            check(sbs_gpu_commit(h_, recolor_passes));
            committed_ = true;
        }
    }

  private:
    void sync_params()
    {
        sbs_xpbd_params p{};
        p.compliance = params_.compliance;
        p.damping = params_.damping;
        p.young_modulus = params_.young_modulus;
        p.poisson_ratio = params_.poisson_ratio;
        p.hooke_coefficient = params_.hooke_coefficient;
        p.collision_compliance = params_.collision_compliance;
        p.collision_damping = params_.collision_damping;
        for (int i = 0; i < 3; ++i) p.gravity[i] = gravity_[i];
        p.gravity_is_force = gravity_is_force_ ? 1 : 0;
        check(sbs_gpu_set_params(h_, &p));
    }

    sbs_gpu* h_ = nullptr;
    bool committed_ = false;
    std::vector<std::size_t> counts_;
    xpbd::simulation_parameters_t params_{};
    vec3 gravity_{0.0, -9.81, 0.0};  // timestep.cpp:39
    bool gravity_is_force_ = true;
};

// timestep_t{dt, iterations, substeps}.step(simulation): one frame on the device
class timestep_t {
  public:
    timestep_t() = default;
    timestep_t(scalar_type dt, std::size_t iterations, std::size_t substeps)
        : dt_(dt), iterations_(iterations), substeps_(substeps)
    {
    }
    void step(simulation_t& simulation)
    {
        simulation.commit_if_needed();
        check(sbs_gpu_step(simulation.handle(), dt_, std::uint32_t(iterations_), std::uint32_t(substeps_)));
    }
    scalar_type& dt() { return dt_; }
    std::size_t& iterations() { return iterations_; }
    std::size_t& substeps() { return substeps_; }

  private:
    scalar_type dt_{0.};
    std::size_t iterations_{0u};
    std::size_t substeps_{0u};
};

}  // namespace physics

namespace geometry {

struct tet_mesh {
    std::vector<vec3> positions;
    std::vector<std::array<index_type, 4>> tets;
};

// get_simple_bar_model(width, height, depth)
inline tet_mesh get_simple_bar_model(index_type w, index_type h, index_type d)
{
    tet_mesh m;
    m.positions.resize(std::size_t(w) * h * d);
    if (w > 1 && h > 1 && d > 1) m.tets.resize(std::size_t(w - 1) * (h - 1) * (d - 1) * 5);
    check(sbs_mesh_bar5(w, h, d, m.positions.empty() ? nullptr : m.positions[0].data(),
                        m.tets.empty() ? nullptr : m.tets[0].data()));
    return m;
}

// tetrahedral_mesh_boundary_t::surface_to_tetrahedral_mesh_index_map()
inline std::vector<index_type> boundary_vertices(tet_mesh const& m)
{
    index_type ns = 0, nt = 0;
    check(sbs_mesh_boundary(index_type(m.positions.size()), m.tets[0].data(), index_type(m.tets.size()), nullptr,
                            nullptr, &ns, &nt));
    std::vector<index_type> s2t(ns);
    std::vector<index_type> tris(std::size_t(nt) * 3);
    check(sbs_mesh_boundary(index_type(m.positions.size()), m.tets[0].data(), index_type(m.tets.size()), s2t.data(),
                            tris.data(), &ns, &nt));
    return s2t;
}

}  // namespace geometry
}  // namespace sbs_amd
