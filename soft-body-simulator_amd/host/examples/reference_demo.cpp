// The reference demo scene (main.cpp:22-124) without the viewer: a 4x4x12-vertex bar of green
// constraints, moved by T(-10, 5, -1) R(pi/2 about (0, 1, 0.2)) S(1, 0.8, 2), falling on a floor
// plane, stepped with timestep_t{0.016, 5 iterations, 1 substep} on the device.
// Usage: reference_demo [frames]; prints "frame x y z" of every particle after the last frame.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "sbs_amd/physics.h"

using namespace sbs_amd;

int main(int argc, char** argv)
{
    int const frames = argc > 1 ? std::atoi(argv[1]) : 60;
    physics::simulation_t simulation{};
    geometry::tet_mesh beam = geometry::get_simple_bar_model(4u, 4u, 12u);

    // Affine T * R * S applied to every rest position (tetrahedral_body.cpp:121-132)
    double const angle = 3.14159 / 2.0;
    double ax = 0.0, ay = 1.0, az = 0.2;
    double const n = std::sqrt(ax * ax + ay * ay + az * az);
    ax /= n, ay /= n, az /= n;
    double const c = std::cos(angle), s = std::sin(angle), t = 1.0 - c;
    double const R[3][3] = {{t * ax * ax + c, t * ax * ay - s * az, t * ax * az + s * ay},
                            {t * ax * ay + s * az, t * ay * ay + c, t * ay * az - s * ax},
                            {t * ax * az - s * ay, t * ay * az + s * ax, t * az * az + c}};
    double const scale[3] = {1.0, 0.8, 2.0}, shift[3] = {-10.0, 5.0, -1.0};
    for (auto& p : beam.positions) {
        double const q[3] = {p[0] * scale[0], p[1] * scale[1], p[2] * scale[2]};
        for (int i = 0; i < 3; ++i) p[i] = R[i][0] * q[0] + R[i][1] * q[1] + R[i][2] * q[2] + shift[i];
    }
    index_type const beam_idx = simulation.add_body(beam.positions);
    simulation.set_collision_surface(beam_idx, geometry::boundary_vertices(beam));
    auto const& prm = simulation.simulation_parameters();
    simulation.add_green_constraints(beam_idx, beam.tets, prm.compliance, prm.damping, prm.young_modulus,
                                     prm.poisson_ratio);
    simulation.add_plane({0.0, 1.0, 0.0}, {0.0, 0.0, 0.0}, {-20.0, -5.0, -20.0}, {20.0, 5.0, 20.0});

    physics::timestep_t timestep{0.016, 5u, 1u};
    for (int f = 0; f < frames; ++f) timestep.step(simulation);
    for (auto const& p : simulation.positions(beam_idx)) std::printf("%d %.9g %.9g %.9g\n", frames, p[0], p[1], p[2]);
    return 0;
}
