/*
 * sbs_gpu.h -- C ABI of the MI355X (gfx950) XPBD time step.
 *
 * Drop-in for the reference's CPU time step: a simulation (bodies of particles, elastic and
 * distance constraints, SDF obstacles) is described through these calls, and sbs_gpu_step()
 * advances it by one frame with the semantics of timestep_t::step (src/physics/timestep.cpp:20-70):
 * contacts are detected once per frame from the frame-start positions, then each substep predicts
 * positions, runs `iterations` Gauss-Seidel sweeps (contacts first, then the elastic constraints,
 * multipliers reset every substep), and commits positions and velocities.
 *
 * The Gauss-Seidel sweep visits each constraint group (one per add_*_constraints call, in call
 * order) color by color; constraints of one color share no particle, so each color is one
 * data-parallel kernel and the result equals a sequential sweep in that order.
 *
 * Conventions: plain pointers and sizes, indices are uint32 (index_type, include/sbs/aliases.h:8),
 * positions are 3 doubles per particle (scalar_type, aliases.h:9; the device computes in FP32 with
 * FP64 reductions). Every call returns SBS_OK (0) or an error code; sbs_gpu_last_error() gives the
 * message of the last failure on the calling thread. Nothing throws across this boundary. A handle
 * is used by one host thread at a time; calls on it are stream ordered and sbs_gpu_step() returns
 * before the device finishes (sbs_gpu_synchronize() or any read waits).
 */
#ifndef SBS_GPU_H
#define SBS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBS_OK 0
#define SBS_ERR_INVALID_ARGUMENT 1
#define SBS_ERR_STATE 2
#define SBS_ERR_HIP 3
#define SBS_ERR_OUT_OF_MEMORY 4
#define SBS_ERR_NO_DEVICE 5
#define SBS_ERR_INTERNAL 6

typedef struct sbs_gpu sbs_gpu;

/* Mirrors xpbd::simulation_parameters_t (include/sbs/physics/xpbd/simulation_parameters.h:10-28)
 * plus the gravity the reference hard-codes at timestep.cpp:39. */
typedef struct sbs_xpbd_params {
    double compliance;           /* green constraint alpha default (1e-4) */
    double damping;              /* green constraint beta default (0) */
    double young_modulus;        /* 1e6 */
    double poisson_ratio;        /* 0.30 */
    double hooke_coefficient;    /* 1 (unused, kept for layout parity) */
    double collision_compliance; /* 1e-8 */
    double collision_damping;    /* 0 (accepted, unused by contact projection, as in the reference) */
    double gravity[3];           /* (0, -9.81, 0) */
    /* 1: the reference's semantics, gravity is a force: f += g, v += f w dt (timestep.cpp:39-40,
     *    particle.cpp:50-53), so acceleration is g/m. 0: gravity is an acceleration applied to
     *    every particle with w > 0 (lumped-mass scenes). */
    int32_t gravity_is_force;
    int32_t reserved;
} sbs_xpbd_params;

/* Fills *p with the reference's defaults (gravity_is_force = 1). */
int sbs_gpu_default_params(sbs_xpbd_params* p);

/* ---- lifecycle ------------------------------------------------------------------------------ */

/* Replaces the construction of simulation_t (include/sbs/physics/simulation.h:15-44) and
 * timestep_t (include/sbs/physics/timestep.h:20-38) on `device`. */
int sbs_gpu_create(int device, sbs_gpu** out);
int sbs_gpu_destroy(sbs_gpu* sim);
const char* sbs_gpu_last_error(void);

/* simulation_t::simulation_parameters() (simulation.h) -- may be called at any time. */
int sbs_gpu_set_params(sbs_gpu* sim, const sbs_xpbd_params* params);

/* ---- scene description (before sbs_gpu_commit) ---------------------------------------------- */

/* A body of n particles at rest positions x0 (3n doubles); replaces tetrahedral_body_t's
 * simulation.add_particle loop (src/physics/tetrahedral_body.cpp:54-64). mass may be NULL (unit
 * mass, particle.cpp:6-9); a particle with mass 0 is fixed (particle.cpp:45-49). */
int sbs_gpu_add_body(sbs_gpu* sim, const double* x0, const double* mass, uint32_t n, uint32_t* body_id);

/* Collision candidates of a body: surface vertex -> body vertex, the map
 * tetrahedral_mesh_boundary_t::from_surface_vertex (src/physics/tetrahedral_mesh_boundary.cpp:55-58)
 * that contact_handler_t uses (src/physics/xpbd/contact_handler.cpp:39-40). */
int sbs_gpu_set_collision_surface(sbs_gpu* sim, uint32_t body, const uint32_t* surface_to_body, uint32_t n);

/* n green_constraint_t on body-local tets (4 indices each), as main.cpp:37-54 adds them
 * (src/physics/xpbd/green_constraint.cpp:11-47). */
int sbs_gpu_add_green_constraints(sbs_gpu* sim, uint32_t body, const uint32_t* tets, uint32_t n, double alpha,
                                  double beta, double young_modulus, double poisson_ratio, uint32_t* group_id);

/* n stable neo-Hookean tet constraints (Macklin & Mueller 2021): deviatoric ||F|| with compliance
 * 1/(mu|V0|) then hydrostatic det F - (1 + mu/lambda) with compliance 1/(lambda|V0|). No reference
 * counterpart; added through the same simulation_t::add_constraint slot. */
int sbs_gpu_add_neohookean_constraints(sbs_gpu* sim, uint32_t body, const uint32_t* tets, uint32_t n, double beta,
                                       double young_modulus, double poisson_ratio, uint32_t* group_id);

/* Batched forms: one group of n tet constraints spanning many bodies -- tet c lies in body
 * bodies[c] with body-local indices tets[4c..4c+3] and its own Young modulus and Poisson ratio.
 * Bodies share no particle, so the group's colors (and launches) cover all of them at once; the
 * result equals adding one group per body. */
int sbs_gpu_add_green_constraints_batched(sbs_gpu* sim, const uint32_t* bodies, const uint32_t* tets, uint32_t n,
                                          double alpha, double beta, const double* young_modulus,
                                          const double* poisson_ratio, uint32_t* group_id);
int sbs_gpu_add_neohookean_constraints_batched(sbs_gpu* sim, const uint32_t* bodies, const uint32_t* tets, uint32_t n,
                                               double beta, const double* young_modulus, const double* poisson_ratio,
                                               uint32_t* group_id);

/* n distance_constraint_t, 4 indices each (b1, v1, b2, v2)
 * (src/physics/xpbd/distance_constraint.cpp:8-22). */
int sbs_gpu_add_distance_constraints(sbs_gpu* sim, const uint32_t* b1_v1_b2_v2, uint32_t n, double alpha,
                                     double beta, uint32_t* group_id);

/* sdf_model_t::from_plane(Hyperplane(normal, point), volume) (src/physics/collision/sdf_model.cpp:181-193).
 * vol_min/vol_max is the SDF's englobing volume; a penetrating vertex outside it gets no contact. */
int sbs_gpu_add_plane(sbs_gpu* sim, const double normal[3], const double point[3], const double vol_min[3],
                      const double vol_max[3]);

/* Analytic sphere obstacle, sd = |p - c| - r (an analytic sdf_model_t, sdf_model.cpp:149-153). */
int sbs_gpu_add_sphere(sbs_gpu* sim, const double center[3], double radius);

/* Partitions a tet group's constraints (in the order they were added) into clusters of
 * consecutive constraints: cluster i is [cluster_offsets[i], cluster_offsets[i+1]), with
 * cluster_offsets[0] = 0 and cluster_offsets[num_clusters] = the group's constraint count. One
 * device thread runs a cluster's constraints one after the other; clusters (not constraints) are
 * colored, so the Gauss-Seidel sweep visits colors, then clusters, then the cluster's constraints
 * in order. A cluster may touch at most 8 particles (e.g. the 6 tets of a hex cell). No reference
 * counterpart: it only reorders the reference's sequential sweep (gauss_seidel_solver.cpp:28-35). */
int sbs_gpu_set_group_clusters(sbs_gpu* sim, uint32_t group, const uint32_t* cluster_offsets, uint32_t num_clusters);

/* Coloring order used by sbs_gpu_commit: */
#define SBS_COLORING_HASH 0  /* device Jones-Plassmann, priorities hash32(index) (default) */
#define SBS_COLORING_INDEX 1 /* sequential first-fit greedy in index order (host, at commit) */
int sbs_gpu_set_coloring(sbs_gpu* sim, int order);

/* Threads per block of the per-constraint kernels (64/128/256) and of the cluster kernels
 * (64/128). Default 256 and 64. */
int sbs_gpu_set_launch_config(sbs_gpu* sim, int block, int cluster_block);

/* Colors every constraint group and uploads the scene. recolor_passes: extra Jones-Plassmann
 * passes of SBS_COLORING_HASH (0 = hash order only). After this, no more scene additions. */
int sbs_gpu_commit(sbs_gpu* sim, uint32_t recolor_passes);

/* ---- time stepping ---------------------------------------------------------------------------- */

/* timestep_t::step (src/physics/timestep.cpp:20-70) with timestep_t's dt/iterations/substeps. */
int sbs_gpu_step(sbs_gpu* sim, double dt, uint32_t iterations, uint32_t substeps);
int sbs_gpu_synchronize(sbs_gpu* sim);

/* 1: capture each frame's launches into a hipGraph and replay it (default 0). */
int sbs_gpu_use_graphs(sbs_gpu* sim, int enable);

/* ---- state access ------------------------------------------------------------------------------ */

/* Positions x and velocities v of a body (3n doubles each, either may be NULL). */
int sbs_gpu_read_body(sbs_gpu* sim, uint32_t body, double* x, double* v);
/* Overwrites x (and the substep start position) and/or v of a body; either may be NULL. */
int sbs_gpu_write_body(sbs_gpu* sim, uint32_t body, const double* x, const double* v);

/* Number of colors of a group, and optionally the color of each of its constraints (in the order
 * they were added). */
int sbs_gpu_group_colors(sbs_gpu* sim, uint32_t group, uint32_t* num_colors, uint32_t* color_of_constraint);

/* Detects contacts from the current positions (what the next step() will use) and reads them:
 * ids gets 4 uint32 per contact (body, surface vertex, body vertex, sdf index), qn 6 doubles
 * (contact point, normal), ordered by (body, surface vertex, sdf). *count gets the total even if
 * larger than cap. */
int sbs_gpu_read_contacts(sbs_gpu* sim, uint32_t* ids, double* qn, uint32_t cap, uint32_t* count);

/* Squared XPBD residual sum (h = C(x) + alpha~ lambda), in FP64, per color of a group, for the
 * state after the last step (dt is that step's substep length). per_color has num_colors entries. */
int sbs_gpu_residual(sbs_gpu* sim, uint32_t group, double dt, double* per_color, uint32_t cap);

/* FP64 totals: kinetic, gravitational potential, elastic energy, momentum x, y, z. */
int sbs_gpu_energy(sbs_gpu* sim, double out[6]);

/* ---- measurement --------------------------------------------------------------------------------- */

#define SBS_KERNEL_PREDICT 0
#define SBS_KERNEL_CONTACT 1
#define SBS_KERNEL_GREEN 2
#define SBS_KERNEL_NEOHOOKEAN 3
#define SBS_KERNEL_DISTANCE 4
#define SBS_KERNEL_COMMIT 5
#define SBS_KERNEL_DETECT 6
#define SBS_KERNEL_COUNT 7

/* 1: bracket every launch with HIP events on the simulation's stream (not with graphs). */
int sbs_gpu_set_timing(sbs_gpu* sim, int enable);
/* Per kernel kind since the last reset: launches, summed launch time (ms), summed work units
 * (constraints, contacts or particles). Waits for the device. */
int sbs_gpu_kernel_stats(sbs_gpu* sim, int kind, uint64_t* launches, double* total_ms, uint64_t* units);
int sbs_gpu_reset_stats(sbs_gpu* sim);

/* ---- host mesh utilities (no device) ---------------------------------------------------------------- */

/* get_simple_bar_model (src/geometry/get_simple_bar_model.cpp:6-122): w*h*d lattice, 5 tets per
 * cell. pos gets 3*w*h*d doubles, tets 4*5*(w-1)(h-1)(d-1) indices; either may be NULL. */
int sbs_mesh_bar5(uint32_t w, uint32_t h, uint32_t d, double* pos, uint32_t* tets);

/* nx*ny*nz cells of edge `edge` from `origin`, 6 tets per cell sharing the cell diagonal. */
int sbs_mesh_box6(uint32_t nx, uint32_t ny, uint32_t nz, double edge, const double origin[3], double* pos,
                  uint32_t* tets);

/* Boundary surface of a tet mesh (src/physics/tetrahedral_mesh_boundary.cpp:65-120 with the face
 * numbering of src/physics/topology.cpp:335-342,904-956). Call with NULL outputs to get the sizes. */
int sbs_mesh_boundary(uint32_t num_vertices, const uint32_t* tets, uint32_t num_tets, uint32_t* surface_to_tet,
                      uint32_t* triangles, uint32_t* num_surface_vertices, uint32_t* num_triangles);

/* Ellipsoid of radii (rx, ry, rz) centred on the origin: the cells of edge `edge` whose centres lie
 * inside it, each split like sbs_mesh_box6; vertices are the used lattice corners in lattice order.
 * Cells (6 consecutive tets each) come grouped by lattice parity class, so that SBS_COLORING_INDEX
 * colors their clusters with 8 balanced colors. Call with NULL pos/tets to get the sizes. */
int sbs_mesh_ellipsoid6(double rx, double ry, double rz, double edge, double* pos, uint32_t* tets,
                        uint32_t* num_vertices, uint32_t* num_tets);

/* Lumped mass: every tet adds density*|V|/4 to each of its vertices. */
int sbs_mesh_lumped_mass(uint32_t num_vertices, const double* pos, const uint32_t* tets, uint32_t num_tets,
                         double density, double* mass);

#ifdef __cplusplus
}
#endif

#endif /* SBS_GPU_H */
