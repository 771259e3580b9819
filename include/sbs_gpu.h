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
#define SBS_ERR_IO 7     /* a file could not be opened, read or written */
#define SBS_ERR_FORMAT 8 /* a file is not a well-formed PLY file */
/* A frame dropped contacts because a surface vertex had more than 8: the state no longer follows
 * the reference's contact set (which keeps every one). Reported by sbs_gpu_step (for frames the
 * device finished before the call) and by sbs_gpu_synchronize; sticky until the handle is
 * destroyed. */
#define SBS_ERR_CONTACT_OVERFLOW 9

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

/* Boundary triangles of a body (3 indices each into its collision surface, in the order of
 * tetrahedral_mesh_boundary_t::triangles, as sbs_mesh_boundary returns them) for the per-frame
 * surface refresh. */
int sbs_gpu_set_surface_triangles(sbs_gpu* sim, uint32_t body, const uint32_t* triangles, uint32_t n);

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

/* sdf_model_t::from_plane(Hyperplane(normal, point), volume) (src/physics/collision/sdf_model.cpp:52-64).
 * vol_min/vol_max is the SDF's englobing volume; a penetrating vertex outside it gets no contact. */
int sbs_gpu_add_plane(sbs_gpu* sim, const double normal[3], const double point[3], const double vol_min[3],
                      const double vol_max[3]);

/* Analytic sphere obstacle, sd = |p - c| - r (an analytic sdf_model_t: the constructor from an analytic
 * function and a volume, sdf_model.cpp:20-24). */
int sbs_gpu_add_sphere(sbs_gpu* sim, const double center[3], double radius);

/* Grid SDF obstacles: sdf_model_t over a Discregrid::CubicLagrangeDiscreteGrid (sdf_model.cpp:17,
 * evaluate :66-75). The grid splits the box [domain_min, domain_max] into resolution[0] x
 * resolution[1] x resolution[2] cells of 32-node cubic Lagrange elements (nodes at the cell corners
 * and at 1/3 and 2/3 of the cell edges, shared between cells); sbs_sdf_grid_node_count and
 * sbs_sdf_grid_node_positions give the node layout. Outside the domain the field is DBL_MAX.
 * sbs_gpu_add_grid_sdf takes node values (environment_body_t(simulation, id, geometry, sdf_model),
 * environment_body.cpp:80-88) and the englobing volume used by the contact broad phase. */
int sbs_gpu_add_grid_sdf(sbs_gpu* sim, const double domain_min[3], const double domain_max[3],
                         const uint32_t resolution[3], const double* nodes, const double vol_min[3],
                         const double vol_max[3]);
/* environment_body_t(simulation, id, geometry, domain, resolution) (environment_body.cpp:12-78): a
 * triangle mesh obstacle. The domain is extended as the reference does (the box of the domain and
 * the vertices, grown by 1e-3 of its diagonal once per vertex), the mesh's signed distance (closest
 * point, angle-weighted pseudonormal sign) is computed at every node on the device, and the
 * extended domain is the englobing volume. */
int sbs_gpu_add_mesh_obstacle(sbs_gpu* sim, const double* positions, uint32_t num_vertices, const uint32_t* triangles,
                              uint32_t num_triangles, const double domain_min[3], const double domain_max[3],
                              const uint32_t resolution[3]);
/* sdf_model_t::evaluate (sdf_model.cpp:66-75) of obstacle `sdf` at n points on the device: signed
 * distances (n) and gradients (3 n). */
int sbs_gpu_evaluate_sdf(sbs_gpu* sim, uint32_t sdf, const double* points, uint32_t n, double* sd, double* grad);

/* 1: every frame ends with tetrahedral_body_t::update_visual_model on the device
 * (src/physics/tetrahedral_body.cpp:157-165, timestep.cpp:61-66): the surface normals accumulate
 * the face normals (p2 - p1) x (p3 - p1) of their triangles in triangle order and are normalized;
 * as in tetrahedral_mesh_boundary_t::compute_normals (src/physics/tetrahedral_mesh_boundary.cpp:122-145)
 * the accumulator keeps the previous frame's normal. Default 0. */
int sbs_gpu_set_surface_refresh(sbs_gpu* sim, int enable);
/* Surface vertex positions (the committed particle positions) and normals of a body, 3 doubles per
 * surface vertex each; either may be NULL. */
int sbs_gpu_read_surface(sbs_gpu* sim, uint32_t body, double* positions, double* normals);
/* Frame dump: the body's surface mesh (surface positions, its boundary triangles and, with the
 * surface refresh on, the accumulated normals) as a PLY triangle mesh (sbs_ply_write). The
 * reference's surface mesh, tetrahedral_mesh_boundary_t, has no file output; this writes the same
 * geometry record its renderer draws from (src/physics/tetrahedral_body.cpp:157-165). */
int sbs_gpu_write_surface_ply(sbs_gpu* sim, uint32_t body, const char* path, int format);

/* Contact broad phase. The reference traverses a KD-tree of the surface points against every SDF's
 * englobing volume (src/physics/collision/bvh_model.cpp:30-100); here each surface vertex tests
 * either every SDF (brute force) or, with the spatial hash, the SDFs whose volumes overlap its grid
 * cell plus every SDF without a finite volume. Both give the same contacts in the same order (SDF
 * order per vertex). At most 8 contacts are kept per surface vertex; more are counted
 * (sbs_gpu_contact_overflow), and once any has been dropped sbs_gpu_step and sbs_gpu_synchronize
 * fail with SBS_ERR_CONTACT_OVERFLOW. */
#define SBS_CONTACT_HASH_AUTO 0 /* hash from 32 obstacles with finite volume on (default) */
#define SBS_CONTACT_HASH_ON 1
#define SBS_CONTACT_HASH_OFF 2
/* Before commit. cell_size: edge of the hash cells (0: the median largest extent of the volumes). */
int sbs_gpu_set_contact_hash(sbs_gpu* sim, int mode, double cell_size);
/* The built hash: cell edge (0 when not hashed), table size 2^table_bits, the slot ranges start/end
 * (2^table_bits words each) into the obstacle index list entries (cap words, *num_entries in total,
 * increasing inside a slot), and the unhashed obstacles every vertex tests (global, *num_global).
 * Any output may be NULL. Cell (i, j, k) of point p is floor(p / cell) per axis in FP64 (clamped to
 * +-2^30); its slot is ((i*73856093) ^ (j*19349663) ^ (k*83492791)) mod 2^table_bits on 32-bit
 * integers (Teschner et al. 2003). */
int sbs_gpu_read_contact_hash(sbs_gpu* sim, double* cell_size, uint32_t* table_bits, uint32_t* start, uint32_t* end,
                              uint32_t* entries, uint32_t cap, uint32_t* num_entries, uint32_t* global,
                              uint32_t* num_global);
/* Contacts dropped so far because a surface vertex had more than its slots (one per SDF plus two
 * with body contacts on, at most 8); a step reports them as SBS_ERR_CONTACT_OVERFLOW. */
int sbs_gpu_contact_overflow(sbs_gpu* sim, uint64_t* dropped);

/* Contacts between the surfaces of deformable bodies. The reference's collision detection visits
 * every pair of collision models (src/physics/collision/brute_force_cd_system.cpp:17-25) but a pair
 * of surface point clouds yields nothing (bvh_model.cpp:30-34): its only contact is a surface
 * particle against an SDF (include/sbs/physics/collision/contact.h:14), with the contact point and
 * normal fixed for the frame (contact_handler.cpp:14-54). Here, once per frame from the positions at
 * its start, each surface vertex p gets up to two more contacts, each with a surface triangle t
 * closer than `radius` r (|p - c|^2 < r^2 in FP64, c the closest point of t, the lowest triangle
 * index on ties): the nearest triangle of another body (p inside or outside it), then the nearest
 * triangle of p's own surface that does not have p as a corner and that p lies more than h/10
 * outside of (a fold of the surface approaching or resting on it; the flat or convex surface around
 * p never holds it). Inside and outside come from the sign of (p - c).N, N the angle-weighted
 * pseudo-normal of the feature of t holding c (Baerentzen and Aanaes 2005). The contact keeps c's
 * barycentric weights in t and the unit normal n along the line from c to p pointing out of the
 * surface (t's normal when p lies on t); a body whose surface encloses a negative volume at rest
 * has its triangles flipped. Every iteration's contact projection takes c at t's corners as they
 * are then and keeps p on the outer side of the plane through c + h n (h = thickness), acting only
 * when p would cross it, p moving by its share of the two-sided projection (denominator
 * w_p + sum_k b_k^2 w_k + alpha~, b_k the weights); the contacts come after p's SDF contacts
 * (contact index: number of SDFs + triangle index in global surface-triangle order). Broad phase: a
 * spatial hash of cells of edge cell_size rebuilt every frame (Teschner et al. 2003), each vertex
 * querying the cells that the box p +- radius overlaps (at most 8 of the 27 around its own), and a
 * list of the triangles whose box covers more than 8 cells. Before commit; 0 selects the defaults: cell_size twice,
 * thickness a tenth of and radius equal to the mean surface edge at rest; thickness < radius <=
 * 0.999 cell_size. */
int sbs_gpu_set_body_contacts(sbs_gpu* sim, int enable, double cell_size, double thickness, double radius);
/* The committed setup: cell edge, thickness, radius, table size 2^table_bits, and the triangles as
 * the detection uses them (4 words each: 3 surface vertices in outward order, body). Any output may
 * be NULL; all 0 when body contacts are off. */
int sbs_gpu_read_body_contact_setup(sbs_gpu* sim, double* cell_size, double* thickness, double* radius,
                                    uint32_t* table_bits, uint32_t* triangles);
/* Broad-phase candidate pairs at the committed positions: per surface vertex (global surface
 * numbering, increasing) the distinct triangles of the hash slots of the 27 cells around its own,
 * and the triangles whose box covers more than 8 cells (tested by every vertex), that do not have it
 * as a corner, in increasing order, as (vertex, triangle) pairs; with pairs NULL
 * *count gets an upper bound (repeats included), else up to cap pairs are written and *count gets
 * their number. */
int sbs_gpu_read_body_contact_candidates(sbs_gpu* sim, uint32_t* pairs, uint64_t cap, uint64_t* count);

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
 * (64/128, or SBS_CLUSTER_BLOCK_STREAMED: 64 threads, and the neo-Hookean hex-cell kernel loads
 * each tet's material just before projecting it instead of all six up front -- fewer registers,
 * more waves in flight, for launches large enough to be bandwidth bound). Default 256 and 64. */
#define SBS_CLUSTER_BLOCK_STREAMED 65
int sbs_gpu_set_launch_config(sbs_gpu* sim, int block, int cluster_block);
/* Default on (no reference counterpart): when the scene's only constraint group is neo-Hookean
 * with hex-cell clusters and warm-started, the first iteration's contact projections run in the
 * warm-start pass's last launch (the per-particle gather of the displacements), each surface
 * vertex's contacts right after its displacement, in place of their own launch. They run on the
 * same positions in the same order as that launch would (gauss_seidel_solver.cpp:28-31), so the
 * results are bit-identical. Off: the contacts have their own launch. */
int sbs_gpu_set_contact_fusion(sbs_gpu* sim, int enable);

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
 * ids gets 4 uint32 per contact (body, surface vertex, body vertex, contact index: the SDF, or the
 * number of SDFs + the surface triangle), qn 6 doubles (the contact point, for a triangle contact
 * the barycentric weights of the closest point in the triangle; the normal), per (body, surface
 * vertex) in projection order. *count gets the total even if larger than cap. */
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
#define SBS_KERNEL_WARM_START 7 /* features, network inference and the warm-start pass */
#define SBS_KERNEL_COUNT 8

/* 1: bracket every launch with HIP events on the simulation's stream (not with graphs). */
int sbs_gpu_set_timing(sbs_gpu* sim, int enable);
/* Per kernel kind since the last reset: launches, summed launch time (ms), summed work units
 * (constraints, contacts or particles). Waits for the device. */
int sbs_gpu_kernel_stats(sbs_gpu* sim, int kind, uint64_t* launches, double* total_ms, uint64_t* units);
int sbs_gpu_reset_stats(sbs_gpu* sim);

/* ---- learned warm start ---------------------------------------------------------------------------- */

/* No reference counterpart: XPBD (as in the reference) starts every substep from the predicted
 * positions with zero multipliers (constraint.cpp:12-16, timestep.cpp:35-43). A warm start instead
 * begins from multipliers lambda_init and the positions they imply, x = x~ + M^-1 grad C(x~)^T lambda_init
 * (one extra colored pass, every gradient taken at the predicted positions x~), so that positions and
 * multipliers start consistent with the implicit step XPBD approximates. XPBD's iterate depends on
 * its starting point, so a warm-started substep is a different (not a bit-equal) approximation of
 * that step; the residuals measure how close each one gets. */

typedef struct sbs_mlp sbs_mlp;

/* Network 16 -> 128 -> 128 -> 128 (ReLU) -> 32 of which num_inputs (<= 13) inputs and num_outputs
 * (<= 4) outputs are used; BF16 matrix-core arithmetic with FP32 accumulation, FP32 master
 * parameters (He initialization from seed, output layer scaled by 0.1). */
int sbs_mlp_create(int device, uint32_t num_inputs, uint32_t num_outputs, uint64_t seed, sbs_mlp** out);
int sbs_mlp_destroy(sbs_mlp* net);
/* Length of the flat FP32 parameter vector: W1[16][128] b1[128] W2[128][128] b2[128] W3[128][128]
 * b3[128] W4[128][32] b4[32], W_l[i][o] multiplying input unit i into output unit o. */
int sbs_mlp_num_params(uint32_t* count);
int sbs_mlp_get_params(sbs_mlp* net, float* params);
int sbs_mlp_set_params(sbs_mlp* net, const float* params);
/* The network computes y = out_scale * f((x - mean) * inv_std); mean, inv_std have num_inputs
 * entries, out_scale num_outputs. Defaults 0, 1, 1. */
int sbs_mlp_set_normalization(sbs_mlp* net, const float* mean, const float* inv_std, const float* out_scale);
int sbs_mlp_get_normalization(sbs_mlp* net, float* mean, float* inv_std, float* out_scale);
/* y[n][num_outputs] for x[n][num_inputs] through the fused inference kernel. */
int sbs_mlp_forward(sbs_mlp* net, const float* x, uint32_t n, float* y);
/* Mean squared error of f(x') against t / out_scale over the n rows and num_outputs outputs, and
 * its gradient with respect to the parameter vector (through the training kernels). */
int sbs_mlp_loss_grad(sbs_mlp* net, const float* x, const float* t, uint32_t n, double* loss, float* grad);
/* Device-resident training rows (num_inputs features, num_outputs targets each). */
int sbs_mlp_dataset_reserve(sbs_mlp* net, uint64_t rows);
int sbs_mlp_dataset_append(sbs_mlp* net, const float* x, const float* t, uint64_t rows);
int sbs_mlp_dataset_size(sbs_mlp* net, uint64_t* rows);
/* Rows [first, first + rows) back to the host (x: rows * num_inputs, t: rows * num_outputs; either
 * may be NULL). */
int sbs_mlp_dataset_read(sbs_mlp* net, uint64_t first, uint64_t rows, float* x, float* t);
int sbs_mlp_dataset_clear(sbs_mlp* net);
/* mean / inverse standard deviation of every input and RMS of every target over (a strided sample
 * of at most 2^20 rows of) the dataset become the normalization. */
int sbs_mlp_fit_normalization(sbs_mlp* net);
/* `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8) on minibatches of `batch` rows drawn with
 * replacement by a counter hash of (seed, step); *loss (may be NULL) gets the last batch's loss. */
int sbs_mlp_train(sbs_mlp* net, uint32_t steps, uint32_t batch, float lr, uint64_t seed, double* loss);

#define SBS_WARM_START_NONE 0     /* reference behaviour: lambda_init = 0 */
#define SBS_WARM_START_PREVIOUS 1 /* lambda_init = the previous substep's final multipliers */
#define SBS_WARM_START_LEARNED 2  /* lambda_init = previous + net(features), net with 13 inputs, 2 outputs */
#define SBS_WARM_START_EXPLICIT 3 /* lambda_init = -C(x~) / alpha~ (zero XPBD residual at x~) */
/* Warm start of the neo-Hookean groups' multipliers (other groups start from zero). The network's
 * 13 inputs per constraint are, at the predicted positions: lambda_prev (deviatoric, hydrostatic),
 * the pair before it, C_dev, C_hyd, log sum_k w_k |grad_k C_dev|^2, the same for C_hyd,
 * log alpha~_dev, log alpha~_hyd, the multipliers -C / alpha~ of both constraints, and the log of
 * the number of tets of the constraint's body (one network serves bodies of every size in a
 * batched group); its 2 outputs are the change of the pair over the substep. The handle keeps
 * `net` (not owned) until the mode changes. */
int sbs_gpu_set_warm_start(sbs_gpu* sim, int mode, sbs_mlp* net);
/* Factor on the warm-start multipliers (default 1): lambda_init = scale * (the mode's multipliers). */
int sbs_gpu_set_warm_start_scale(sbs_gpu* sim, double scale);
/* Separate factors for the neo-Hookean deviatoric and hydrostatic multipliers (no reference
 * counterpart; sbs_gpu_set_warm_start_scale sets both to one value). */
int sbs_gpu_set_warm_start_scales(sbs_gpu* sim, double deviatoric, double hydrostatic);
/* While stride > 0, every substep of sbs_gpu_step appends one training row (the 13 features at the
 * substep start, the change of the multiplier pair over the substep) for every stride-th
 * constraint of each neo-Hookean group to net's dataset, until its capacity is reached. Graphs
 * are not used while capturing. stride 0 stops. */
int sbs_gpu_capture(sbs_gpu* sim, sbs_mlp* net, uint32_t stride);
/* Device copy of the dynamic state (positions, velocities, normals, multipliers and their warm-start
 * history) and its restoration, in slot 0. No reference counterpart (the reference keeps no copies
 * of its state); the tuner and the bench compare solvers from identical states with them. History
 * buffers created after a snapshot was taken are restored as zeros, their contents at that time. */
int sbs_gpu_snapshot(sbs_gpu* sim);
int sbs_gpu_restore(sbs_gpu* sim);
/* The same in slot `slot` (0 <= slot < SBS_SNAPSHOT_SLOTS); slots are independent. */
#define SBS_SNAPSHOT_SLOTS 2
int sbs_gpu_snapshot_slot(sbs_gpu* sim, int slot);
int sbs_gpu_restore_slot(sbs_gpu* sim, int slot);
/* With a warm start and limit > 0, each sbs_gpu_step copies the state it needs to repeat the frame
 * (its own copy: a snapshot taken with sbs_gpu_snapshot is kept) and, when the FP64 residual norm
 * sqrt(sum over neo-Hookean groups of sum h^2) after the frame exceeds limit or is not a number (a
 * frame that diverged), restores
 * it and repeats the frame without warm start (the plain solver as fallback) with `iterations`
 * iterations per substep (0: the step's own count). *fallbacks counts the repeated frames. */
int sbs_gpu_set_fallback(sbs_gpu* sim, double limit, uint32_t iterations);
int sbs_gpu_fallback_count(sbs_gpu* sim, uint64_t* fallbacks);

/* ---- device primitives of the contact hash (host arrays in and out) ----------------------------------- */

/* out[i] = in[0] + ... + in[i-1] for i <= n (n + 1 words out). */
int sbs_util_exclusive_scan(int device, const uint32_t* in, uint64_t n, uint32_t* out);
/* Stable sort of the pairs by the low `bits` bits of the keys (LSD radix sort, 8 bits per pass). */
int sbs_util_sort_pairs(int device, const uint32_t* keys, const uint32_t* vals, uint64_t n, int bits,
                        uint32_t* out_keys, uint32_t* out_vals);
/* The same sort without any allocation; no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops): keys and vals are
 * device pointers of n words, sorted in place on `stream` (a hipStream_t; NULL: the null stream) and
 * returned before the sort runs (the result is there when the stream reaches it). At most
 * SBS_SORT_NOALLOC_MAX_PAIRS pairs; more are rejected. Its work buffers are one device array per GPU
 * that the library holds until the process ends (about 10.5 MB), shared by every call on that GPU:
 * calls that overlap on one GPU race, so issue them on one stream. */
#define SBS_SORT_NOALLOC_MAX_PAIRS (1u << 20)
int sbs_util_sort_pairs_noalloc(int device, uint32_t* keys, uint32_t* vals, uint64_t n, int bits, void* stream);

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

/* The same boundary surface computed on GPU `device` (radix sorts and prefix sums over the 4T
 * faces; at most 2^21 vertices). Outputs are host arrays sized by a first call with NULL outputs,
 * or at most 3 * 4 * num_tets triangle indices and num_vertices surface vertices. */
int sbs_gpu_extract_boundary(int device, uint32_t num_vertices, const uint32_t* tets, uint32_t num_tets,
                             uint32_t* surface_to_tet, uint32_t* triangles, uint32_t* num_surface_vertices,
                             uint32_t* num_triangles);

/* ---------------------------------------------------------------- PLY files
 * The geometry record of src/io/ply.cpp (common::geometry_t, include/sbs/common/geometry.h:10-30):
 * 3 position floats per vertex, 3 or 4 indices per triangle or tet, optional per-vertex normals
 * (3 floats), uvs (2 floats) and colors (3 bytes). Arrays a read returns are owned by the library
 * and freed by sbs_geometry_release. */
#define SBS_GEOMETRY_UNSET (-1) /* the file has neither a face nor a tet element */
#define SBS_GEOMETRY_TRIANGLE 0
#define SBS_GEOMETRY_TETRAHEDRON 1
#define SBS_PLY_ASCII 0 /* io::ply_format_t, include/sbs/io/ply.h:25 */
#define SBS_PLY_BINARY_LITTLE_ENDIAN 1
#define SBS_PLY_BINARY_BIG_ENDIAN 2

typedef struct sbs_geometry {
    float* positions;
    uint64_t num_positions; /* floats */
    int32_t* indices;
    uint64_t num_indices;
    float* normals;
    uint64_t num_normals;
    float* uvs;
    uint64_t num_uvs;
    uint8_t* colors;
    uint64_t num_colors;
    int32_t geometry_type; /* SBS_GEOMETRY_* */
} sbs_geometry;

/* io::read_ply(path) (src/io/ply.cpp:303-319, 321-775): ASCII, little- and big-endian binary.
 * The path must end in ".ply". Fails with SBS_ERR_IO or SBS_ERR_FORMAT where read_ply returns no
 * value. */
int sbs_ply_read(const char* path, sbs_geometry* out);
/* io::read_ply(std::istream&) over a file image in memory. */
int sbs_ply_read_memory(const void* data, uint64_t size, sbs_geometry* out);
/* Frees the arrays of a geometry returned by a read and clears the record. */
int sbs_geometry_release(sbs_geometry* geometry);
/* io::write_ply(path, geometry, format) (src/io/ply.cpp:25-39, 41-296). */
int sbs_ply_write(const char* path, const sbs_geometry* geometry, int format);
/* io::write_ply(std::ostream&, ...) into a caller buffer: *size gets the file size; with a NULL
 * buffer nothing is written. */
int sbs_ply_write_memory(const sbs_geometry* geometry, int format, void* buffer, uint64_t capacity, uint64_t* size);

/* Grid SDF helpers (no simulation handle). Node count and node positions of a cubic Lagrange grid
 * (Discregrid's node numbering: corners, then x, y and z edge nodes), environment_body_t's extended
 * domain (environment_body.cpp:52-64), and the node values of a triangle mesh's signed distance
 * computed on GPU `device` (what CubicLagrangeDiscreteGrid::addFunction stores for
 * MeshDistance::signedDistanceCached, environment_body.cpp:66-73). */
int sbs_sdf_grid_node_count(const uint32_t resolution[3], uint64_t* count);
int sbs_sdf_grid_node_positions(const double domain_min[3], const double domain_max[3], const uint32_t resolution[3],
                                double* positions);
int sbs_sdf_environment_domain(const double* positions, uint32_t num_vertices, const double domain_min[3],
                               const double domain_max[3], double out_min[3], double out_max[3]);
int sbs_sdf_grid_from_mesh(int device, const double* positions, uint32_t num_vertices, const uint32_t* triangles,
                           uint32_t num_triangles, const double domain_min[3], const double domain_max[3],
                           const uint32_t resolution[3], double* nodes);

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
