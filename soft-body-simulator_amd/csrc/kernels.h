// Launch interface of the gfx950 kernels (implemented in kernels.hip). All launches go to the
// given stream; none allocates memory.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "mlp.h"
#include "sbs_gpu.h"

namespace sbs {

// Particle positions: x, y, z and the inverse mass in w, in FP64 (see kernels.hip).
using pos4 = double4;

// Analytic SDF obstacle in device memory (FP64, like the contact positions).
struct sdf_dev {
    int kind;         // 0 plane, 1 sphere, 2 grid
    uint32_t res[3];  // grid: cells per axis
    double a[4];      // plane: n.xyz, offset | sphere: c.xyz, r
    double vmin[3], vmax[3];
    // grid: domain, cell size and its inverse, node values (cubic Lagrange grid layout)
    double gmin[3], gmax[3], cell[3], inv_cell[3];
    const double* nodes;
};

// Arrays of one tet-constraint group in color order. Material layout (3 float4 per tet):
//   m0 = (r0.x, r0.y, r0.z, r1.x), m1 = (r1.y, r1.z, r2.x, r2.y), m2 = (r2.z, ...)
// with r_k the rows of DmInv. green:      m2 = (r2.z, V0 signed, mu, lambda)
//                             neo-Hookean: m2 = (r2.z, 1/(mu|V0|), 1/(lambda|V0|), 1 + mu/lambda)
struct tet_group_dev {
    const int4* idx;
    const float4* m0;
    const float4* m1;
    const float4* m2;
    float* lambda;  // green: 1 per tet, neo-Hookean: 2 per tet (deviatoric, hydrostatic)
};

// Clusters of a tet group: cluster c runs constraints [off[c], off[c+1]) one after the other in
// one thread; its (at most 8) particles are v[c].ids (-1 = unused slot) and every constraint names
// its 4 particles by local slot (loc: 4 x 8 bits). Constraint arrays are in execution order.
struct cluster_dev {
    const uint32_t* off;
    const int4* v0;  // slots 0..3
    const int4* v1;  // slots 4..7
    const uint32_t* loc;
};

constexpr int kClusterSlots = 8;

// The clusters of one color split over the 8 XCDs. Workgroups are dispatched to the XCDs round-robin
// (workgroup b runs on XCD b % 8), so the hex kernels run the color's clusters [off[x], off[x+1])
// in workgroups b with b % 8 == x. The host orders every color's clusters by spatial region (8
// regions of the group's clusters in Morton order), so an XCD keeps working on the same particles
// from color to color and from launch to launch, and finds them in its own L2.
struct xcd_split {
    uint32_t off[9];
};
constexpr int kXcds = 8;

struct distance_group_dev {
    const uint2* idx;
    const float* rest;
    float* lambda;
};

struct contact_dev {
    const uint32_t* surf;  // surface vertex -> global particle
    uint32_t S;            // surface vertex count
    uint32_t slots;        // contact slots per surface vertex
    pos4* q;               // [slot * S + s]: contact point, w = sdf index
    float4* n;             // normal
    float* lambda;
    uint32_t* count;       // contacts per surface vertex
    uint32_t* overflow;    // contacts dropped because a vertex had more than `slots`
    // contacts with surface triangles (index >= nsdf): q.xyz holds the barycentric weights of the
    // triangle's closest point, and the plane point is taken from the surface positions copied at the
    // start of the contact launch (snap[s], per surface vertex) plus body_h along n
    const pos4* snap;
    const uint4* btris;
    uint32_t nsdf;
    double body_h;
};

void launch_predict(hipStream_t s, pos4* xi, const pos4* xn, float4* v, uint32_t n, float gx, float gy,
                    float gz, int gravity_is_force, float dt, int block);
void launch_commit(hipStream_t s, const pos4* xi, pos4* xn, float4* v, uint32_t n, float dt, int block);
// Particle rows: a permutation of the particles in workgroups of kParticleRowBlock rows, workgroup
// b holding particles of XCD region b mod 8 (padding rows 0xffffffff), so that a per-particle pass
// runs on the XCD whose L2 holds the region's particles. With prow (nprows rows), the commit and
// prediction run over the rows (kParticleRowBlock threads per workgroup), else over [0, n) with
// `block` threads.
constexpr int kParticleRowBlock = 256;
void launch_commit_predict(hipStream_t s, pos4* xi, pos4* xn, float4* v, uint32_t n, float gx, float gy,
                           float gz, int gravity_is_force, float dt, int block, const uint32_t* prow = nullptr,
                           uint32_t nprows = 0);

void launch_green(hipStream_t s, tet_group_dev g, pos4* xi, const pos4* xn, uint32_t begin, uint32_t count,
                  float alpha_tilde, float gamma, int first, int block);
// Neo-Hookean launches take a mode: kSolveFirst (first iteration, multipliers start at 0),
// kSolveNext (multipliers as stored), or a warm-start pass (no solve): kApplyWarmStart adds
// M^-1 grad C^T lambda of the stored multipliers to xi, kApplyExplicit first sets the multipliers
// to -C / alpha~ and stores them; both take every gradient at the predicted positions, passed as
// `xn` (a copy of xi made before the pass), so the pass is a sum independent of visiting order.
// In those modes beta_dt carries the factor the deviatoric multipliers are scaled by (and stored),
// hyd_scale the hydrostatic multipliers' factor (unused by the solve modes).
constexpr int kSolveNext = 0, kSolveFirst = 1, kApplyWarmStart = 2, kApplyExplicit = 3;
void launch_neohookean(hipStream_t s, tet_group_dev g, pos4* xi, const pos4* xn, uint32_t begin,
                       uint32_t count, float inv_dt2, float beta_dt, int mode, int block, float hyd_scale = 1.0f);
// One thread per cluster of clusters [begin, begin+count).
void launch_green_clusters(hipStream_t s, tet_group_dev g, cluster_dev cl, pos4* xi, const pos4* xn,
                           uint32_t begin, uint32_t count, float alpha_tilde, float gamma, int first, int block);
void launch_neohookean_clusters(hipStream_t s, tet_group_dev g, cluster_dev cl, pos4* xi, const pos4* xn,
                                uint32_t begin, uint32_t count, float inv_dt2, float beta_dt, int mode, int block,
                                float hyd_scale = 1.0f);

// Clusters that are the 6 tets of one hex cell of the box split, in its slot pattern
// {0123, 0413, 0253, 0563, 0743, 0673}: particles held in registers. The color's constraints
// start at tbase, slot-major: tet k of the color's i-th cluster is tbase + k * count + i.
void launch_green_hex(hipStream_t s, tet_group_dev g, cluster_dev cl, pos4* xi, const pos4* xn, uint32_t begin,
                      uint32_t count, uint32_t tbase, xcd_split const& sp, float alpha_tilde, float gamma, int first,
                      int block);
// Device copies of up to kMaxSpans arrays in one launch (sizes in bytes, multiples of 4; 16-byte
// aligned starts).
constexpr int kMaxSpans = 8;
struct copy_spans {
    void* dst[kMaxSpans];
    const void* src[kMaxSpans];
    uint64_t bytes[kMaxSpans];
    int n;
};
void launch_copy_spans(hipStream_t s, copy_spans const& spans);
// Residual partial sums (as launch_neohookean_residual's) of a hex-cluster neo-Hookean group, one
// thread per row of cinfo (block sums of its cluster's six tets' h^2; partial gets
// ceil(nrows / 64) values). cinfo as at launch_hex_warm_start.
void launch_neohookean_hex_residual(hipStream_t s, tet_group_dev g, cluster_dev cl, const pos4* x,
                                    const uint4* cinfo, uint32_t nrows, float inv_dt2, double* partial);
// out[0] = the sum of partial[0, n) by one block in a fixed order (out may be pinned host memory).
void launch_sum_partials(hipStream_t s, const double* partial, uint32_t n, double* out);
// *dst = *src by one thread (dst may be pinned host memory): publishes a device counter to the host
// without a copy node in a frame graph.
void launch_publish_u32(hipStream_t s, const uint32_t* src, uint32_t* dst);

void launch_neohookean_hex(hipStream_t s, tet_group_dev g, cluster_dev cl, pos4* xi, const pos4* xn,
                           uint32_t begin, uint32_t count, uint32_t tbase, xcd_split const& sp, float inv_dt2,
                           float beta_dt, int mode, int block, float hyd_scale = 1.0f);
// cluster_block value selecting the hex kernel that loads each tet's material just before
// projecting it (fewer registers, more waves) instead of prefetching all six.
constexpr int kHexStreamed = SBS_CLUSTER_BLOCK_STREAMED;
constexpr uint32_t kHexPatternLoc[6] = {0x03020100u, 0x03010400u, 0x03050200u, 0x03060500u, 0x03040700u, 0x03070600u};

void launch_distance(hipStream_t s, distance_group_dev g, pos4* xi, const pos4* xn, uint32_t begin,
                     uint32_t count, float alpha_tilde, float gamma, int first, int block);

void launch_detect(hipStream_t s, contact_dev c, const pos4* x, const sdf_dev* sdfs, uint32_t nsdf, int block);
void launch_contact(hipStream_t s, contact_dev c, pos4* xi, float alpha_tilde, int first, int block);
// snap[s] = xi[surf[s]] for s < S: the surface positions the triangle contacts of a contact launch
// read.
void launch_snapshot_surface(hipStream_t s, const uint32_t* surf, const pos4* xi, pos4* snap, uint32_t S);

// Residual partial sums: one double per block of `block` threads over [begin, begin+count).
void launch_green_residual(hipStream_t s, tet_group_dev g, const pos4* x, uint32_t begin, uint32_t count,
                           float alpha_tilde, double* partial, int block);
void launch_neohookean_residual(hipStream_t s, tet_group_dev g, const pos4* x, uint32_t begin, uint32_t count,
                                float inv_dt2, double* partial, int block);
void launch_distance_residual(hipStream_t s, distance_group_dev g, const pos4* x, uint32_t begin,
                              uint32_t count, float alpha_tilde, double* partial, int block);

// Energy partials: per block 6 doubles (kinetic, potential, elastic, px, py, pz).
void launch_particle_energy(hipStream_t s, const pos4* x, const float4* v, const float* mass, uint32_t n,
                            float gx, float gy, float gz, int gravity_is_force, double* partial, int block);
void launch_green_energy(hipStream_t s, tet_group_dev g, const pos4* x, uint32_t count, double* partial,
                         int block);
void launch_neohookean_energy(hipStream_t s, tet_group_dev g, const pos4* x, uint32_t count, double* partial,
                              int block);

// One Jones-Plassmann round over n elements; element t touches vertices
// verts[elem_off[t] .. elem_off[t+1]) (elem_off == nullptr: `arity` vertices each).
void launch_jp_round(hipStream_t s, uint32_t n, uint32_t arity, const uint32_t* elem_off, const uint32_t* verts,
                     const uint32_t* inc_off,
                     const uint32_t* inc, const uint64_t* key, const uint32_t* color_in, uint32_t* color_out,
                     uint32_t* remaining, uint32_t* overflow, int block);
void launch_jp_keys(hipStream_t s, uint32_t n, const uint32_t* prev, uint64_t* key, int block);

void launch_copy_pos(hipStream_t s, const pos4* a, pos4* b, uint32_t n);

// The warm-start network's side of a hex group's warm-start pass: per tet (execution order) the
// multiplier history hist (the pair one substep before the stored one) and the body-size feature;
// raw [nrows][40] feature rows (capture); the network's packed BF16 weights, biases and
// normalization (learned mode).
struct warm_net_dev {
    float2* hist;
    const float* body_size;
    float* raw;
    const mlp::bf16* pack;
    const float* bias;
    const mlp::norm_dev* norm;
};
// net_mode of launch_hex_warm_start: the pass alone; also the features of every cluster into raw
// (training capture); or the network's predicted change added to the multipliers first.
constexpr int kWarmNetNone = 0, kWarmNetFeatures = 1, kWarmNetLearned = 2;

// Warm-start pass of a hex-cluster neo-Hookean group in two launches instead of one per color:
// every cluster's slot displacements at the predicted positions xt into disp ([8][nrows]: slot k of
// the cluster in row r at k * nrows + r), then for each particle p < n, xi[p] += the FP64 sum over
// csr[csr_off[p] .. csr_off[p+1]) (indices into disp, in cluster execution order) -- the same
// additions in the same order as the per-color pass; with csr_off null, csr is a dense [n][8] table
// of those indices (unused: 0xffffffff). nrows is a multiple of 64; per 64-row workgroup b,
// cinfo[2b] = (its first cluster's index in its color, the color's cluster count, the color's first
// constraint, the first cluster) and cinfo[2b + 1].x = its rows in use, t < that running cluster
// cinfo[2b].w + t.
// With surface_of (dense list only; particle -> surface vertex or -1) the gather also runs the first
// iteration's contact projections (compliance c_at) of the surface vertices, in place of that
// iteration's contact launch. With prow (dense list only), the gather runs over the particle rows.
// With net_mode kWarmNetFeatures or kWarmNetLearned, the cluster kernel forms the cluster's 39
// network features at the predicted positions: alpha~ lambda_prev of its 6 tets' (deviatoric,
// hydrostatic) pairs (12), alpha~ lambda_prev2 (12), the 12 constraint values C at xt, log alpha~
// of tet 0 (deviatoric, hydrostatic) and the body-size feature, where lambda_prev is the stored
// multiplier pair (the previous substep's final) and lambda_prev2 net.hist, which receives
// lambda_prev. kWarmNetFeatures writes them to net.raw[row]; kWarmNetLearned normalizes them, runs
// the network (matrix cores, two 32-row tiles per workgroup) and adds out_scale * y_k / alpha~_k to
// multiplier k before the pass scales them.
void launch_hex_warm_start(hipStream_t s, tet_group_dev g, cluster_dev cl, const pos4* xt, pos4* xi,
                           const uint4* cinfo, uint32_t nrows, float inv_dt2, int expl, float2 scale,
                           float4* disp, const uint32_t* csr_off, const uint32_t* csr, uint32_t n,
                           const int* surface_of = nullptr, contact_dev const& c = {}, float c_at = 0.0f,
                           const uint32_t* prow = nullptr, uint32_t nprows = 0, int net_mode = kWarmNetNone,
                           warm_net_dev const& net = {});
// Training rows of a hex group after a captured substep: for j < rows, cluster ci = offset + j *
// stride (rowof: cluster -> its row of raw and cinfo), ds_x[base + j] = raw[rowof[ci]] (40 values)
// and ds_t[base + j] = the change of alpha~ lambda of its 12 multipliers over the substep (12
// values); rows holding a non-finite value are stored as zeros.
void launch_hex_capture(hipStream_t s, tet_group_dev g, const uint4* cinfo, const uint32_t* rowof, const float* raw,
                        uint32_t stride, uint32_t offset, float inv_dt2, float* ds_x, float* ds_t, uint64_t base,
                        uint32_t rows);

// Surface refresh (surface.hip): for every surface vertex s < S (particle surf[s]), normal[s] +=
// the face normals of its triangles tris[inc[off[s] .. off[s+1])] (surface vertex indices, in
// triangle order), then normal[s] is normalized; positions are read from x.
void launch_surface_refresh(hipStream_t s, const uint32_t* surf, const pos4* x, const uint4* tris, const uint32_t* off,
                            const uint32_t* inc, double4* normal, uint32_t S);

// Fills a buffer with a 32-bit value (graph-safe replacement of a memset).
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t value, uint32_t n);

constexpr uint32_t kUncolored = 0xffffffffu;
constexpr uint32_t kMaxColors = 256;

}  // namespace sbs
