// C ABI of the gfx950 XPBD time step (declared in include/sbs_gpu.h).
//
// The handle owns one HIP stream and every device buffer. sbs_gpu_commit() colors each constraint
// group on the device (Jones-Plassmann), lays the group's data out in color order and uploads the
// particles; sbs_gpu_step() then issues, per frame: contact detection, and per substep the
// prediction, `iterations` x (contact projection, one launch per color of every group), and the
// velocity update (fused with the next substep's prediction).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <memory>
#include <unordered_map>
#include <vector>

#include "abi_util.h"
#include "body_contact.h"
#include "contact_hash.h"
#include "grid_sdf.h"
#include "kernels.h"
#include "scan_sort.h"
#include "mlp.h"
#include "mlp_handle.h"
#include "sbs_gpu.h"

namespace sbs {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(std::string const& msg) { g_last_error = msg; }

struct body_t {
    uint32_t offset = 0, n = 0;
    std::vector<double> x0, mass;
    std::vector<uint32_t> surface;
    std::vector<uint32_t> triangles;  // boundary triangles, 3 surface vertex indices each
};

struct group_t {
    int kind = 0;  // 0 green, 1 neo-Hookean, 2 distance
    uint32_t body = 0, count = 0;
    double alpha = 0, beta = 0, young = 0, poisson = 0;
    std::vector<uint32_t> conn;  // 4 body-local ids per tet, or (b1, v1, b2, v2) per distance
    std::vector<uint32_t> cbody;        // per-tet body of a batched group (empty: `body` for all)
    std::vector<double> cyoung, cpoisson;  // per-tet material of a batched group (empty: uniform)
    std::vector<uint32_t> cluster_off;  // empty: every constraint is its own cluster
    // after commit
    uint32_t ncolors = 0;
    std::vector<uint32_t> color;          // per constraint, insertion order
    std::vector<uint32_t> color_offsets;  // ncolors + 1, constraint positions in execution order
    std::vector<uint32_t> cluster_color_offsets;  // ncolors + 1, cluster positions in execution order
    std::vector<xcd_split> xsplit;  // per color: its clusters by XCD (offsets inside the color)
    dev_buf<uint32_t> cl_off, cl_loc;
    dev_buf<int4> cl_v0, cl_v1;
    bool hex = false;  // every cluster is a hex cell in the fixed slot pattern
    dev_buf<int4> idx;
    dev_buf<float4> m0, m1, m2;
    dev_buf<uint2> didx;
    dev_buf<float> rest;
    dev_buf<float> lambda;
    // multiplier warm start (neo-Hookean groups)
    dev_buf<float2> ws_hist;     // the multiplier pairs one substep before the stored ones
    dev_buf<float> ws_raw;       // hex groups: [ws_rows][40] network features at the substep start (capture)
    dev_buf<float> ws_bsize;     // per constraint: log of the number of tets of its body
    dev_buf<uint32_t> ws_rowof_dev;  // hex groups: cluster -> its row (capture)
    dev_buf<float4> ws_disp;     // hex groups: slot displacements of the warm-start pass, [8][rows]
    dev_buf<uint32_t> ws_csr_off, ws_csr;  // particle -> slot list
    // hex groups, the all-color launches: rows in 64-row workgroups, workgroup b holding clusters
    // of XCD region b mod 8, all of one color; per workgroup two records (first cluster's index in
    // its color, clusters in the color, the color's first constraint, first cluster) and (rows in
    // use, 0, 0, 0)
    dev_buf<uint4> ws_cinfo;
    uint32_t ws_rows = 0;
    std::vector<uint32_t> ws_rowof;  // cluster -> its row
    bool clustered() const { return !cluster_off.empty(); }
    uint32_t body_of(uint32_t c) const { return cbody.empty() ? body : cbody[c]; }
    double young_of(uint32_t c) const { return cyoung.empty() ? young : cyoung[c]; }
    double poisson_of(uint32_t c) const { return cpoisson.empty() ? poisson : cpoisson[c]; }
    uint32_t num_elements() const { return clustered() ? uint32_t(cluster_off.size() - 1) : count; }
    tet_group_dev tet() const { return {idx.p, m0.p, m1.p, m2.p, lambda.p}; }
    cluster_dev clusters() const { return {cl_off.p, cl_v0.p, cl_v1.p, cl_loc.p}; }
    distance_group_dev dist() const { return {didx.p, rest.p, lambda.p}; }
};

constexpr std::size_t kMaxContactSlots = 8;   // contacts kept per surface vertex
// shape of the warm-start network: the features and multiplier changes of one hex cluster
constexpr int kWarmStartInputs = 39, kWarmStartOutputs = 12;
constexpr std::size_t kHashMinObstacles = 32;  // automatic mode hashes from this many bounded obstacles on
constexpr uint64_t kMaxHashCells = 1u << 16;   // cells of one hashed obstacle volume, at most

struct timing_rec {
    int kind;
    uint64_t launches, units;
    hipEvent_t a, b;
};

}  // namespace sbs

using namespace sbs;

struct sbs_gpu {
    int device = 0;
    hipStream_t stream = nullptr;
    sbs_xpbd_params params{};
    std::vector<body_t> bodies;
    std::vector<group_t> groups;
    std::vector<sdf_dev> sdfs;
    std::vector<std::unique_ptr<dev_buf<double>>> grid_nodes;  // node values of grid SDFs
    bool committed = false;
    uint32_t V = 0;

    dev_buf<pos4> xi, xn;
    dev_buf<pos4> xt;  // predicted positions of the substep (warm-start pass)
    dev_buf<float4> vel;
    dev_buf<float> mass;

    dev_buf<uint32_t> surf;
    uint32_t S = 0, slots = 0;
    // surface refresh: triangles (global surface indices), vertex -> triangle lists, normals
    bool surface_refresh = false;
    uint32_t ntris = 0;
    dev_buf<uint4> stris;
    dev_buf<uint32_t> stri_off, stri_inc;
    dev_buf<double4> snormal;
    dev_buf<pos4> cq;
    dev_buf<float4> cn;
    dev_buf<float> clam;
    dev_buf<uint32_t> ccount;
    dev_buf<sdf_dev> dsdf;
    dev_buf<uint32_t> coverflow;  // contacts dropped for want of slots (cumulative)
    uint32_t* overflow_host = nullptr;  // pinned: coverflow as of the last frame the device finished
    // spatial hash of the obstacles (contact broad phase)
    int hash_mode = SBS_CONTACT_HASH_AUTO;
    double hash_cell_size = 0.0;  // requested (0: automatic)
    bool use_hash = false;
    double hash_h = 0.0;
    int hash_bits = 0;
    uint32_t hash_pairs = 0;
    std::vector<uint32_t> hashed_ids, global_ids;
    dev_buf<uint32_t> hstart, hend, hentries, hcount, hglobal, hids, hws;

    dev_buf<double> partial;

    // contacts between bodies (surface vertex against surface triangle); sbs_gpu_set_body_contacts
    bool body_contacts = false;
    double body_cell_request = 0.0, body_thickness_request = 0.0, body_radius_request = 0.0;  // 0: defaults
    double body_cell = 0.0, body_thickness = 0.0, body_radius = 0.0;
    int body_bits = 0;
    dev_buf<uint4> btris;
    dev_buf<uint32_t> bkeys, bvals, bstart, bend, bws, bsbody, bbig, bnbig, btoff;
    dev_buf<uint4> bnbr;                 // edge neighbours of each triangle
    dev_buf<uint32_t> bvt_off, bvt;      // per surface vertex: its (triangle, corner) entries
    dev_buf<pos4> btn, bsph, bvn;        // per frame: triangle unit normals and bounding spheres, vertex pseudo-normals
    dev_buf<int4> btlo;                  // per frame: each triangle's lowest cell
    dev_buf<uint4> btpx, berec;          // per frame: corner particles and body; hash entry records
    dev_buf<float4> besph;               // per frame: hash entry spheres
    dev_buf<pos4> csnap;  // surface positions at the start of a contact launch (triangle contacts)
    body_contact_dev body_view()
    {
        return {btris.p, uint32_t(btris.n), 1.0 / body_cell, body_thickness, body_radius, (1u << body_bits) - 1u,
                uint32_t(sdfs.size()), bkeys.p, bvals.p, bstart.p, bend.p, bbig.p, bnbig.p, bsbody.p, btoff.p,
                bnbr.p, bvt_off.p, bvt.p, btn.p, bsph.p, btlo.p, btpx.p, berec.p, besph.p, bvn.p};
    }
    void setup_body_contacts(std::vector<uint4> const& tr, std::vector<pos4> const& x);

    int block = 256;
    int cluster_block = 64;
    bool fuse_first_contacts = true;  // the first iteration's contacts in the warm-start gather
    dev_buf<int> surface_of_dev;      // particle -> surface vertex, or -1
    dev_buf<uint32_t> prow;           // particle rows by XCD region (built with the hex warm-start tables)
    int coloring_order = SBS_COLORING_HASH;

    bool timing = false;
    std::vector<timing_rec> pending;
    std::vector<hipEvent_t> event_pool;
    uint64_t st_launches[SBS_KERNEL_COUNT] = {};
    double st_ms[SBS_KERNEL_COUNT] = {};
    uint64_t st_units[SBS_KERNEL_COUNT] = {};

    bool use_graphs = false;
    // instantiated frame graphs by launch key, most recently used last: a checked warm-started frame
    // and the plain frame that replaces a rejected one keep a graph each
    static constexpr std::size_t kGraphCache = 4;
    std::vector<std::pair<std::array<double, 16>, hipGraphExec_t>> graphs;
    void drop_graphs()
    {
        if (graphs.empty()) return;
        // a cached graph may still be running on the stream: let it finish first
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
    }

    int ws_mode = SBS_WARM_START_NONE;
    double ws_scale = 1.0;      // deviatoric multipliers' warm-start factor
    double ws_scale_hyd = 1.0;  // hydrostatic multipliers' factor
    sbs_mlp* ws_net = nullptr;
    sbs_mlp* cap_net = nullptr;
    uint32_t cap_stride = 0;
    uint64_t cap_substeps = 0;  // captured substeps so far (moves the captured clusters' offset)
    // snapshot slots: copies of the dynamic state (committed positions, velocities, normals, and per
    // group the multipliers and their warm-start history)
    struct snap_slot {
        dev_buf<pos4> xn;
        dev_buf<float4> vel;
        dev_buf<double4> normal;
        std::vector<dev_buf<float>> lambda;
        std::vector<dev_buf<float2>> hist;
        bool taken = false;
    };
    std::array<snap_slot, SBS_SNAPSHOT_SLOTS> snaps;
    // the per-frame fallback check's own state copy (the public snapshot is left alone): committed
    // positions, velocities, normals, and in learned mode the multiplier history. The plain frame
    // that replaces a rejected one writes every multiplier before reading it, so they are not kept.
    dev_buf<pos4> chk_xn;
    dev_buf<float4> chk_vel;
    dev_buf<double4> chk_normal;
    std::vector<dev_buf<float2>> chk_hist;  // per group
    double* resid_host = nullptr;  // pinned: the last frame's squared residual norm (fallback check)
    dev_buf<double> check_partial;  // its per-block sums (its own buffer: frame graphs hold the pointer)
    double fb_limit = 0.0;
    uint32_t fb_iterations = 0;  // 0: the step's own iteration count
    uint64_t fb_count = 0;

    ~sbs_gpu()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        for (auto& r : pending) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : event_pool) (void)hipEventDestroy(e);
        if (resid_host) (void)hipHostFree(resid_host);
        if (overflow_host) (void)hipHostFree(overflow_host);
        if (stream) (void)hipStreamDestroy(stream);
    }

    hipEvent_t take_event()
    {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }

    // Runs `launch` (which issues `launches` back-to-back kernels of one kind) on the stream,
    // bracketed by one pair of events when timing is on.
    template <class F>
    void timed(int kind, uint64_t units, F&& launch, uint64_t launches = 1)
    {
        if (!timing) {
            launch();
            return;
        }
        timing_rec r{kind, launches, units, take_event(), take_event()};
        hip_check(hipEventRecord(r.a, stream), "hipEventRecord");
        launch();
        hip_check(hipEventRecord(r.b, stream), "hipEventRecord");
        pending.push_back(r);
    }

    void harvest()
    {
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (auto& r : pending) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
            st_launches[r.kind] += r.launches;
            st_ms[r.kind] += ms;
            st_units[r.kind] += r.units;
            event_pool.push_back(r.a);
            event_pool.push_back(r.b);
        }
        pending.clear();
    }

    contact_dev contacts()
    {
        return {surf.p, S, slots, cq.p, cn.p, clam.p, ccount.p, coverflow.p, csnap.p, btris.p, uint32_t(sdfs.size()),
                body_thickness};
    }
    contact_hash_dev hash_view() const
    {
        return {hstart.p, hend.p, hentries.p, hglobal.p, uint32_t(global_ids.size()), (1u << hash_bits) - 1u,
                1.0 / hash_h};
    }
    void build_hash();

    void set_current() { hip_check(hipSetDevice(device), "hipSetDevice"); }
    bool all_nh_hex() const
    {
        for (auto const& g : groups)
            if (g.kind == 1 && g.count && !g.hex) return false;
        return true;
    }
    // SBS_ERR_CONTACT_OVERFLOW once a finished frame has dropped a contact (the pinned word is
    // written by the frame after its detection; read without waiting)
    void check_overflow() const
    {
        uint32_t const n = overflow_host ? *reinterpret_cast<volatile const uint32_t*>(overflow_host) : 0u;
        if (n)
            throw api_error(SBS_ERR_CONTACT_OVERFLOW,
                            std::to_string(n) + " contacts were dropped because a surface vertex had more than " +
                                std::to_string(slots) +
                                " (its contact slots; the state no longer follows the reference's contact set)");
    }

    void commit(uint32_t recolor_passes);
    void color_group(group_t& g, uint32_t recolor_passes);
    void upload_group(group_t& g);
    void enqueue_frame(double dt, uint32_t iterations, uint32_t substeps, int warm, bool check = false);
    void run_frame(double dt, uint32_t iterations, uint32_t substeps, int warm, bool check = false);
    void detect();
    void launch_group_colors(group_t const& g, double dt, int mode, const pos4* xref);
    void ensure_warm_buffers(bool raw);
    void ensure_cinfo(group_t& g);
    void snapshot(int slot);
    void prepare_snapshot(int slot);
    void enqueue_snapshot(int slot);
    void enqueue_check_snapshot();
    void restore(int slot);
    void restore_check();
    std::size_t residual_blocks() const;
    void prepare_check();
    void enqueue_residual(double dt);
};

namespace {

uint32_t global_vertex(sbs_gpu const& s, uint32_t body, uint32_t v)
{
    require(body < s.bodies.size(), "body index out of range");
    require(v < s.bodies[body].n, "vertex index out of range");
    return s.bodies[body].offset + v;
}

void lame(double young, double poisson, double* mu, double* lam)
{
    *mu = young / (2.0 * (1.0 + poisson));
    *lam = (young * poisson) / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
}

}  // namespace

// Colors the elements of a group (its clusters, or its constraints when it has none) so that
// elements of one color share no particle. SBS_COLORING_HASH: device Jones-Plassmann, pass 0 keyed
// on hash32(index), each recolor pass keyed on (previous color, hash32(index)) -- never more
// colors than the pass before. SBS_COLORING_INDEX: sequential first-fit greedy in index order
// (inherently serial; computed on the host at commit).
void sbs_gpu::color_group(group_t& g, uint32_t recolor_passes)
{
    uint32_t const ne = g.num_elements();
    // element -> particles (CSR), duplicates within an element removed
    std::vector<uint32_t> eoff(std::size_t(ne) + 1, 0), verts;
    verts.reserve(std::size_t(g.count) * 4);
    auto constraint_particles = [&](uint32_t c, uint32_t out[4]) -> int {
        if (g.kind == 2) {
            out[0] = global_vertex(*this, g.conn[4 * c], g.conn[4 * c + 1]);
            out[1] = global_vertex(*this, g.conn[4 * c + 2], g.conn[4 * c + 3]);
            return 2;
        }
        for (int k = 0; k < 4; ++k) out[k] = global_vertex(*this, g.body_of(c), g.conn[4 * std::size_t(c) + k]);
        return 4;
    };
    for (uint32_t e = 0; e < ne; ++e) {
        uint32_t const c0 = g.clustered() ? g.cluster_off[e] : e;
        uint32_t const c1 = g.clustered() ? g.cluster_off[e + 1] : e + 1;
        std::size_t const start = verts.size();
        for (uint32_t c = c0; c < c1; ++c) {
            uint32_t pv[4];
            int const np = constraint_particles(c, pv);
            for (int k = 0; k < np; ++k)
                if (std::find(verts.begin() + start, verts.end(), pv[k]) == verts.end()) verts.push_back(pv[k]);
        }
        eoff[e + 1] = uint32_t(verts.size());
    }
    // particle -> elements incidence (CSR)
    std::vector<uint32_t> off(std::size_t(V) + 1, 0), inc(verts.size());
    for (uint32_t x : verts) off[x + 1]++;
    for (std::size_t i = 0; i < V; ++i) off[i + 1] += off[i];
    {
        std::vector<uint32_t> fill(off.begin(), off.end() - 1);
        for (uint32_t e = 0; e < ne; ++e)
            for (uint32_t a = eoff[e]; a < eoff[e + 1]; ++a) inc[fill[verts[a]]++] = e;
    }
    std::vector<uint32_t> ecolor(ne);
    if (coloring_order == SBS_COLORING_INDEX) {
        std::vector<uint32_t> stamp(kMaxColors + 1, UINT32_MAX);
        for (uint32_t e = 0; e < ne; ++e) {
            for (uint32_t a = eoff[e]; a < eoff[e + 1]; ++a)
                for (uint32_t j = off[verts[a]]; j < off[verts[a] + 1]; ++j) {
                    uint32_t const u = inc[j];
                    if (u < e) stamp[ecolor[u]] = e;
                }
            uint32_t c = 0;
            while (stamp[c] == e) ++c;
            require(c < kMaxColors, "constraint group needs more than 256 colors");
            ecolor[e] = c;
        }
    } else {
        dev_buf<uint32_t> deoff, dverts, doff, dinc, ca, cb, prev, counters;
        dev_buf<uint64_t> key;
        deoff.upload(eoff.data(), eoff.size(), stream);
        dverts.upload(verts.data(), verts.size(), stream);
        doff.upload(off.data(), off.size(), stream);
        dinc.upload(inc.data(), inc.size(), stream);
        ca.alloc(ne);
        cb.alloc(ne);
        prev.alloc(ne);
        key.alloc(ne);
        counters.alloc(2);
        for (uint32_t pass = 0; pass <= recolor_passes; ++pass) {
            launch_jp_keys(stream, ne, pass == 0 ? nullptr : prev.p, key.p, 256);
            launch_fill_u32(stream, ca.p, kUncolored, ne);
            uint32_t h[2] = {1, 0};
            uint32_t rounds = 0;
            while (h[0] != 0) {
                require(++rounds < 1000000, "coloring did not terminate", SBS_ERR_INTERNAL);
                launch_fill_u32(stream, counters.p, 0u, 2);
                launch_jp_round(stream, ne, 0, deoff.p, dverts.p, doff.p, dinc.p, key.p, ca.p, cb.p, counters.p,
                                counters.p + 1, 256);
                hip_check(hipMemcpyAsync(h, counters.p, sizeof h, hipMemcpyDeviceToHost, stream), "download counters");
                hip_check(hipStreamSynchronize(stream), "coloring round");
                require(h[1] == 0, "constraint group needs more than 256 colors");
                std::swap(ca, cb);
            }
            hip_check(hipMemcpyAsync(prev.p, ca.p, ne * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream), "copy colors");
        }
        hip_check(hipMemcpyAsync(ecolor.data(), prev.p, ne * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
                  "download colors");
        hip_check(hipStreamSynchronize(stream), "coloring");
    }
    g.ncolors = 0;
    for (uint32_t c : ecolor) g.ncolors = std::max(g.ncolors, c + 1);
    g.color.resize(g.count);
    for (uint32_t e = 0; e < ne; ++e) {
        uint32_t const c0 = g.clustered() ? g.cluster_off[e] : e;
        uint32_t const c1 = g.clustered() ? g.cluster_off[e + 1] : e + 1;
        for (uint32_t c = c0; c < c1; ++c) g.color[c] = ecolor[e];
    }
}

// Lays a group out in color order (stable by insertion index) and uploads it in FP32; the rest
// shape inverse and volume are computed in FP64 first (green_constraint.cpp:38-46).
void sbs_gpu::upload_group(group_t& g)
{
    uint32_t const n = g.count;
    uint32_t const ne = g.num_elements();
    auto e_begin = [&](uint32_t e) { return g.clustered() ? g.cluster_off[e] : e; };
    auto e_end = [&](uint32_t e) { return g.clustered() ? g.cluster_off[e + 1] : e + 1; };
    // execution order: elements by color (stable by index), constraints of an element in order
    g.cluster_color_offsets.assign(g.ncolors + 1, 0);
    g.color_offsets.assign(g.ncolors + 1, 0);
    for (uint32_t e = 0; e < ne; ++e) {
        uint32_t const c = g.color[e_begin(e)];
        g.cluster_color_offsets[c + 1]++;
        g.color_offsets[c + 1] += e_end(e) - e_begin(e);
    }
    for (uint32_t c = 0; c < g.ncolors; ++c) {
        g.cluster_color_offsets[c + 1] += g.cluster_color_offsets[c];
        g.color_offsets[c + 1] += g.color_offsets[c];
    }
    std::vector<uint32_t> eorder(ne);
    {
        std::vector<uint32_t> fill(g.cluster_color_offsets.begin(), g.cluster_color_offsets.end() - 1);
        for (uint32_t e = 0; e < ne; ++e) eorder[fill[g.color[e_begin(e)]]++] = e;
    }
    if (g.clustered()) {
        // XCD affinity: the clusters in Morton order of their first constraint's rest centroid,
        // split into 8 regions of equal count; inside each color, clusters by region, then by
        // index (the caller's order, which keeps neighbouring clusters' particle gathers
        // coalesced), with the region boundaries as the color's XCD split. Elements of one color
        // are independent, so their order does not change the result.
        std::vector<std::array<double, 3>> cen(ne);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (uint32_t e = 0; e < ne; ++e) {
            uint32_t const c = e_begin(e);
            body_t const& b = bodies[g.body_of(c)];
            for (int d = 0; d < 3; ++d) {
                double m = 0.0;
                for (int k = 0; k < 4; ++k) m += b.x0[3 * std::size_t(g.conn[4 * std::size_t(c) + k]) + d];
                cen[e][d] = 0.25 * m;
                lo[d] = std::min(lo[d], cen[e][d]);
                hi[d] = std::max(hi[d], cen[e][d]);
            }
        }
        auto spread = [](uint64_t v) {
            v &= 0x1fffffull;
            v = (v | (v << 32)) & 0x1f00000000ffffull;
            v = (v | (v << 16)) & 0x1f0000ff0000ffull;
            v = (v | (v << 8)) & 0x100f00f00f00f00full;
            v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
            v = (v | (v << 2)) & 0x1249249249249249ull;
            return v;
        };
        std::vector<std::pair<uint64_t, uint32_t>> key(ne);
        for (uint32_t e = 0; e < ne; ++e) {
            uint64_t code = 0;
            for (int d = 0; d < 3; ++d) {
                double const ext = hi[d] - lo[d];
                double const u = ext > 0.0 ? (cen[e][d] - lo[d]) / ext : 0.0;
                code |= spread(uint64_t(std::min(2097151.0, std::max(0.0, u * 2097152.0)))) << d;
            }
            key[e] = {code, e};
        }
        std::sort(key.begin(), key.end());
        std::vector<uint32_t> rank(ne), region(ne);
        for (uint32_t r = 0; r < ne; ++r) rank[key[r].second] = r;
        for (uint32_t e = 0; e < ne; ++e) region[e] = uint32_t(uint64_t(rank[e]) * kXcds / ne);
        g.xsplit.assign(g.ncolors, xcd_split{});
        for (uint32_t c = 0; c < g.ncolors; ++c) {
            uint32_t const cb = g.cluster_color_offsets[c], ce = g.cluster_color_offsets[c + 1];
            std::stable_sort(eorder.begin() + cb, eorder.begin() + ce,
                             [&](uint32_t a, uint32_t b) { return region[a] < region[b]; });
            xcd_split& sp = g.xsplit[c];
            uint32_t j = cb;
            for (int x = 0; x <= kXcds; ++x) {
                while (j < ce && int(region[eorder[j]]) < x) ++j;
                sp.off[x] = j - cb;
            }
            sp.off[kXcds] = ce - cb;
        }
    }
    std::vector<uint32_t> order;
    order.reserve(n);
    for (uint32_t e : eorder)
        for (uint32_t c = e_begin(e); c < e_end(e); ++c) order.push_back(c);
    if (g.clustered()) {
        std::vector<uint32_t> off(ne + 1, 0), loc(n);
        std::vector<int4> v0(ne), v1(ne);
        uint32_t pos = 0;
        for (uint32_t j = 0; j < ne; ++j) {
            uint32_t const e = eorder[j];
            int slots[kClusterSlots];
            int used = 0;
            for (uint32_t c = e_begin(e); c < e_end(e); ++c, ++pos) {
                uint32_t l = 0;
                body_t const& b = bodies[g.body_of(c)];
                for (int k = 0; k < 4; ++k) {
                    int const gv = int(b.offset + g.conn[4 * std::size_t(c) + k]);
                    int s = 0;
                    while (s < used && slots[s] != gv) ++s;
                    if (s == used) {
                        require(used < kClusterSlots, "a cluster touches more than 8 particles");
                        slots[used++] = gv;
                    }
                    l |= uint32_t(s) << (8 * k);
                }
                loc[pos] = l;
            }
            for (int s = used; s < kClusterSlots; ++s) slots[s] = -1;
            v0[j] = make_int4(slots[0], slots[1], slots[2], slots[3]);
            v1[j] = make_int4(slots[4], slots[5], slots[6], slots[7]);
            off[j + 1] = pos;
        }
        g.hex = true;
        for (uint32_t j = 0; j < ne && g.hex; ++j) {
            if (off[j + 1] - off[j] != 6) g.hex = false;
            for (uint32_t k = 0; k < 6 && g.hex; ++k)
                if (loc[off[j] + k] != kHexPatternLoc[k]) g.hex = false;
            if (g.hex && v1[j].w < 0) g.hex = false;
        }
        if (g.hex) {
            // slot-major constraint layout inside each color: the k-th tet of the color's j-th
            // cluster sits at color_offsets[c] + k * n_c + j, so lane j's loads coalesce
            order.clear();
            for (uint32_t c = 0; c < g.ncolors; ++c) {
                uint32_t const cb = g.cluster_color_offsets[c], ce = g.cluster_color_offsets[c + 1];
                for (uint32_t k = 0; k < 6; ++k)
                    for (uint32_t j = cb; j < ce; ++j) order.push_back(e_begin(eorder[j]) + k);
            }
        }
        g.cl_off.upload(off.data(), off.size(), stream);
        g.cl_loc.upload(loc.data(), loc.size(), stream);
        g.cl_v0.upload(v0.data(), v0.size(), stream);
        g.cl_v1.upload(v1.data(), v1.size(), stream);
    }
    if (g.kind == 2) {
        std::vector<uint2> idx(n);
        std::vector<float> rest(n);
        for (uint32_t j = 0; j < n; ++j) {
            uint32_t const* e = &g.conn[4 * std::size_t(order[j])];
            uint32_t const a = global_vertex(*this, e[0], e[1]), b = global_vertex(*this, e[2], e[3]);
            idx[j] = make_uint2(a, b);
            double const* pa = &bodies[e[0]].x0[3 * std::size_t(e[1])];
            double const* pb = &bodies[e[2]].x0[3 * std::size_t(e[3])];
            double const d[3] = {pa[0] - pb[0], pa[1] - pb[1], pa[2] - pb[2]};
            rest[j] = float(std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
        }
        g.didx.upload(idx.data(), n, stream);
        g.rest.upload(rest.data(), n, stream);
        g.lambda.alloc(n);
    } else {
        std::vector<int4> idx(n);
        std::vector<float4> m0(n), m1(n), m2(g.kind == 1 ? 0 : n);
        std::vector<float2> m2h(g.kind == 1 ? n + (n & 1u) : 0, make_float2(0.0f, 0.0f));
        for (uint32_t j = 0; j < n; ++j) {
            uint32_t const* t = &g.conn[4 * std::size_t(order[j])];
            body_t const& b = bodies[g.body_of(order[j])];
            double mu, lam;
            lame(g.young_of(order[j]), g.poisson_of(order[j]), &mu, &lam);
            double D[3][3];  // D[r][c] = (p_c - p_4)[r]
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) D[r][c] = b.x0[3 * std::size_t(t[c]) + r] - b.x0[3 * std::size_t(t[3]) + r];
            double A[3][3];  // adjugate
            A[0][0] = D[1][1] * D[2][2] - D[1][2] * D[2][1];
            A[0][1] = D[0][2] * D[2][1] - D[0][1] * D[2][2];
            A[0][2] = D[0][1] * D[1][2] - D[0][2] * D[1][1];
            A[1][0] = D[1][2] * D[2][0] - D[1][0] * D[2][2];
            A[1][1] = D[0][0] * D[2][2] - D[0][2] * D[2][0];
            A[1][2] = D[0][2] * D[1][0] - D[0][0] * D[1][2];
            A[2][0] = D[1][0] * D[2][1] - D[1][1] * D[2][0];
            A[2][1] = D[0][1] * D[2][0] - D[0][0] * D[2][1];
            A[2][2] = D[0][0] * D[1][1] - D[0][1] * D[1][0];
            double const det = D[0][0] * A[0][0] + D[0][1] * A[1][0] + D[0][2] * A[2][0];
            require(det != 0.0, "degenerate tet in rest shape");
            double R[3][3];  // DmInv, rows r_k
            for (int i = 0; i < 9; ++i) R[i / 3][i % 3] = A[i / 3][i % 3] / det;
            double const V0 = det / 6.0;
            idx[j] = make_int4(int(b.offset + t[0]), int(b.offset + t[1]), int(b.offset + t[2]), int(b.offset + t[3]));
            if (g.kind == 0) {  // green: the rows of DmInv, V0, mu, lambda
                float r[9];
                for (int i = 0; i < 9; ++i) r[i] = float(R[i / 3][i % 3]);
                m0[j] = make_float4(r[0], r[1], r[2], r[3]);
                m1[j] = make_float4(r[4], r[5], r[6], r[7]);
                m2[j] = make_float4(r[8], float(V0), float(mu), float(lam));
            } else {  // neo-Hookean: Q = R R^T, det R and the compliances (the kernels' nh_rest), 40 B
                auto q = [&](int a, int c) {
                    return float(R[a][0] * R[c][0] + R[a][1] * R[c][1] + R[a][2] * R[c][2]);
                };
                double const av = std::fabs(V0);
                m0[j] = make_float4(q(0, 0), q(1, 1), q(2, 2), q(0, 1));
                m1[j] = make_float4(q(0, 2), q(1, 2), float(1.0 / det), float(1.0 / (mu * av)));
                m2h[j] = make_float2(float(1.0 / (lam * av)), float(1.0 + mu / lam));
            }
        }
        g.idx.upload(idx.data(), n, stream);
        g.m0.upload(m0.data(), n, stream);
        g.m1.upload(m1.data(), n, stream);
        if (g.kind == 1)  // float2 per tet, in float4 storage
            g.m2.upload(reinterpret_cast<const float4*>(m2h.data()), m2h.size() / 2, stream);
        else
            g.m2.upload(m2.data(), n, stream);
        g.lambda.alloc(std::size_t(n) * (g.kind == 1 ? 2 : 1));
        if (g.kind == 1) {  // warm-start feature: log of the number of the group's tets in each tet's body
            std::vector<uint32_t> per_body(bodies.size(), 0);
            for (uint32_t c = 0; c < n; ++c) ++per_body[g.body_of(c)];
            std::vector<float> bs(n);
            for (uint32_t j = 0; j < n; ++j) bs[j] = float(std::log(double(per_body[g.body_of(order[j])])));
            g.ws_bsize.upload(bs.data(), n, stream);
        }
    }
    // multipliers start at zero (the "previous" multipliers of a warm-started first substep)
    if (g.lambda.n) hip_check(hipMemsetAsync(g.lambda.p, 0, g.lambda.n * sizeof(float), stream), "clear multipliers");
    hip_check(hipStreamSynchronize(stream), "upload group");
}

void sbs_gpu::commit(uint32_t recolor_passes)
{
    require(!committed, "sbs_gpu_commit called twice", SBS_ERR_STATE);
    set_current();
    V = 0;
    for (auto& b : bodies) {
        b.offset = V;
        V += b.n;
    }
    std::vector<pos4> x(V);
    std::vector<float4> v(V, make_float4(0, 0, 0, 0));
    std::vector<float> m(V);
    for (auto const& b : bodies)
        for (uint32_t i = 0; i < b.n; ++i) {
            double const mi = b.mass.empty() ? 1.0 : b.mass[i];
            double const w = mi > 0.0 ? 1.0 / mi : 0.0;  // particle.cpp:39-44
            x[b.offset + i] = make_double4(b.x0[3 * i], b.x0[3 * i + 1], b.x0[3 * i + 2], w);
            m[b.offset + i] = float(mi);
        }
    xi.upload(x.data(), V, stream);
    xn.upload(x.data(), V, stream);
    vel.upload(v.data(), V, stream);
    mass.upload(m.data(), V, stream);

    for (auto& g : groups) {
        if (g.count == 0) {
            g.ncolors = 0;
            g.color_offsets.assign(1, 0);
            continue;
        }
        color_group(g, recolor_passes);
        upload_group(g);
    }

    std::vector<uint32_t> sv;
    for (auto const& b : bodies)
        for (uint32_t s : b.surface) sv.push_back(b.offset + s);
    S = uint32_t(sv.size());
    std::vector<uint4> tr;  // boundary triangles of all bodies in global surface numbering, w = body
    {  // and each vertex's triangles
        uint32_t base = 0;
        for (uint32_t bi = 0; bi < bodies.size(); ++bi) {
            body_t const& b = bodies[bi];
            for (std::size_t t = 0; t + 2 < b.triangles.size(); t += 3)
                tr.push_back(make_uint4(base + b.triangles[t], base + b.triangles[t + 1], base + b.triangles[t + 2], bi));
            base += uint32_t(b.surface.size());
        }
        ntris = uint32_t(tr.size());
        if (ntris) {
            std::vector<uint32_t> off(std::size_t(S) + 1, 0), inc(3 * std::size_t(ntris));
            for (auto const& t : tr) {
                off[t.x + 1]++;
                off[t.y + 1]++;
                off[t.z + 1]++;
            }
            for (std::size_t i = 0; i < S; ++i) off[i + 1] += off[i];
            std::vector<uint32_t> fill(off.begin(), off.end() - 1);
            for (uint32_t j = 0; j < ntris; ++j) {
                inc[fill[tr[j].x]++] = j;
                inc[fill[tr[j].y]++] = j;
                inc[fill[tr[j].z]++] = j;
            }
            stris.upload(tr.data(), tr.size(), stream);
            stri_off.upload(off.data(), off.size(), stream);
            stri_inc.upload(inc.data(), inc.size(), stream);
            snormal.alloc(S);
            hip_check(hipMemsetAsync(snormal.p, 0, S * sizeof(double4), stream), "clear normals");
        }
    }
    // one more contact slot for the nearest surface triangle of another body (or of the vertex's own)
    bool const bodies_collide = body_contacts && S && !tr.empty();
    slots = uint32_t(std::min<std::size_t>(sdfs.size() + (bodies_collide ? 2 : 0), kMaxContactSlots));
    if (S) surf.upload(sv.data(), S, stream);
    {
        std::vector<int> so(V, -1);
        for (uint32_t s = 0; s < S; ++s) so[sv[s]] = int(s);
        surface_of_dev.upload(so.data(), so.size(), stream);
    }
    if (S && slots) {
        cq.alloc(std::size_t(S) * slots);
        cn.alloc(std::size_t(S) * slots);
        clam.alloc(std::size_t(S) * slots);
        ccount.alloc(S);
        coverflow.alloc(1);
        hip_check(hipMemsetAsync(coverflow.p, 0, sizeof(uint32_t), stream), "clear overflow");
        void* p = nullptr;
        hip_check(hipHostMalloc(&p, sizeof(uint32_t), hipHostMallocDefault), "hipHostMalloc");
        overflow_host = static_cast<uint32_t*>(p);
        *overflow_host = 0;
        dsdf.upload(sdfs.data(), sdfs.size(), stream);
        build_hash();
        if (bodies_collide) setup_body_contacts(tr, x);
    }
    hip_check(hipStreamSynchronize(stream), "commit");
    committed = true;
}

void sbs_gpu::detect()
{
    if (S && slots) {
        contact_dev const c = contacts();
        if (use_hash)
            timed(SBS_KERNEL_DETECT, S, [&] { launch_detect_hashed(stream, c, xn.p, dsdf.p, hash_view(), block); });
        else
            timed(SBS_KERNEL_DETECT, S, [&] { launch_detect(stream, c, xn.p, dsdf.p, uint32_t(sdfs.size()), block); });
        if (btris.n)  // after the SDF contacts of every vertex: the nearest surface triangle
            timed(SBS_KERNEL_DETECT, S, [&] {
                body_contact_dev const b = body_view();
                build_body_hash(stream, b, surf.p, xn.p, body_bits, bws.p);
                launch_detect_bodies(stream, c, b, xn.p, block);
            });
    }
}

// The surface triangles as contact obstacles: each body's triangles oriented outward (a body whose
// surface encloses a negative volume at rest has every triangle flipped), the cell edge (default
// twice the mean surface edge at rest), the thickness (default a tenth of the mean edge) and the
// radius (default the mean edge), a table of at least 4 slots per triangle, and the per-frame
// buffers.
void sbs_gpu::setup_body_contacts(std::vector<uint4> const& tr, std::vector<pos4> const& x)
{
    std::vector<uint32_t> sv, sbody;
    for (uint32_t bi = 0; bi < bodies.size(); ++bi)
        for (uint32_t s : bodies[bi].surface) {
            sv.push_back(bodies[bi].offset + s);
            sbody.push_back(bi);
        }
    bsbody.upload(sbody.data(), sbody.size(), stream);
    {  // the triangles come body by body
        std::vector<uint32_t> toff(bodies.size() + 1, 0);
        for (std::size_t t = 0; t < tr.size(); ++t) {
            require(t == 0 || tr[t].w >= tr[t - 1].w, "surface triangles must be grouped by body");
            ++toff[tr[t].w + 1];
        }
        for (std::size_t k = 0; k < bodies.size(); ++k) toff[k + 1] += toff[k];
        btoff.upload(toff.data(), toff.size(), stream);
    }
    std::vector<double> vol(bodies.size(), 0.0);
    double edge_sum = 0.0;
    auto P = [&](uint32_t s) { pos4 const& q = x[sv[s]]; return std::array<double, 3>{q.x, q.y, q.z}; };
    for (auto const& t : tr) {
        auto const a = P(t.x), b = P(t.y), c = P(t.z);
        vol[t.w] += a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
        auto len = [](std::array<double, 3> const& u, std::array<double, 3> const& v) {
            double const dx = u[0] - v[0], dy = u[1] - v[1], dz = u[2] - v[2];
            return std::sqrt(dx * dx + dy * dy + dz * dz);
        };
        edge_sum += len(a, b) + len(b, c) + len(c, a);
    }
    double const mean_edge = edge_sum / (3.0 * double(tr.size()));
    body_cell = body_cell_request > 0.0 ? body_cell_request : 2.0 * mean_edge;
    body_thickness = body_thickness_request > 0.0 ? body_thickness_request : 0.1 * mean_edge;
    body_radius = body_radius_request > 0.0 ? body_radius_request : mean_edge;
    require(body_thickness < body_radius && body_radius <= 0.999 * body_cell,
            "body contacts need thickness < radius <= 0.999 cell size");
    std::vector<uint4> oriented(tr);
    for (auto& t : oriented)
        if (vol[t.w] < 0.0) std::swap(t.y, t.z);
    body_bits = 10;
    // two slots per triangle: the occupied cells (about 0.15 per triangle) stay sparse, and up to 4M
    // triangles the keys with their padding bit fit 24 bits, three radix passes
    while ((uint64_t(1) << body_bits) < 2 * uint64_t(tr.size()) && body_bits < 28) ++body_bits;
    std::size_t const n = std::size_t(kTriCells) * tr.size();
    require(n < (std::size_t(1) << 31), "too many surface triangles for the body contact hash");
    btris.upload(oriented.data(), oriented.size(), stream);
    // the triangle across each edge (the oriented edge b->a of the neighbour matches a->b), and each
    // surface vertex's (triangle, corner) entries in increasing triangle order
    std::vector<uint4> nbr(oriented.size(), make_uint4(~0u, ~0u, ~0u, 0u));
    {
        std::unordered_map<uint64_t, uint32_t> edge;
        edge.reserve(3 * oriented.size());
        auto key = [](uint32_t a, uint32_t b) { return (uint64_t(a) << 32) | b; };
        for (uint32_t t = 0; t < oriented.size(); ++t) {
            uint32_t const v[3] = {oriented[t].x, oriented[t].y, oriented[t].z};
            for (int k = 0; k < 3; ++k) edge[key(v[k], v[(k + 1) % 3])] = t;
        }
        for (uint32_t t = 0; t < oriented.size(); ++t) {
            uint32_t const v[3] = {oriented[t].x, oriented[t].y, oriented[t].z};
            uint32_t o[3];
            for (int k = 0; k < 3; ++k) {
                auto it = edge.find(key(v[(k + 1) % 3], v[k]));
                o[k] = it == edge.end() ? ~0u : it->second;
            }
            nbr[t] = make_uint4(o[0], o[1], o[2], 0u);
        }
    }
    std::vector<uint32_t> vt_off(std::size_t(S) + 1, 0), vt(3 * oriented.size());
    for (auto const& t : oriented) ++vt_off[t.x + 1], ++vt_off[t.y + 1], ++vt_off[t.z + 1];
    for (std::size_t s = 0; s < S; ++s) vt_off[s + 1] += vt_off[s];
    {
        std::vector<uint32_t> fill(vt_off.begin(), vt_off.end() - 1);
        for (uint32_t t = 0; t < oriented.size(); ++t) {
            uint32_t const v[3] = {oriented[t].x, oriented[t].y, oriented[t].z};
            for (uint32_t k = 0; k < 3; ++k) vt[fill[v[k]]++] = 3 * t + k;
        }
    }
    bnbr.upload(nbr.data(), nbr.size(), stream);
    bvt_off.upload(vt_off.data(), vt_off.size(), stream);
    bvt.upload(vt.data(), vt.size(), stream);
    btn.alloc(oriented.size());
    bsph.alloc(oriented.size());
    btlo.alloc(oriented.size());
    btpx.alloc(oriented.size());
    berec.alloc(std::size_t(kTriCells) * oriented.size());
    besph.alloc(std::size_t(kTriCells) * oriented.size());
    bvn.alloc(S);
    bbig.alloc(oriented.size());
    bnbig.alloc(1);
    bkeys.alloc(n);
    bvals.alloc(n);
    bstart.alloc(std::size_t(1) << body_bits);
    bend.alloc(std::size_t(1) << body_bits);
    bws.alloc(body_contact_workspace(uint32_t(tr.size())));
    csnap.alloc(S);
}

// Obstacles with a finite volume of at most kMaxHashCells cells are hashed; the others (unbounded
// planes, huge volumes) go to the list every vertex tests. Automatic cell edge: the median of the
// hashed volumes' largest extent. Table: the smallest power of two >= 2 x (cells over all hashed
// volumes), at least 2^10.
void sbs_gpu::build_hash()
{
    use_hash = false;
    hashed_ids.clear();
    global_ids.clear();
    if (hash_mode == SBS_CONTACT_HASH_OFF) return;
    std::vector<double> ext;
    for (uint32_t k = 0; k < sdfs.size(); ++k) {
        sdf_dev const& f = sdfs[k];
        bool bounded = true;
        double e = 0.0;
        for (int d = 0; d < 3; ++d) {
            bounded = bounded && std::fabs(f.vmin[d]) < 1e30 && std::fabs(f.vmax[d]) < 1e30 && f.vmax[d] >= f.vmin[d];
            e = std::max(e, f.vmax[d] - f.vmin[d]);
        }
        if (bounded) ext.push_back(e);
    }
    if (ext.empty() || (hash_mode == SBS_CONTACT_HASH_AUTO && ext.size() < kHashMinObstacles)) return;
    std::vector<double> sorted = ext;
    std::nth_element(sorted.begin(), sorted.begin() + sorted.size() / 2, sorted.end());
    hash_h = hash_cell_size > 0.0 ? hash_cell_size : std::max(sorted[sorted.size() / 2], 1e-6);
    double const inv_h = 1.0 / hash_h;
    uint64_t pairs = 0;
    for (uint32_t k = 0; k < sdfs.size(); ++k) {
        sdf_dev const& f = sdfs[k];
        bool bounded = true;
        uint64_t cells = 1;
        for (int d = 0; d < 3; ++d) {
            bounded = bounded && std::fabs(f.vmin[d]) < 1e30 && std::fabs(f.vmax[d]) < 1e30 && f.vmax[d] >= f.vmin[d];
            if (!bounded) break;
            cells *= uint64_t(hash_cell(f.vmax[d], inv_h) - hash_cell(f.vmin[d], inv_h) + 1);
        }
        if (bounded && cells <= kMaxHashCells) {
            hashed_ids.push_back(k);
            pairs += cells;
        } else {
            global_ids.push_back(k);
        }
    }
    require(pairs < (uint64_t(1) << 31), "obstacle volumes cover too many hash cells; use a larger cell size");
    hash_pairs = uint32_t(pairs);
    hash_bits = 10;
    while ((uint64_t(1) << hash_bits) < 2 * pairs && hash_bits < 28) ++hash_bits;
    std::size_t const T = std::size_t(1) << hash_bits;
    hstart.alloc(T);
    hend.alloc(T);
    hentries.alloc(std::max<std::size_t>(pairs, 1));
    hcount.alloc(1);
    hids.upload(hashed_ids.data(), hashed_ids.size(), stream);
    hglobal.upload(global_ids.data(), global_ids.size(), stream);
    hws.alloc(contact_hash_workspace(uint32_t(hashed_ids.size()), pairs));
    build_contact_hash(stream, dsdf.p, hids.p, uint32_t(hashed_ids.size()), pairs, inv_h, hash_bits, hstart.p, hend.p,
                       hentries.p, hcount.p, hws.p);
    hip_check(hipGetLastError(), "contact hash build");
    hip_check(hipStreamSynchronize(stream), "contact hash build");
    hws.release();
    use_hash = true;
}

// One launch per color of a group (mode: kSolveFirst / kSolveNext, or for neo-Hookean groups a
// warm-start pass). xref: the substep start positions (damping), or for a warm-start pass the
// copy of the predicted positions the gradients are taken at.
void sbs_gpu::launch_group_colors(group_t const& g, double dt, int mode, const pos4* xref)
{
    double const dt2 = dt * dt;
    double const at = g.alpha / dt2;
    double const gamma = at * (g.beta * dt2) / dt;
    int const first = mode == kSolveFirst;
    // warm-start passes carry the multipliers' scale in place of the damping term
    float const bdt = mode >= kApplyWarmStart ? float(ws_scale) : float(g.beta * dt);
    for (uint32_t col = 0; col < g.ncolors; ++col) {
        uint32_t const b = g.color_offsets[col], n = g.color_offsets[col + 1] - b;
        uint32_t const cb = g.cluster_color_offsets[col];
        uint32_t const cn = g.cluster_color_offsets[col + 1] - cb;
        if (g.hex && g.kind == 0)
            launch_green_hex(stream, g.tet(), g.clusters(), xi.p, xref, cb, cn, b, g.xsplit[col], float(at),
                             float(gamma), first, cluster_block);
        else if (g.hex && g.kind == 1)
            launch_neohookean_hex(stream, g.tet(), g.clusters(), xi.p, xref, cb, cn, b, g.xsplit[col],
                                  float(1.0 / dt2), bdt, mode, cluster_block, float(ws_scale_hyd));
        else if (g.clustered() && g.kind == 0)
            launch_green_clusters(stream, g.tet(), g.clusters(), xi.p, xref, cb, cn, float(at), float(gamma), first,
                                  cluster_block);
        else if (g.clustered() && g.kind == 1)
            launch_neohookean_clusters(stream, g.tet(), g.clusters(), xi.p, xref, cb, cn, float(1.0 / dt2),
                                       bdt, mode, cluster_block, float(ws_scale_hyd));
        else if (g.kind == 0)
            launch_green(stream, g.tet(), xi.p, xref, b, n, float(at), float(gamma), first, block);
        else if (g.kind == 1)
            launch_neohookean(stream, g.tet(), xi.p, xref, b, n, float(1.0 / dt2), bdt, mode, block,
                              float(ws_scale_hyd));
        else
            launch_distance(stream, g.dist(), xi.p, xref, b, n, float(at), float(gamma), first, block);
    }
}

// check: the frame starts with a snapshot of the state and ends with the residual for the
// fallback decision (the caller prepared both).
void sbs_gpu::enqueue_frame(double dt_frame, uint32_t iterations, uint32_t substeps, int warm, bool check)
{
    if (check) enqueue_check_snapshot();
    double const dt = dt_frame / double(substeps);
    double const dt2 = dt * dt;
    float const g0 = float(params.gravity[0]), g1 = float(params.gravity[1]), g2 = float(params.gravity[2]);
    int const force = params.gravity_is_force;
    detect();  // timestep.cpp:29-30
    bool const have_contacts = S && slots;
    if (have_contacts) launch_publish_u32(stream, coverflow.p, overflow_host);
    contact_dev const c = contacts();
    float const c_at = float(params.collision_compliance / dt2);
    bool const capture = cap_net != nullptr && cap_stride > 0;
    for (uint32_t s = 0; s < substeps; ++s) {
        if (s == 0)
            timed(SBS_KERNEL_PREDICT, V, [&] { launch_predict(stream, xi.p, xn.p, vel.p, V, g0, g1, g2, force, float(dt), block); });
        else
            timed(SBS_KERNEL_PREDICT, V,
                  [&] {
                      launch_commit_predict(stream, xi.p, xn.p, vel.p, V, g0, g1, g2, force, float(dt), block, prow.p,
                                            uint32_t(prow.n));
                  });
        // multiplier warm start of the neo-Hookean groups: lambda_init (previous, plus the network's
        // change, or -C(x~)/alpha~), then x = x~ + M^-1 grad C(x~)^T lambda_init
        bool warmed = false;
        // every gradient of the pass is taken at x~: a copy of it is needed unless a single hex group
        // (whose displacements are all computed before any is added) is the only neo-Hookean group
        uint32_t nh = 0;
        bool colored_pass = false;
        for (auto const& g : groups) {
            nh += g.kind == 1;
            colored_pass = colored_pass || (g.kind == 1 && !g.hex);
        }
        bool const copy_xt = warm != SBS_WARM_START_NONE && (colored_pass || nh > 1);
        // the first iteration's contacts ride in the warm-start pass's gather when it is the last
        // launch before them: one group, hex clusters, the dense gather
        bool const fused0 = have_contacts && fuse_first_contacts && iterations > 0 && warm != SBS_WARM_START_NONE &&
                            groups.size() == 1 && !csnap.p &&
                            groups[0].kind == 1 && groups[0].hex && !groups[0].ws_csr_off.p;
        if (copy_xt) timed(SBS_KERNEL_WARM_START, V, [&] { launch_copy_pos(stream, xi.p, xt.p, V); });
        const pos4* const xtilde = copy_xt ? xt.p : xi.p;
        for (auto& g : groups) {
            if (g.kind != 1) continue;
            bool const learned = warm == SBS_WARM_START_LEARNED;
            if (warm != SBS_WARM_START_NONE) {
                int const mode = warm == SBS_WARM_START_EXPLICIT ? kApplyExplicit : kApplyWarmStart;
                if (g.hex) {
                    // the network's features (capture) or its predicted multiplier change (learned)
                    // are formed inside the pass's cluster kernel
                    int const net_mode = learned ? kWarmNetLearned : (capture ? kWarmNetFeatures : kWarmNetNone);
                    warm_net_dev wn{g.ws_hist.p, g.ws_bsize.p, capture ? g.ws_raw.p : nullptr,
                                    learned ? ws_net->pack.p : nullptr, learned ? ws_net->bias.p : nullptr,
                                    learned ? ws_net->norm.p : nullptr};
                    timed(SBS_KERNEL_WARM_START, g.count, [&] {
                        launch_hex_warm_start(stream, g.tet(), g.clusters(), xtilde, xi.p, g.ws_cinfo.p,
                                              g.ws_rows, float(1.0 / dt2), mode == kApplyExplicit,
                                              make_float2(float(ws_scale), float(ws_scale_hyd)), g.ws_disp.p,
                                              g.ws_csr_off.p, g.ws_csr.p, V, fused0 ? surface_of_dev.p : nullptr, c,
                                              c_at, prow.p, uint32_t(prow.n), net_mode, wn);
                    }, 2);
                } else {
                    timed(SBS_KERNEL_WARM_START, g.count, [&] { launch_group_colors(g, dt, mode, xt.p); }, g.ncolors);
                }
                warmed = true;
            }
        }
        for (uint32_t it = 0; it < iterations; ++it) {
            int const first = it == 0;
            if (have_contacts && !(first && fused0 && warmed))
                timed(SBS_KERNEL_CONTACT, S, [&] {
                    if (csnap.p) launch_snapshot_surface(stream, surf.p, xi.p, csnap.p, S);
                    launch_contact(stream, c, xi.p, c_at, first, block);
                });
            for (auto const& g : groups) {
                int const kind = g.kind == 0 ? SBS_KERNEL_GREEN : (g.kind == 1 ? SBS_KERNEL_NEOHOOKEAN : SBS_KERNEL_DISTANCE);
                int const mode = (first && !(warmed && g.kind == 1)) ? kSolveFirst : kSolveNext;
                // the colors of a group run back to back; timing brackets the whole run
                timed(kind, g.count, [&] { launch_group_colors(g, dt, mode, xn.p); }, g.ncolors);
            }
        }
        if (capture) {
            for (auto& g : groups) {
                if (g.kind != 1) continue;
                // clusters offset + j * stride; the offset moves from substep to substep
                uint32_t const nc = g.num_elements();
                uint32_t const offset = uint32_t((uint64_t(cap_substeps++) * 7919u) % cap_stride) % nc;
                uint64_t const want = (uint64_t(nc - offset) + cap_stride - 1) / cap_stride;
                uint64_t const room = cap_net->ds_cap - cap_net->ds_size;
                uint32_t const rows = uint32_t(std::min(want, room));
                launch_hex_capture(stream, g.tet(), g.ws_cinfo.p, g.ws_rowof_dev.p, g.ws_raw.p, cap_stride, offset,
                                   float(1.0 / dt2), cap_net->ds_x.p, cap_net->ds_t.p, cap_net->ds_size, rows);
                cap_net->ds_size += rows;
            }
        }
    }
    timed(SBS_KERNEL_COMMIT, V, [&] { launch_commit(stream, xi.p, xn.p, vel.p, V, float(dt), block); });
    if (surface_refresh && ntris)  // timestep.cpp:61-66
        launch_surface_refresh(stream, surf.p, xn.p, stris.p, stri_off.p, stri_inc.p, snormal.p, S);
    if (check) enqueue_residual(dt);
}

// The rows of a hex group's all-color launches (warm-start displacements, residual). The clusters of
// XCD region x (every color's share, color by color, in runs of at most 64 consecutive clusters of
// one color) fill the 64-row workgroups b with b mod 8 = x, so that a workgroup runs on the XCD
// whose L2 holds its region's particles, as in the per-color launches; a workgroup's record gives
// its run (the kernels read it once per workgroup), unused rows and workgroups are padding. Uploaded
// once; frame graphs hold the pointer.
void sbs_gpu::ensure_cinfo(group_t& g)
{
    if (g.ws_cinfo.p || !g.hex) return;
    uint32_t const nc = g.num_elements();
    // per region, its workgroups: each a run of at most 64 consecutive clusters of one color
    std::vector<std::vector<std::array<uint32_t, 5>>> region(kXcds);
    for (uint32_t c = 0; c < g.ncolors; ++c) {
        uint32_t const cb = g.cluster_color_offsets[c], cn = g.cluster_color_offsets[c + 1] - cb;
        xcd_split sp{};
        if (c < g.xsplit.size())
            sp = g.xsplit[c];
        else
            for (int x = 0; x <= kXcds; ++x) sp.off[x] = uint32_t(uint64_t(cn) * uint32_t(x) / kXcds);
        for (int x = 0; x < kXcds; ++x)
            for (uint32_t j = sp.off[x]; j < sp.off[x + 1]; j += 64)
                region[x].push_back({j, cn, g.color_offsets[c], cb + j, std::min<uint32_t>(64u, sp.off[x + 1] - j)});
    }
    std::size_t per = 1;
    for (auto const& r : region) per = std::max(per, r.size());
    std::size_t const blocks = per * kXcds;
    require(blocks * 64 < 0xffffffffull / 8, "too many clusters for the warm-start rows");
    std::vector<uint4> rec(2 * blocks, make_uint4(0u, 0u, 0u, 0u));
    g.ws_rowof.assign(nc, 0u);
    for (int x = 0; x < kXcds; ++x)
        for (std::size_t q = 0; q < region[x].size(); ++q) {
            auto const& w = region[x][q];
            std::size_t const b = q * kXcds + std::size_t(x);
            rec[2 * b] = make_uint4(w[0], w[1], w[2], w[3]);
            rec[2 * b + 1] = make_uint4(w[4], 0u, 0u, 0u);
            for (uint32_t t = 0; t < w[4]; ++t) g.ws_rowof[w[3] + t] = uint32_t(b * 64 + t);
        }
    g.ws_rows = uint32_t(blocks * 64);
    g.ws_cinfo.upload(rec.data(), rec.size(), stream);
}

void sbs_gpu::ensure_warm_buffers(bool raw)
{
    if (!xt.p) xt.alloc(V);
    std::vector<int8_t> preg;  // particle -> XCD region of the first hex cluster (execution order) holding it
    for (auto& g : groups) {
        if (g.kind != 1) continue;
        if (g.hex && !g.ws_csr.p) {
            // particle -> (cluster, slot) list in cluster execution order (a counting sort)
            uint32_t const nc = g.num_elements();
            std::vector<int4> v0(nc), v1(nc);
            hip_check(hipMemcpy(v0.data(), g.cl_v0.p, nc * sizeof(int4), hipMemcpyDeviceToHost), "read clusters");
            hip_check(hipMemcpy(v1.data(), g.cl_v1.p, nc * sizeof(int4), hipMemcpyDeviceToHost), "read clusters");
            ensure_cinfo(g);
            std::size_t const rows = g.ws_rows;  // slot k of the cluster in row r: disp[k * rows + r]
            std::vector<uint32_t> off(std::size_t(V) + 1, 0), ent(std::size_t(nc) * 8);
            auto slot = [&](uint32_t c, int k) {
                int4 const& a = k < 4 ? v0[c] : v1[c];
                int const j = k & 3;
                return j == 0 ? a.x : (j == 1 ? a.y : (j == 2 ? a.z : a.w));
            };
            for (uint32_t c = 0; c < nc; ++c)
                for (int k = 0; k < 8; ++k) off[std::size_t(slot(c, k)) + 1]++;
            uint32_t widest = 0;
            for (std::size_t p = 0; p < V; ++p) widest = std::max(widest, off[p + 1]);
            for (std::size_t p = 0; p < V; ++p) off[p + 1] += off[p];
            std::vector<uint32_t> fill(off.begin(), off.end() - 1);
            for (uint32_t c = 0; c < nc; ++c)
                for (int k = 0; k < 8; ++k) ent[fill[slot(c, k)]++] = uint32_t(k * rows + g.ws_rowof[c]);
            if (!prow.p) {
                preg.resize(V, -1);
                for (uint32_t c = 0; c < nc; ++c)
                    for (int k = 0; k < 8; ++k) {
                        int8_t& r = preg[std::size_t(slot(c, k))];
                        if (r < 0) r = int8_t((g.ws_rowof[c] / 64) % kXcds);
                    }
            }
            if (widest <= 8) {  // a dense row of 8 per particle, same order (ws_csr_off stays empty)
                std::vector<uint32_t> dense(std::size_t(V) * 8, 0xffffffffu);
                for (std::size_t p = 0; p < V; ++p)
                    for (uint32_t j = off[p]; j < off[p + 1]; ++j) dense[8 * p + (j - off[p])] = ent[j];
                g.ws_csr.upload(dense.data(), dense.size(), stream);
            } else {
                g.ws_csr_off.upload(off.data(), off.size(), stream);
                g.ws_csr.upload(ent.data(), ent.size(), stream);
            }
            g.ws_disp.alloc(rows * 8);
        }
        if (!g.ws_hist.p && g.count) {
            g.ws_hist.alloc(g.count);
            hip_check(hipMemsetAsync(g.ws_hist.p, 0, g.count * sizeof(float2), stream), "clear history");
        }
        if (raw && g.hex && !g.ws_raw.p) {
            g.ws_raw.alloc(std::size_t(g.ws_rows) * mlp::kMaxInputs);
            g.ws_rowof_dev.upload(g.ws_rowof.data(), g.ws_rowof.size(), stream);
        }
    }
    if (!prow.p && !preg.empty()) {
        // per region its particles in index order (particles of no hex cluster by index eighths),
        // in workgroups of kParticleRowBlock rows with workgroup b on region b mod 8
        std::vector<std::vector<uint32_t>> region(kXcds);
        for (std::size_t p = 0; p < V; ++p)
            region[preg[p] >= 0 ? preg[p] : int(p * kXcds / V)].push_back(uint32_t(p));
        std::size_t blocks = 0;
        for (auto const& r : region) blocks = std::max(blocks, (r.size() + kParticleRowBlock - 1) / kParticleRowBlock);
        std::vector<uint32_t> rows(blocks * kXcds * kParticleRowBlock, 0xffffffffu);
        for (int x = 0; x < kXcds; ++x)
            for (std::size_t q = 0; q < region[x].size(); ++q)
                rows[((q / kParticleRowBlock) * kXcds + std::size_t(x)) * kParticleRowBlock + q % kParticleRowBlock] =
                    region[x][q];
        prow.upload(rows.data(), rows.size(), stream);
    }
}

// Buffers of a snapshot slot (allocated outside any graph capture).
void sbs_gpu::prepare_snapshot(int slot)
{
    snap_slot& sn = snaps[slot];
    if (!sn.xn.p) {
        sn.xn.alloc(V);
        sn.vel.alloc(V);
    }
    if (snormal.p && !sn.normal.p) sn.normal.alloc(S);
    sn.lambda.resize(groups.size());
    sn.hist.resize(groups.size());
    for (std::size_t i = 0; i < groups.size(); ++i) {
        group_t const& g = groups[i];
        if (g.lambda.n && !sn.lambda[i].p) sn.lambda[i].alloc(g.lambda.n);
        if (g.ws_hist.p && !sn.hist[i].p) sn.hist[i].alloc(g.ws_hist.n);
    }
}

// Copies of the dynamic state (committed positions, velocities, normals, multipliers and their
// history) in one launch.
void sbs_gpu::enqueue_snapshot(int slot)
{
    snap_slot& sn = snaps[slot];
    copy_spans sp{};
    auto add = [&](void* d, const void* src, std::size_t bytes) {
        if (!bytes) return;
        if (sp.n == kMaxSpans) {
            launch_copy_spans(stream, sp);
            sp = copy_spans{};
        }
        sp.dst[sp.n] = d;
        sp.src[sp.n] = src;
        sp.bytes[sp.n] = bytes;
        ++sp.n;
    };
    require(sn.xn.p && sn.vel.p && (!snormal.p || sn.normal.p) && sn.lambda.size() == groups.size(),
            "snapshot buffers not prepared", SBS_ERR_STATE);
    for (std::size_t i = 0; i < groups.size(); ++i)
        require((!groups[i].lambda.n || sn.lambda[i].p) && (!groups[i].ws_hist.p || sn.hist[i].p),
                "snapshot buffers not prepared", SBS_ERR_STATE);
    add(sn.xn.p, xn.p, V * sizeof(pos4));
    add(sn.vel.p, vel.p, V * sizeof(float4));
    if (snormal.p) add(sn.normal.p, snormal.p, S * sizeof(double4));
    for (std::size_t i = 0; i < groups.size(); ++i) {
        group_t const& g = groups[i];
        if (g.lambda.n) add(sn.lambda[i].p, g.lambda.p, g.lambda.n * sizeof(float));
        if (g.ws_hist.p) add(sn.hist[i].p, g.ws_hist.p, g.ws_hist.n * sizeof(float2));
    }
    launch_copy_spans(stream, sp);
    sn.taken = true;
}

void sbs_gpu::snapshot(int slot)
{
    prepare_snapshot(slot);
    enqueue_snapshot(slot);
}

void sbs_gpu::enqueue_check_snapshot()
{
    copy_spans sp{};
    auto add = [&](void* d, const void* src, std::size_t bytes) {
        require(d != nullptr && sp.n < kMaxSpans, "check buffers not prepared", SBS_ERR_STATE);
        sp.dst[sp.n] = d;
        sp.src[sp.n] = src;
        sp.bytes[sp.n] = bytes;
        ++sp.n;
    };
    add(chk_xn.p, xn.p, V * sizeof(pos4));
    add(chk_vel.p, vel.p, V * sizeof(float4));
    if (snormal.p) add(chk_normal.p, snormal.p, S * sizeof(double4));
    if (ws_mode == SBS_WARM_START_LEARNED)
        for (std::size_t i = 0; i < groups.size(); ++i)
            if (groups[i].ws_hist.p) {
                if (sp.n == kMaxSpans) {
                    launch_copy_spans(stream, sp);
                    sp = copy_spans{};
                }
                add(chk_hist[i].p, groups[i].ws_hist.p, groups[i].ws_hist.n * sizeof(float2));
            }
    launch_copy_spans(stream, sp);
}

void sbs_gpu::restore_check()
{
    hip_check(hipMemcpyAsync(xn.p, chk_xn.p, V * sizeof(pos4), hipMemcpyDeviceToDevice, stream), "restore");
    hip_check(hipMemcpyAsync(xi.p, chk_xn.p, V * sizeof(pos4), hipMemcpyDeviceToDevice, stream), "restore");
    hip_check(hipMemcpyAsync(vel.p, chk_vel.p, V * sizeof(float4), hipMemcpyDeviceToDevice, stream), "restore");
    if (snormal.p)
        hip_check(hipMemcpyAsync(snormal.p, chk_normal.p, S * sizeof(double4), hipMemcpyDeviceToDevice, stream),
                  "restore");
    if (ws_mode == SBS_WARM_START_LEARNED)
        for (std::size_t i = 0; i < groups.size(); ++i)
            if (groups[i].ws_hist.p && chk_hist[i].p)
                hip_check(hipMemcpyAsync(groups[i].ws_hist.p, chk_hist[i].p, groups[i].ws_hist.n * sizeof(float2),
                                         hipMemcpyDeviceToDevice, stream), "restore");
}

void sbs_gpu::restore(int slot)
{
    snap_slot const& sn = snaps[slot];
    require(sn.taken, "no snapshot has been taken in this slot", SBS_ERR_STATE);
    hip_check(hipMemcpyAsync(xn.p, sn.xn.p, V * sizeof(pos4), hipMemcpyDeviceToDevice, stream), "restore");
    hip_check(hipMemcpyAsync(xi.p, sn.xn.p, V * sizeof(pos4), hipMemcpyDeviceToDevice, stream), "restore");
    hip_check(hipMemcpyAsync(vel.p, sn.vel.p, V * sizeof(float4), hipMemcpyDeviceToDevice, stream), "restore");
    if (sn.normal.p)
        hip_check(hipMemcpyAsync(snormal.p, sn.normal.p, S * sizeof(double4), hipMemcpyDeviceToDevice, stream),
                  "restore");
    for (std::size_t i = 0; i < groups.size(); ++i) {
        group_t& g = groups[i];
        if (i < sn.lambda.size() && sn.lambda[i].p)
            hip_check(hipMemcpyAsync(g.lambda.p, sn.lambda[i].p, g.lambda.n * sizeof(float), hipMemcpyDeviceToDevice,
                                     stream), "restore");
        if (g.ws_hist.p && i < sn.hist.size() && sn.hist[i].p)
            hip_check(hipMemcpyAsync(g.ws_hist.p, sn.hist[i].p, g.ws_hist.n * sizeof(float2), hipMemcpyDeviceToDevice,
                                     stream), "restore");
        else if (g.ws_hist.p)  // created (as zeros) after the snapshot was taken: zeros at snapshot time
            hip_check(hipMemsetAsync(g.ws_hist.p, 0, g.ws_hist.n * sizeof(float2), stream), "restore");
    }
}

// sqrt of the summed squared FP64 residuals of every neo-Hookean group (state after the last step).
// Blocks of the check's residual: per hex-cluster neo-Hookean group one per 256 clusters, per other
// neo-Hookean group one per 256 tets.
std::size_t sbs_gpu::residual_blocks() const
{
    std::size_t total = 0;
    for (auto const& g : groups)
        if (g.kind == 1) total += g.hex ? (g.ws_rows + 63) / 64 : (g.count + 255) / 256;
    return total;
}

// Buffers of the per-frame fallback check (allocated outside any graph capture).
void sbs_gpu::prepare_check()
{
    if (!chk_xn.p) {
        chk_xn.alloc(V);
        chk_vel.alloc(V);
    }
    if (snormal.p && !chk_normal.p) chk_normal.alloc(S);
    chk_hist.resize(groups.size());
    for (auto& g : groups)
        if (g.kind == 1) ensure_cinfo(g);
    if (ws_mode == SBS_WARM_START_LEARNED)
        for (std::size_t i = 0; i < groups.size(); ++i)
            if (groups[i].ws_hist.p && !chk_hist[i].p) chk_hist[i].alloc(groups[i].ws_hist.n);
    std::size_t const total = std::max<std::size_t>(residual_blocks(), 1);
    if (check_partial.n < total) {
        drop_graphs();  // a cached frame graph may hold the old buffer
        check_partial.alloc(total);
    }
    if (!resid_host) {
        void* p = nullptr;
        hip_check(hipHostMalloc(&p, sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        resid_host = static_cast<double*>(p);
    }
}

// resid_host[0] = sum over the neo-Hookean groups of sum h^2 at the committed positions: one
// launch per group over all its colors (per cluster for hex groups), then one block sums the
// partial sums in order.
void sbs_gpu::enqueue_residual(double dt)
{
    double const dt2 = dt * dt;
    std::size_t at = 0;
    for (auto const& g : groups) {
        if (g.kind != 1) continue;
        if (g.hex) {
            uint32_t const rows = g.ws_rows;
            launch_neohookean_hex_residual(stream, g.tet(), g.clusters(), xn.p, g.ws_cinfo.p, rows, float(1.0 / dt2),
                                           check_partial.p + at);
            at += (rows + 63) / 64;
        } else {
            launch_neohookean_residual(stream, g.tet(), xn.p, 0, g.count, float(1.0 / dt2), check_partial.p + at, 256);
            at += (g.count + 255) / 256;
        }
    }
    launch_sum_partials(stream, check_partial.p, uint32_t(at), resid_host);
}

void sbs_gpu::run_frame(double dt, uint32_t iterations, uint32_t substeps, int warm, bool check)
{
    bool const capture = cap_net != nullptr && cap_stride > 0;
    if (use_graphs && !timing && !capture) {
        std::array<double, 16> key{};
        key[0] = dt;
        key[1] = iterations;
        key[2] = substeps;
        key[3] = params.gravity[0];
        key[4] = params.gravity[1];
        key[5] = params.gravity[2];
        key[6] = params.gravity_is_force;
        key[7] = params.collision_compliance;
        key[8] = block;
        key[9] = cluster_block;
        key[10] = warm;
        key[11] = double(reinterpret_cast<uintptr_t>(warm == SBS_WARM_START_LEARNED ? ws_net : nullptr));
        key[12] = ws_scale;
        key[13] = surface_refresh;
        key[14] = ws_scale_hyd;
        key[15] = check + 2 * fuse_first_contacts;
        auto hit = std::find_if(graphs.begin(), graphs.end(), [&](auto const& g) { return g.first == key; });
        if (hit != graphs.end()) {
            std::rotate(hit, hit + 1, graphs.end());  // most recently used last
        } else {
            if (graphs.size() == kGraphCache) {
                // the least recently used graph may still be running on the stream
                hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
                (void)hipGraphExecDestroy(graphs.front().second);
                graphs.erase(graphs.begin());
            }
            hipGraph_t graph;
            hipGraphExec_t graph_exec = nullptr;
            hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                enqueue_frame(dt, iterations, substeps, warm, check);
            } catch (...) {
                hipGraph_t g2 = nullptr;
                (void)hipStreamEndCapture(stream, &g2);
                if (g2) (void)hipGraphDestroy(g2);
                throw;
            }
            hip_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
            hipError_t const e = hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            hip_check(e, "hipGraphInstantiate");
            graphs.emplace_back(key, graph_exec);
        }
        hip_check(hipGraphLaunch(graphs.back().second, stream), "hipGraphLaunch");
    } else {
        enqueue_frame(dt, iterations, substeps, warm, check);
        hip_check(hipGetLastError(), "kernel launch");
    }
}

// ---------------------------------------------------------------------------------------------

extern "C" {

const char* sbs_gpu_last_error(void) { return g_last_error.c_str(); }

int sbs_gpu_default_params(sbs_xpbd_params* p)
{
    return guarded([&] {
        require(p != nullptr, "params is null");
        *p = sbs_xpbd_params{};
        p->compliance = 1e-4;
        p->damping = 0.0;
        p->young_modulus = 1e6;
        p->poisson_ratio = 0.30;
        p->hooke_coefficient = 1.0;
        p->collision_compliance = 1e-8;
        p->collision_damping = 0.0;
        p->gravity[0] = 0.0;
        p->gravity[1] = -9.81;
        p->gravity[2] = 0.0;
        p->gravity_is_force = 1;
    });
}

int sbs_gpu_create(int device, sbs_gpu** out)
{
    return guarded([&] {
        require(out != nullptr, "out is null");
        *out = nullptr;
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        require(count > 0, "no HIP device", SBS_ERR_NO_DEVICE);
        require(device >= 0 && device < count, "device index out of range", SBS_ERR_NO_DEVICE);
        auto* s = new sbs_gpu();
        s->device = device;
        try {
            s->set_current();
            hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
            sbs_gpu_default_params(&s->params);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int sbs_gpu_destroy(sbs_gpu* sim)
{
    return guarded([&] {
        if (!sim) return;
        (void)hipSetDevice(sim->device);
        delete sim;
    });
}

int sbs_gpu_set_params(sbs_gpu* sim, const sbs_xpbd_params* params)
{
    return guarded([&] {
        require(sim && params, "null argument");
        sim->params = *params;
    });
}

int sbs_gpu_add_body(sbs_gpu* sim, const double* x0, const double* mass, uint32_t n, uint32_t* body_id)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(x0 != nullptr || n == 0, "x0 is null");
        body_t b;
        b.n = n;
        b.x0.assign(x0, x0 + 3 * std::size_t(n));
        if (mass) {
            b.mass.assign(mass, mass + n);
            for (double m : b.mass) require(m >= 0.0 && std::isfinite(m), "mass must be finite and >= 0");
        }
        for (double v : b.x0) require(std::isfinite(v), "positions must be finite");
        sim->bodies.push_back(std::move(b));
        if (body_id) *body_id = uint32_t(sim->bodies.size() - 1);
    });
}

int sbs_gpu_set_collision_surface(sbs_gpu* sim, uint32_t body, const uint32_t* surface_to_body, uint32_t n)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(body < sim->bodies.size(), "body index out of range");
        require(surface_to_body != nullptr || n == 0, "surface map is null");
        for (uint32_t i = 0; i < n; ++i) require(surface_to_body[i] < sim->bodies[body].n, "surface vertex out of range");
        sim->bodies[body].surface.assign(surface_to_body, surface_to_body + n);
    });
}

int sbs_gpu_set_surface_triangles(sbs_gpu* sim, uint32_t body, const uint32_t* triangles, uint32_t n)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(body < sim->bodies.size(), "body index out of range");
        require(triangles != nullptr || n == 0, "triangles is null");
        std::size_t const ns = sim->bodies[body].surface.size();
        for (std::size_t i = 0; i < 3 * std::size_t(n); ++i)
            require(triangles[i] < ns, "triangle vertex is not a surface vertex (set the collision surface first)");
        sim->bodies[body].triangles.assign(triangles, triangles + 3 * std::size_t(n));
    });
}

static int add_tet_group(sbs_gpu* sim, int kind, uint32_t body, const uint32_t* tets, uint32_t n, double alpha,
                         double beta, double young, double poisson, uint32_t* group_id)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(body < sim->bodies.size(), "body index out of range");
        require(tets != nullptr || n == 0, "tets is null");
        require(poisson > -1.0 && poisson < 0.5, "poisson ratio must be in (-1, 0.5)");
        require(young > 0.0, "young modulus must be positive");
        for (std::size_t i = 0; i < 4 * std::size_t(n); ++i) require(tets[i] < sim->bodies[body].n, "tet vertex out of range");
        group_t g;
        g.kind = kind;
        g.body = body;
        g.count = n;
        g.alpha = alpha;
        g.beta = beta;
        g.young = young;
        g.poisson = poisson;
        g.conn.assign(tets, tets + 4 * std::size_t(n));
        sim->groups.push_back(std::move(g));
        if (group_id) *group_id = uint32_t(sim->groups.size() - 1);
    });
}

static int add_tet_group_batched(sbs_gpu* sim, int kind, const uint32_t* bodies, const uint32_t* tets, uint32_t n,
                                 double alpha, double beta, const double* young, const double* poisson,
                                 uint32_t* group_id)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(n == 0 || (bodies && tets && young && poisson), "null argument");
        for (uint32_t c = 0; c < n; ++c) {
            require(bodies[c] < sim->bodies.size(), "body index out of range");
            require(poisson[c] > -1.0 && poisson[c] < 0.5, "poisson ratio must be in (-1, 0.5)");
            require(young[c] > 0.0, "young modulus must be positive");
            for (int k = 0; k < 4; ++k) require(tets[4 * std::size_t(c) + k] < sim->bodies[bodies[c]].n, "tet vertex out of range");
        }
        group_t g;
        g.kind = kind;
        g.body = n ? bodies[0] : 0;
        g.count = n;
        g.alpha = alpha;
        g.beta = beta;
        g.young = n ? young[0] : 1.0;
        g.poisson = n ? poisson[0] : 0.0;
        g.conn.assign(tets, tets + 4 * std::size_t(n));
        g.cbody.assign(bodies, bodies + n);
        g.cyoung.assign(young, young + n);
        g.cpoisson.assign(poisson, poisson + n);
        sim->groups.push_back(std::move(g));
        if (group_id) *group_id = uint32_t(sim->groups.size() - 1);
    });
}

int sbs_gpu_add_green_constraints_batched(sbs_gpu* sim, const uint32_t* bodies, const uint32_t* tets, uint32_t n,
                                          double alpha, double beta, const double* young_modulus,
                                          const double* poisson_ratio, uint32_t* group_id)
{
    return add_tet_group_batched(sim, 0, bodies, tets, n, alpha, beta, young_modulus, poisson_ratio, group_id);
}

int sbs_gpu_add_neohookean_constraints_batched(sbs_gpu* sim, const uint32_t* bodies, const uint32_t* tets, uint32_t n,
                                               double beta, const double* young_modulus, const double* poisson_ratio,
                                               uint32_t* group_id)
{
    return add_tet_group_batched(sim, 1, bodies, tets, n, 0.0, beta, young_modulus, poisson_ratio, group_id);
}

int sbs_gpu_add_green_constraints(sbs_gpu* sim, uint32_t body, const uint32_t* tets, uint32_t n, double alpha,
                                  double beta, double young_modulus, double poisson_ratio, uint32_t* group_id)
{
    return add_tet_group(sim, 0, body, tets, n, alpha, beta, young_modulus, poisson_ratio, group_id);
}

int sbs_gpu_add_neohookean_constraints(sbs_gpu* sim, uint32_t body, const uint32_t* tets, uint32_t n, double beta,
                                       double young_modulus, double poisson_ratio, uint32_t* group_id)
{
    return add_tet_group(sim, 1, body, tets, n, 0.0, beta, young_modulus, poisson_ratio, group_id);
}

int sbs_gpu_add_distance_constraints(sbs_gpu* sim, const uint32_t* e, uint32_t n, double alpha, double beta,
                                     uint32_t* group_id)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(e != nullptr || n == 0, "constraint array is null");
        for (uint32_t i = 0; i < n; ++i) {
            require(e[4 * i] < sim->bodies.size() && e[4 * i + 2] < sim->bodies.size(), "body index out of range");
            require(e[4 * i + 1] < sim->bodies[e[4 * i]].n && e[4 * i + 3] < sim->bodies[e[4 * i + 2]].n,
                    "vertex index out of range");
            require(!(e[4 * i] == e[4 * i + 2] && e[4 * i + 1] == e[4 * i + 3]), "distance constraint joins a particle to itself");
        }
        group_t g;
        g.kind = 2;
        g.count = n;
        g.alpha = alpha;
        g.beta = beta;
        g.conn.assign(e, e + 4 * std::size_t(n));
        sim->groups.push_back(std::move(g));
        if (group_id) *group_id = uint32_t(sim->groups.size() - 1);
    });
}

int sbs_gpu_add_plane(sbs_gpu* sim, const double normal[3], const double point[3], const double vol_min[3],
                      const double vol_max[3])
{
    return guarded([&] {
        require(sim && normal && point && vol_min && vol_max, "null argument");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        sdf_dev f{};
        f.kind = 0;
        double const off = -(normal[0] * point[0] + normal[1] * point[1] + normal[2] * point[2]);
        for (int i = 0; i < 3; ++i) {
            f.a[i] = normal[i];
            f.vmin[i] = vol_min[i];
            f.vmax[i] = vol_max[i];
        }
        f.a[3] = off;
        sim->sdfs.push_back(f);
    });
}

int sbs_gpu_add_sphere(sbs_gpu* sim, const double center[3], double radius)
{
    return guarded([&] {
        require(sim && center, "null argument");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(radius > 0.0, "radius must be positive");
        sdf_dev f{};
        f.kind = 1;
        for (int i = 0; i < 3; ++i) {
            f.a[i] = center[i];
            f.vmin[i] = center[i] - radius;
            f.vmax[i] = center[i] + radius;
        }
        f.a[3] = radius;
        sim->sdfs.push_back(f);
    });
}

namespace {

sdf_dev grid_record(const double dmin[3], const double dmax[3], const uint32_t res[3], const double* dev_nodes,
                    const double vol_min[3], const double vol_max[3])
{
    sdf_dev f{};
    f.kind = 2;
    grid_cell_size(dmin, dmax, res, f.cell, f.inv_cell);
    for (int d = 0; d < 3; ++d) {
        f.res[d] = res[d];
        f.gmin[d] = dmin[d];
        f.gmax[d] = dmax[d];
        f.vmin[d] = vol_min[d];
        f.vmax[d] = vol_max[d];
    }
    f.nodes = dev_nodes;
    return f;
}

}  // namespace

int sbs_gpu_add_grid_sdf(sbs_gpu* sim, const double domain_min[3], const double domain_max[3],
                         const uint32_t resolution[3], const double* nodes, const double vol_min[3],
                         const double vol_max[3])
{
    return guarded([&] {
        require(sim && nodes && vol_min && vol_max, "null argument");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        check_grid(domain_min, domain_max, resolution);
        uint64_t const n = grid_dims(resolution).nodes();
        auto buf = std::make_unique<dev_buf<double>>();
        buf->upload(nodes, n, sim->stream);
        hip_check(hipStreamSynchronize(sim->stream), "upload grid nodes");
        sim->sdfs.push_back(grid_record(domain_min, domain_max, resolution, buf->p, vol_min, vol_max));
        sim->grid_nodes.push_back(std::move(buf));
    });
}

int sbs_gpu_add_mesh_obstacle(sbs_gpu* sim, const double* positions, uint32_t num_vertices, const uint32_t* triangles,
                              uint32_t num_triangles, const double domain_min[3], const double domain_max[3],
                              const uint32_t resolution[3])
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(domain_min && domain_max, "null domain");
        require(num_triangles > 0, "the mesh has no triangles");
        double lo[3], hi[3];
        environment_domain(positions, num_vertices, domain_min, domain_max, lo, hi);
        check_grid(lo, hi, resolution);
        auto buf = std::make_unique<dev_buf<double>>();
        buf->alloc(grid_dims(resolution).nodes());
        build_mesh_grid(sim->stream, positions, num_vertices, triangles, num_triangles, lo, hi, resolution, buf->p);
        sim->sdfs.push_back(grid_record(lo, hi, resolution, buf->p, lo, hi));
        sim->grid_nodes.push_back(std::move(buf));
    });
}

int sbs_gpu_evaluate_sdf(sbs_gpu* sim, uint32_t sdf, const double* points, uint32_t n, double* sd, double* grad)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sdf < sim->sdfs.size(), "sdf index out of range");
        require((points && sd && grad) || n == 0, "null argument");
        evaluate_sdf(sim->stream, sim->sdfs[sdf], points, n, sd, grad);
    });
}

int sbs_gpu_set_group_clusters(sbs_gpu* sim, uint32_t group, const uint32_t* cluster_offsets, uint32_t num_clusters)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(group < sim->groups.size(), "group index out of range");
        group_t& g = sim->groups[group];
        require(g.kind != 2, "clusters are supported for tet constraint groups");
        if (num_clusters == 0) {
            require(g.count == 0 || cluster_offsets == nullptr, "a non-empty group needs at least one cluster");
            g.cluster_off.clear();
            return;
        }
        require(cluster_offsets != nullptr, "cluster_offsets is null");
        require(cluster_offsets[0] == 0 && cluster_offsets[num_clusters] == g.count,
                "cluster offsets must start at 0 and end at the group's constraint count");
        for (uint32_t i = 0; i < num_clusters; ++i)
            require(cluster_offsets[i + 1] > cluster_offsets[i], "clusters must be non-empty and increasing");
        g.cluster_off.assign(cluster_offsets, cluster_offsets + num_clusters + 1);
    });
}

int sbs_gpu_set_coloring(sbs_gpu* sim, int order)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(order == SBS_COLORING_HASH || order == SBS_COLORING_INDEX, "unknown coloring order");
        sim->coloring_order = order;
    });
}

int sbs_gpu_set_contact_fusion(sbs_gpu* sim, int enable)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        sim->fuse_first_contacts = enable != 0;
    });
}

int sbs_gpu_set_launch_config(sbs_gpu* sim, int block, int cluster_block)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(block == 64 || block == 128 || block == 256, "block must be 64, 128 or 256");
        require(cluster_block == 64 || cluster_block == 128 || cluster_block == SBS_CLUSTER_BLOCK_STREAMED,
                "cluster_block must be 64, 128 or SBS_CLUSTER_BLOCK_STREAMED");
        sim->block = block;
        sim->cluster_block = cluster_block;
    });
}

int sbs_gpu_commit(sbs_gpu* sim, uint32_t recolor_passes)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(recolor_passes <= 64, "recolor_passes must be <= 64");
        sim->commit(recolor_passes);
    });
}

int sbs_gpu_step(sbs_gpu* sim, double dt, uint32_t iterations, uint32_t substeps)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        require(dt > 0.0 && std::isfinite(dt), "dt must be positive");
        require(substeps >= 1, "substeps must be >= 1");
        sim->check_overflow();
        require(!(sim->cap_net && sim->cap_stride) || sim->ws_mode != SBS_WARM_START_NONE,
                "capture needs a warm start: its pass forms the features", SBS_ERR_STATE);
        sim->set_current();
        bool const check = sim->fb_limit > 0.0 && sim->ws_mode != SBS_WARM_START_NONE;
        // with the check the frame itself snapshots the state and ends with its residual
        if (check) sim->prepare_check();
        sim->run_frame(dt, iterations, substeps, sim->ws_mode, check);
        if (check) {
            hip_check(hipStreamSynchronize(sim->stream), "fallback check");
            double const r = std::sqrt(*sim->resid_host);
            // !(r <= limit): a frame whose residual is NaN (one that diverged) is repeated as well
            if (!(r <= sim->fb_limit)) {
                sim->restore_check();
                sim->run_frame(dt, sim->fb_iterations ? sim->fb_iterations : iterations, substeps, SBS_WARM_START_NONE);
                ++sim->fb_count;
            }
        }
        if (sim->cap_net && sim->cap_stride) hip_check(hipStreamSynchronize(sim->stream), "capture");
    });
}

int sbs_gpu_synchronize(sbs_gpu* sim)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        hip_check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
        sim->check_overflow();
    });
}

int sbs_gpu_use_graphs(sbs_gpu* sim, int enable)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        sim->use_graphs = enable != 0;
    });
}

int sbs_gpu_read_body(sbs_gpu* sim, uint32_t body, double* x, double* v)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        require(body < sim->bodies.size(), "body index out of range");
        body_t const& b = sim->bodies[body];
        std::vector<pos4> hx(b.n);
        std::vector<float4> hv(b.n);
        if (b.n) {
            hip_check(hipMemcpyAsync(hx.data(), sim->xn.p + b.offset, b.n * sizeof(pos4), hipMemcpyDeviceToHost,
                                     sim->stream), "read positions");
            hip_check(hipMemcpyAsync(hv.data(), sim->vel.p + b.offset, b.n * sizeof(float4), hipMemcpyDeviceToHost,
                                     sim->stream), "read velocities");
        }
        hip_check(hipStreamSynchronize(sim->stream), "read body");
        for (uint32_t i = 0; i < b.n; ++i) {
            if (x) {
                x[3 * i] = hx[i].x;
                x[3 * i + 1] = hx[i].y;
                x[3 * i + 2] = hx[i].z;
            }
            if (v) {
                v[3 * i] = hv[i].x;
                v[3 * i + 1] = hv[i].y;
                v[3 * i + 2] = hv[i].z;
            }
        }
    });
}

int sbs_gpu_write_body(sbs_gpu* sim, uint32_t body, const double* x, const double* v)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        require(body < sim->bodies.size(), "body index out of range");
        body_t const& b = sim->bodies[body];
        if (b.n == 0) return;
        if (x) {
            std::vector<pos4> hx(b.n);
            hip_check(hipMemcpyAsync(hx.data(), sim->xn.p + b.offset, b.n * sizeof(pos4), hipMemcpyDeviceToHost,
                                     sim->stream), "read positions");
            hip_check(hipStreamSynchronize(sim->stream), "write body");
            for (uint32_t i = 0; i < b.n; ++i) {
                require(std::isfinite(x[3 * i]) && std::isfinite(x[3 * i + 1]) && std::isfinite(x[3 * i + 2]),
                        "positions must be finite");
                hx[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], hx[i].w);
            }
            hip_check(hipMemcpyAsync(sim->xn.p + b.offset, hx.data(), b.n * sizeof(pos4), hipMemcpyHostToDevice,
                                     sim->stream), "write positions");
            hip_check(hipMemcpyAsync(sim->xi.p + b.offset, hx.data(), b.n * sizeof(pos4), hipMemcpyHostToDevice,
                                     sim->stream), "write positions");
            hip_check(hipStreamSynchronize(sim->stream), "write body");
        }
        if (v) {
            std::vector<float4> hv(b.n);
            for (uint32_t i = 0; i < b.n; ++i) hv[i] = make_float4(float(v[3 * i]), float(v[3 * i + 1]), float(v[3 * i + 2]), 0.f);
            hip_check(hipMemcpyAsync(sim->vel.p + b.offset, hv.data(), b.n * sizeof(float4), hipMemcpyHostToDevice,
                                     sim->stream), "write velocities");
            hip_check(hipStreamSynchronize(sim->stream), "write body");
        }
    });
}

int sbs_gpu_group_colors(sbs_gpu* sim, uint32_t group, uint32_t* num_colors, uint32_t* color_of_constraint)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        require(group < sim->groups.size(), "group index out of range");
        group_t const& g = sim->groups[group];
        if (num_colors) *num_colors = g.ncolors;
        if (color_of_constraint && !g.color.empty())
            std::memcpy(color_of_constraint, g.color.data(), g.color.size() * sizeof(uint32_t));
    });
}

int sbs_util_exclusive_scan(int device, const uint32_t* in, uint64_t n, uint32_t* out)
{
    return guarded([&] {
        require(n == 0 || (in != nullptr && out != nullptr), "null argument");
        require(out != nullptr, "out is null");
        hip_check(hipSetDevice(device), "hipSetDevice");
        hipStream_t s;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        try {
            dev_buf<uint32_t> din, dout, ws;
            din.upload(in, n, s);
            dout.alloc(n + 1);
            ws.alloc(scan_workspace(n));
            exclusive_scan(s, din.p, dout.p, n, ws.p);
            hip_check(hipGetLastError(), "scan");
            hip_check(hipMemcpyAsync(out, dout.p, (n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read scan");
            hip_check(hipStreamSynchronize(s), "scan");
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        hip_check(hipStreamDestroy(s), "hipStreamDestroy");
    });
}

int sbs_util_sort_pairs(int device, const uint32_t* keys, const uint32_t* vals, uint64_t n, int bits,
                        uint32_t* out_keys, uint32_t* out_vals)
{
    return guarded([&] {
        require(n == 0 || (keys && vals && out_keys && out_vals), "null argument");
        require(bits >= 0 && bits <= 32, "bits must be 0..32");
        require(n < (uint64_t(1) << 31), "too many pairs");
        hip_check(hipSetDevice(device), "hipSetDevice");
        if (n == 0) return;
        hipStream_t s;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        try {
            dev_buf<uint32_t> dk, dv, ws;
            dk.upload(keys, n, s);
            dv.upload(vals, n, s);
            ws.alloc(sort_workspace(n));
            radix_sort_pairs(s, dk.p, dv.p, n, bits, ws.p);
            hip_check(hipGetLastError(), "sort");
            hip_check(hipMemcpyAsync(out_keys, dk.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read sort");
            hip_check(hipMemcpyAsync(out_vals, dv.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read sort");
            hip_check(hipStreamSynchronize(s), "sort");
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        hip_check(hipStreamDestroy(s), "hipStreamDestroy");
    });
}

int sbs_gpu_set_surface_refresh(sbs_gpu* sim, int enable)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        sim->surface_refresh = enable != 0;
    });
}

namespace {

void read_surface(sbs_gpu* sim, uint32_t body, double* positions, double* normals)
{
    require(sim != nullptr, "sim is null");
    require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
    require(body < sim->bodies.size(), "body index out of range");
    uint32_t base = 0;
    for (uint32_t b = 0; b < body; ++b) base += uint32_t(sim->bodies[b].surface.size());
    body_t const& bd = sim->bodies[body];
    std::size_t const n = bd.surface.size();
    if (n == 0) return;
    if (positions) {
        std::vector<pos4> x(bd.n);
        hip_check(hipMemcpyAsync(x.data(), sim->xn.p + bd.offset, bd.n * sizeof(pos4), hipMemcpyDeviceToHost,
                                 sim->stream), "read surface");
        hip_check(hipStreamSynchronize(sim->stream), "read surface");
        for (std::size_t i = 0; i < n; ++i) {
            pos4 const& p = x[bd.surface[i]];
            positions[3 * i] = p.x;
            positions[3 * i + 1] = p.y;
            positions[3 * i + 2] = p.z;
        }
    }
    if (normals) {
        require(sim->snormal.p != nullptr, "the body has no surface triangles", SBS_ERR_STATE);
        std::vector<double4> nv(n);
        hip_check(hipMemcpyAsync(nv.data(), sim->snormal.p + base, n * sizeof(double4), hipMemcpyDeviceToHost,
                                 sim->stream), "read normals");
        hip_check(hipStreamSynchronize(sim->stream), "read normals");
        for (std::size_t i = 0; i < n; ++i) {
            normals[3 * i] = nv[i].x;
            normals[3 * i + 1] = nv[i].y;
            normals[3 * i + 2] = nv[i].z;
        }
    }
}

}  // namespace

int sbs_gpu_read_surface(sbs_gpu* sim, uint32_t body, double* positions, double* normals)
{
    return guarded([&] { read_surface(sim, body, positions, normals); });
}

int sbs_gpu_write_surface_ply(sbs_gpu* sim, uint32_t body, const char* path, int format)
{
    return guarded([&] {
        require(path != nullptr, "path is null");
        read_surface(sim, body, nullptr, nullptr);  // argument and state checks
        body_t const& bd = sim->bodies[body];
        require(!bd.triangles.empty(), "the body has no surface triangles (sbs_gpu_set_surface_triangles)",
                SBS_ERR_STATE);
        std::size_t const n = bd.surface.size();
        std::vector<double> x(3 * n), nv;
        bool const with_normals = sim->surface_refresh;
        if (with_normals) nv.resize(3 * n);
        read_surface(sim, body, x.data(), with_normals ? nv.data() : nullptr);
        // the surface mesh holds float positions and normals (common::geometry_t)
        std::vector<float> xf(x.begin(), x.end()), nf(nv.begin(), nv.end());
        std::vector<int32_t> tri(bd.triangles.begin(), bd.triangles.end());
        sbs_geometry g{};
        g.positions = xf.data();
        g.num_positions = xf.size();
        g.indices = tri.data();
        g.num_indices = tri.size();
        g.normals = nf.empty() ? nullptr : nf.data();
        g.num_normals = nf.size();
        g.geometry_type = SBS_GEOMETRY_TRIANGLE;
        int const code = sbs_ply_write(path, &g, format);
        if (code != SBS_OK) throw api_error(code, sbs_gpu_last_error());
    });
}

int sbs_gpu_set_contact_hash(sbs_gpu* sim, int mode, double cell_size)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "scene already committed", SBS_ERR_STATE);
        require(mode == SBS_CONTACT_HASH_AUTO || mode == SBS_CONTACT_HASH_ON || mode == SBS_CONTACT_HASH_OFF,
                "unknown contact hash mode");
        require(cell_size >= 0.0 && std::isfinite(cell_size), "cell_size must be finite and >= 0");
        sim->hash_mode = mode;
        sim->hash_cell_size = cell_size;
    });
}

int sbs_gpu_read_contact_hash(sbs_gpu* sim, double* cell_size, uint32_t* table_bits, uint32_t* start, uint32_t* end,
                              uint32_t* entries, uint32_t cap, uint32_t* num_entries, uint32_t* global,
                              uint32_t* num_global)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        if (cell_size) *cell_size = sim->use_hash ? sim->hash_h : 0.0;
        if (table_bits) *table_bits = sim->use_hash ? uint32_t(sim->hash_bits) : 0u;
        if (num_global) *num_global = uint32_t(sim->global_ids.size());
        if (global) std::copy(sim->global_ids.begin(), sim->global_ids.end(), global);
        uint32_t n = 0;
        if (sim->use_hash) {
            hip_check(hipMemcpyAsync(&n, sim->hcount.p, sizeof(uint32_t), hipMemcpyDeviceToHost, sim->stream), "read hash");
            std::size_t const T = std::size_t(1) << sim->hash_bits;
            if (start)
                hip_check(hipMemcpyAsync(start, sim->hstart.p, T * sizeof(uint32_t), hipMemcpyDeviceToHost, sim->stream),
                          "read hash");
            if (end)
                hip_check(hipMemcpyAsync(end, sim->hend.p, T * sizeof(uint32_t), hipMemcpyDeviceToHost, sim->stream),
                          "read hash");
            hip_check(hipStreamSynchronize(sim->stream), "read hash");
            if (entries && n)
                hip_check(hipMemcpy(entries, sim->hentries.p, std::min(n, cap) * sizeof(uint32_t), hipMemcpyDeviceToHost),
                          "read hash");
        }
        if (num_entries) *num_entries = n;
    });
}

int sbs_gpu_set_body_contacts(sbs_gpu* sim, int enable, double cell_size, double thickness, double radius)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(!sim->committed, "body contacts are set before sbs_gpu_commit", SBS_ERR_STATE);
        require(std::isfinite(cell_size) && cell_size >= 0.0 && std::isfinite(thickness) && thickness >= 0.0 &&
                    std::isfinite(radius) && radius >= 0.0,
                "cell_size, thickness and radius must be finite and >= 0");
        sim->body_contacts = enable != 0;
        sim->body_cell_request = cell_size;
        sim->body_thickness_request = thickness;
        sim->body_radius_request = radius;
    });
}

int sbs_gpu_read_body_contact_setup(sbs_gpu* sim, double* cell_size, double* thickness, double* radius,
                                    uint32_t* table_bits, uint32_t* triangles)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        if (cell_size) *cell_size = sim->btris.n ? sim->body_cell : 0.0;
        if (thickness) *thickness = sim->btris.n ? sim->body_thickness : 0.0;
        if (radius) *radius = sim->btris.n ? sim->body_radius : 0.0;
        if (table_bits) *table_bits = uint32_t(sim->btris.n ? sim->body_bits : 0);
        if (triangles && sim->btris.n)
            hip_check(hipMemcpy(triangles, sim->btris.p, sim->btris.n * sizeof(uint4), hipMemcpyDeviceToHost),
                      "read triangles");
    });
}

int sbs_gpu_read_body_contact_candidates(sbs_gpu* sim, uint32_t* pairs, uint64_t cap, uint64_t* count)
{
    return guarded([&] {
        require(sim != nullptr && count != nullptr, "null argument");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        sim->set_current();
        *count = 0;
        if (!sim->btris.n) return;
        body_contact_dev const b = sim->body_view();
        contact_dev const c = sim->contacts();
        build_body_hash(sim->stream, b, sim->surf.p, sim->xn.p, sim->body_bits, sim->bws.p);
        dev_buf<uint32_t> counts, offs, ws;
        counts.alloc(sim->S);
        offs.alloc(std::size_t(sim->S) + 1);
        ws.alloc(scan_workspace(sim->S));
        launch_body_candidates(sim->stream, c, b, sim->xn.p, nullptr, counts.p, nullptr);
        exclusive_scan(sim->stream, counts.p, offs.p, sim->S, ws.p);
        uint32_t total = 0;
        hip_check(hipMemcpyAsync(&total, offs.p + sim->S, sizeof(uint32_t), hipMemcpyDeviceToHost, sim->stream),
                  "read candidates");
        hip_check(hipStreamSynchronize(sim->stream), "read candidates");
        *count = total;  // with repeats; the canonical count follows when pairs are read
        if (!total) return;
        dev_buf<uint32_t> out;
        out.alloc(2 * std::size_t(total));
        launch_body_candidates(sim->stream, c, b, sim->xn.p, offs.p, counts.p, out.p);
        std::vector<uint64_t> raw(total);
        {
            std::vector<uint32_t> h(2 * std::size_t(total));
            hip_check(hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, sim->stream),
                      "read candidates");
            hip_check(hipStreamSynchronize(sim->stream), "read candidates");
            for (std::size_t i = 0; i < total; ++i) raw[i] = (uint64_t(h[2 * i]) << 32) | h[2 * i + 1];
        }
        // the canonical list: (vertex, triangle) increasing, each pair once
        std::sort(raw.begin(), raw.end());
        raw.erase(std::unique(raw.begin(), raw.end()), raw.end());
        *count = raw.size();
        std::size_t const n = std::min<uint64_t>(raw.size(), cap);
        for (std::size_t i = 0; i < n; ++i) {
            pairs[2 * i] = uint32_t(raw[i] >> 32);
            pairs[2 * i + 1] = uint32_t(raw[i]);
        }
    });
}

int sbs_gpu_contact_overflow(sbs_gpu* sim, uint64_t* dropped)
{
    return guarded([&] {
        require(sim != nullptr && dropped != nullptr, "null argument");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        uint32_t n = 0;
        if (sim->coverflow.p) {
            hip_check(hipMemcpyAsync(&n, sim->coverflow.p, sizeof(uint32_t), hipMemcpyDeviceToHost, sim->stream),
                      "read overflow");
            hip_check(hipStreamSynchronize(sim->stream), "read overflow");
        }
        *dropped = n;
    });
}

int sbs_gpu_read_contacts(sbs_gpu* sim, uint32_t* ids, double* qn, uint32_t cap, uint32_t* count)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        require(count != nullptr, "count is null");
        *count = 0;
        if (!(sim->S && sim->slots)) return;
        sim->detect();
        std::vector<uint32_t> cnt(sim->S);
        std::vector<pos4> q(std::size_t(sim->S) * sim->slots);
        std::vector<float4> n(q.size());
        hip_check(hipMemcpyAsync(cnt.data(), sim->ccount.p, sim->S * sizeof(uint32_t), hipMemcpyDeviceToHost, sim->stream), "read contacts");
        hip_check(hipMemcpyAsync(q.data(), sim->cq.p, q.size() * sizeof(pos4), hipMemcpyDeviceToHost, sim->stream), "read contacts");
        hip_check(hipMemcpyAsync(n.data(), sim->cn.p, n.size() * sizeof(float4), hipMemcpyDeviceToHost, sim->stream), "read contacts");
        hip_check(hipStreamSynchronize(sim->stream), "read contacts");
        if (sim->timing) sim->harvest();
        uint32_t k = 0, s = 0;
        for (uint32_t bi = 0; bi < sim->bodies.size(); ++bi) {
            body_t const& b = sim->bodies[bi];
            for (uint32_t j = 0; j < b.surface.size(); ++j, ++s) {
                for (uint32_t c = 0; c < cnt[s]; ++c, ++k) {
                    if (k >= cap) continue;
                    std::size_t const slot = std::size_t(c) * sim->S + s;
                    if (ids) {
                        ids[4 * std::size_t(k)] = bi;
                        ids[4 * std::size_t(k) + 1] = j;
                        ids[4 * std::size_t(k) + 2] = b.surface[j];
                        ids[4 * std::size_t(k) + 3] = uint32_t(q[slot].w);
                    }
                    if (qn) {
                        double* o = qn + 6 * std::size_t(k);
                        o[0] = q[slot].x;
                        o[1] = q[slot].y;
                        o[2] = q[slot].z;
                        o[3] = n[slot].x;
                        o[4] = n[slot].y;
                        o[5] = n[slot].z;
                    }
                }
            }
        }
        *count = k;
    });
}

int sbs_gpu_residual(sbs_gpu* sim, uint32_t group, double dt, double* per_color, uint32_t cap)
{
    return guarded([&] {
        require(sim != nullptr && per_color != nullptr, "null argument");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        require(group < sim->groups.size(), "group index out of range");
        require(dt > 0.0, "dt must be positive");
        group_t const& g = sim->groups[group];
        require(cap >= g.ncolors, "per_color has fewer entries than the group has colors");
        double const dt2 = dt * dt;
        for (uint32_t c = 0; c < cap; ++c) per_color[c] = 0.0;
        for (uint32_t col = 0; col < g.ncolors; ++col) {
            uint32_t const b = g.color_offsets[col], n = g.color_offsets[col + 1] - b;
            uint32_t const blocks = (n + 255) / 256;
            if (sim->partial.n < blocks) sim->partial.alloc(std::max<std::size_t>(blocks, 1024));
            if (g.kind == 0)
                launch_green_residual(sim->stream, g.tet(), sim->xn.p, b, n, float(g.alpha / dt2), sim->partial.p, 256);
            else if (g.kind == 1)
                launch_neohookean_residual(sim->stream, g.tet(), sim->xn.p, b, n, float(1.0 / dt2), sim->partial.p, 256);
            else
                launch_distance_residual(sim->stream, g.dist(), sim->xn.p, b, n, float(g.alpha / dt2), sim->partial.p, 256);
            std::vector<double> h(blocks);
            hip_check(hipMemcpyAsync(h.data(), sim->partial.p, blocks * sizeof(double), hipMemcpyDeviceToHost, sim->stream),
                      "read residual");
            hip_check(hipStreamSynchronize(sim->stream), "residual");
            double acc = 0.0;
            for (double x : h) acc += x;
            per_color[col] = acc;
        }
    });
}

int sbs_gpu_energy(sbs_gpu* sim, double out[6])
{
    return guarded([&] {
        require(sim != nullptr && out != nullptr, "null argument");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        for (int i = 0; i < 6; ++i) out[i] = 0.0;
        auto run = [&](uint32_t n, auto&& launch) {
            if (!n) return;
            uint32_t const blocks = (n + 255) / 256;
            if (sim->partial.n < std::size_t(blocks) * 6) sim->partial.alloc(std::max<std::size_t>(std::size_t(blocks) * 6, 1024));
            launch();
            std::vector<double> h(std::size_t(blocks) * 6);
            hip_check(hipMemcpyAsync(h.data(), sim->partial.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, sim->stream),
                      "read energy");
            hip_check(hipStreamSynchronize(sim->stream), "energy");
            for (std::size_t i = 0; i < h.size(); ++i) out[i % 6] += h[i];
        };
        sbs_xpbd_params const& p = sim->params;
        run(sim->V, [&] {
            launch_particle_energy(sim->stream, sim->xn.p, sim->vel.p, sim->mass.p, sim->V, float(p.gravity[0]),
                                   float(p.gravity[1]), float(p.gravity[2]), p.gravity_is_force, sim->partial.p, 256);
        });
        for (auto const& g : sim->groups) {
            if (g.kind == 0) run(g.count, [&] { launch_green_energy(sim->stream, g.tet(), sim->xn.p, g.count, sim->partial.p, 256); });
            if (g.kind == 1)
                run(g.count, [&] { launch_neohookean_energy(sim->stream, g.tet(), sim->xn.p, g.count, sim->partial.p, 256); });
        }
    });
}

int sbs_gpu_set_timing(sbs_gpu* sim, int enable)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        if (sim->timing && !enable) sim->harvest();
        sim->timing = enable != 0;
    });
}

int sbs_gpu_kernel_stats(sbs_gpu* sim, int kind, uint64_t* launches, double* total_ms, uint64_t* units)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(kind >= 0 && kind < SBS_KERNEL_COUNT, "kernel kind out of range");
        sim->harvest();
        if (launches) *launches = sim->st_launches[kind];
        if (total_ms) *total_ms = sim->st_ms[kind];
        if (units) *units = sim->st_units[kind];
    });
}

int sbs_gpu_reset_stats(sbs_gpu* sim)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        sim->harvest();
        for (int k = 0; k < SBS_KERNEL_COUNT; ++k) {
            sim->st_launches[k] = 0;
            sim->st_ms[k] = 0;
            sim->st_units[k] = 0;
        }
    });
}

int sbs_gpu_set_warm_start(sbs_gpu* sim, int mode, sbs_mlp* net)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        require(mode >= SBS_WARM_START_NONE && mode <= SBS_WARM_START_EXPLICIT, "unknown warm start mode");
        if (mode == SBS_WARM_START_LEARNED) {
            require(net != nullptr, "the learned warm start needs a network");
            require(net->nin == uint32_t(kWarmStartInputs) && net->nout == uint32_t(kWarmStartOutputs),
                    "the warm-start network must have 39 inputs and 12 outputs");
            require(sim->all_nh_hex(), "the learned warm start needs hex-cell clusters in every neo-Hookean group",
                    SBS_ERR_STATE);
            require(net->device == sim->device, "network and simulation are on different devices");
        }
        sim->set_current();
        // a cached frame graph holds the network's buffers by address: a network created where a
        // destroyed one lived must not replay the old graph
        if (mode == SBS_WARM_START_LEARNED) sim->drop_graphs();
        sim->ensure_warm_buffers(false);
        sim->ws_mode = mode;
        sim->ws_net = mode == SBS_WARM_START_LEARNED ? net : nullptr;
        hip_check(hipStreamSynchronize(sim->stream), "warm start");
    });
}

int sbs_gpu_capture(sbs_gpu* sim, sbs_mlp* net, uint32_t stride)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        sim->set_current();
        if (stride == 0) {
            hip_check(hipStreamSynchronize(sim->stream), "capture");
            sim->cap_net = nullptr;
            sim->cap_stride = 0;
            return;
        }
        require(net != nullptr, "net is null");
        require(net->nin == uint32_t(kWarmStartInputs) && net->nout == uint32_t(kWarmStartOutputs),
                "the warm-start network must have 39 inputs and 12 outputs");
        require(sim->all_nh_hex(), "capture needs hex-cell clusters in every neo-Hookean group", SBS_ERR_STATE);
        require(net->device == sim->device, "network and simulation are on different devices");
        sim->ensure_warm_buffers(true);
        hip_check(hipStreamSynchronize(net->stream), "capture");
        sim->cap_net = net;
        sim->cap_stride = stride;
    });
}

int sbs_gpu_set_warm_start_scale(sbs_gpu* sim, double scale)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(std::isfinite(scale) && scale >= 0.0, "scale must be finite and >= 0");
        sim->ws_scale = scale;
        sim->ws_scale_hyd = scale;
    });
}

int sbs_gpu_set_warm_start_scales(sbs_gpu* sim, double deviatoric, double hydrostatic)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(std::isfinite(deviatoric) && deviatoric >= 0.0 && std::isfinite(hydrostatic) && hydrostatic >= 0.0,
                "scales must be finite and >= 0");
        sim->ws_scale = deviatoric;
        sim->ws_scale_hyd = hydrostatic;
    });
}

int sbs_gpu_snapshot(sbs_gpu* sim) { return sbs_gpu_snapshot_slot(sim, 0); }

int sbs_gpu_restore(sbs_gpu* sim) { return sbs_gpu_restore_slot(sim, 0); }

int sbs_gpu_snapshot_slot(sbs_gpu* sim, int slot)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        require(slot >= 0 && slot < SBS_SNAPSHOT_SLOTS, "snapshot slot out of range");
        sim->set_current();
        sim->snapshot(slot);
    });
}

int sbs_gpu_restore_slot(sbs_gpu* sim, int slot)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(sim->committed, "sbs_gpu_commit has not been called", SBS_ERR_STATE);
        require(slot >= 0 && slot < SBS_SNAPSHOT_SLOTS, "snapshot slot out of range");
        sim->set_current();
        sim->restore(slot);
    });
}

int sbs_gpu_set_fallback(sbs_gpu* sim, double limit, uint32_t iterations)
{
    return guarded([&] {
        require(sim != nullptr, "sim is null");
        require(limit >= 0.0 && std::isfinite(limit), "limit must be finite and >= 0");
        sim->fb_limit = limit;
        sim->fb_iterations = iterations;
    });
}

int sbs_gpu_fallback_count(sbs_gpu* sim, uint64_t* fallbacks)
{
    return guarded([&] {
        require(sim != nullptr && fallbacks != nullptr, "null argument");
        *fallbacks = sim->fb_count;
    });
}

}  // extern "C"
