// Spatial hash of the SDF obstacles and the hashed contact detection (gfx950).
//
// Replaces the reference's broad phase -- a KD-tree of bounding spheres over the surface points,
// traversed against each SDF's englobing volume (bvh_model.cpp:30-100) -- for scenes with many
// obstacles. Cells are cubes of edge h; cell (i, j, k) = floor(p / h) per axis in FP64, and its slot
// in a table of 2^b entries is the spatial hash of Teschner et al. 2003, "Optimized Spatial
// Hashing for Collision Detection of Deformable Objects": (i*73856093 ^ j*19349663 ^ k*83492791)
// mod 2^b, on 32-bit two's-complement integers.
//
// Build (once per obstacle set): every hashed obstacle counts the cells its volume overlaps, an
// exclusive prefix sum places its (slot, obstacle) pairs, a stable radix sort groups them by slot
// (keeping obstacle order inside a slot), a second prefix sum drops repeated pairs (two cells of one
// obstacle in one slot), and each slot's range [start, end) is marked. Detection: one thread per
// surface vertex merges its slot's obstacle list with the list of unhashed (unbounded) obstacles in
// increasing obstacle index and applies the exact contact test of the brute-force path, so the
// contacts, and their order, are the brute force's.
#include "contact_hash.h"
#include "scan_sort.h"
#include "sdf.h"

namespace sbs {

namespace {

unsigned grid_for(std::size_t n, int block) { return unsigned((n + std::size_t(block) - 1) / std::size_t(block)); }

__global__ void k_obstacle_count(const sdf_dev* __restrict__ sdfs, const uint32_t* __restrict__ ids, uint32_t n,
                                 double inv_h, uint32_t* __restrict__ count, int4* __restrict__ lo,
                                 int4* __restrict__ ext)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sdf_dev const f = sdfs[ids[i]];
    int a[3], e[3];
    for (int d = 0; d < 3; ++d) {
        a[d] = hash_cell(f.vmin[d], inv_h);
        e[d] = hash_cell(f.vmax[d], inv_h) - a[d] + 1;
    }
    lo[i] = make_int4(a[0], a[1], a[2], 0);
    ext[i] = make_int4(e[0], e[1], e[2], 0);
    count[i] = uint32_t(e[0]) * uint32_t(e[1]) * uint32_t(e[2]);
}

__global__ void k_obstacle_emit(const uint32_t* __restrict__ ids, uint32_t n, const uint32_t* __restrict__ offs,
                                const int4* __restrict__ lo, const int4* __restrict__ ext, uint32_t mask,
                                uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int4 const a = lo[i], e = ext[i];
    uint32_t o = offs[i];
    for (int z = 0; z < e.z; ++z)
        for (int y = 0; y < e.y; ++y)
            for (int x = 0; x < e.x; ++x) {
                keys[o] = hash_slot(a.x + x, a.y + y, a.z + z, mask);
                vals[o] = ids[i];
                ++o;
            }
}

__global__ void k_first_flags(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t n,
                              uint32_t* __restrict__ flag)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || keys[i] != keys[i - 1] || vals[i] != vals[i - 1]) ? 1u : 0u;
}

__global__ void k_compact(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                          const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, uint32_t n,
                          uint32_t* __restrict__ okeys, uint32_t* __restrict__ ovals)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    okeys[pos[i]] = keys[i];
    ovals[pos[i]] = vals[i];
}

__global__ void k_fill(uint32_t* __restrict__ p, uint32_t v, uint32_t n)
{
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_slot_ranges(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ count, uint32_t* start,
                              uint32_t* end)
{
    uint32_t const n = *count;
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t const k = keys[i];
    if (i == 0 || keys[i - 1] != k) start[k] = i;
    if (i + 1 == n || keys[i + 1] != k) end[k] = i + 1;
}

// One thread per surface vertex (see the file comment).
__global__ void k_detect_hashed(contact_dev c, const pos4* __restrict__ x, const sdf_dev* __restrict__ sdfs,
                                contact_hash_dev h)
{
    uint32_t const s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c.S) return;
    pos4 const p = x[c.surf[s]];
    uint32_t const slot = hash_slot(hash_cell(p.x, h.inv_h), hash_cell(p.y, h.inv_h), hash_cell(p.z, h.inv_h), h.mask);
    uint32_t a = h.start[slot], ae = h.end[slot], b = 0;
    uint32_t cnt = 0;
    while (a < ae || b < h.nglobal) {
        uint32_t const ka = a < ae ? h.entries[a] : 0xffffffffu;
        uint32_t const kb = b < h.nglobal ? h.global[b] : 0xffffffffu;
        uint32_t const k = ka < kb ? ka : kb;
        if (ka == k) ++a;
        if (kb == k) ++b;
        double q[3], n[3];
        if (!sdf_contact(sdfs[k], p, q, n)) continue;
        if (cnt == c.slots) {
            atomicAdd(c.overflow, 1u);
            continue;
        }
        store_contact(c, s, cnt++, k, q, n);
    }
    c.count[s] = cnt;
}

}  // namespace

namespace {

// Word offsets of the workspace areas (int4 areas 4-word aligned).
struct hash_ws {
    std::size_t count, offs, lo, ext, keys, vals, flag, pos, ckeys, tmp, total;
    hash_ws(uint32_t nh, std::size_t n)
    {
        count = 0;
        offs = count + nh;
        lo = (offs + nh + 1 + 3) / 4 * 4;
        ext = lo + 4 * std::size_t(nh);
        keys = ext + 4 * std::size_t(nh);
        vals = keys + n;
        flag = vals + n;
        pos = flag + n;
        ckeys = pos + n + 1;
        tmp = ckeys + n;
        std::size_t const scan_ws = scan_workspace(n > nh ? n : nh);
        std::size_t const sort_ws = sort_workspace(n);
        total = tmp + (scan_ws > sort_ws ? scan_ws : sort_ws);
    }
};

__global__ void k_copy_word(const uint32_t* __restrict__ a, uint32_t* __restrict__ b) { *b = *a; }

}  // namespace

std::size_t contact_hash_workspace(uint32_t nhashed, uint64_t pairs) { return hash_ws(nhashed, pairs).total; }

void build_contact_hash(hipStream_t s, const sdf_dev* sdfs, const uint32_t* hashed_ids, uint32_t nhashed,
                        uint64_t pairs, double inv_h, int bits, uint32_t* start, uint32_t* end, uint32_t* entries,
                        uint32_t* num_entries, uint32_t* ws)
{
    uint32_t const T = 1u << bits, mask = T - 1;
    std::size_t const n = pairs;
    hash_ws const w(nhashed, n);
    uint32_t* count = ws + w.count;
    uint32_t* offs = ws + w.offs;
    int4* lo = reinterpret_cast<int4*>(ws + w.lo);
    int4* ext = reinterpret_cast<int4*>(ws + w.ext);
    uint32_t* keys = ws + w.keys;
    uint32_t* vals = ws + w.vals;
    uint32_t* flag = ws + w.flag;
    uint32_t* pos = ws + w.pos;
    uint32_t* ckeys = ws + w.ckeys;
    uint32_t* tmp = ws + w.tmp;
    k_fill<<<grid_for(T, 256), 256, 0, s>>>(start, 0u, T);
    k_fill<<<grid_for(T, 256), 256, 0, s>>>(end, 0u, T);
    if (nhashed == 0 || n == 0) {
        k_fill<<<1, 1, 0, s>>>(num_entries, 0u, 1);
        return;
    }
    k_obstacle_count<<<grid_for(nhashed, 256), 256, 0, s>>>(sdfs, hashed_ids, nhashed, inv_h, count, lo, ext);
    exclusive_scan(s, count, offs, nhashed, tmp);
    k_obstacle_emit<<<grid_for(nhashed, 256), 256, 0, s>>>(hashed_ids, nhashed, offs, lo, ext, mask, keys, vals);
    radix_sort_pairs(s, keys, vals, n, bits, tmp);
    k_first_flags<<<grid_for(n, 256), 256, 0, s>>>(keys, vals, uint32_t(n), flag);
    exclusive_scan(s, flag, pos, n, tmp);
    k_compact<<<grid_for(n, 256), 256, 0, s>>>(keys, vals, flag, pos, uint32_t(n), ckeys, entries);
    k_copy_word<<<1, 1, 0, s>>>(pos + n, num_entries);
    k_slot_ranges<<<grid_for(n, 256), 256, 0, s>>>(ckeys, num_entries, start, end);
}

void launch_detect_hashed(hipStream_t s, contact_dev c, const pos4* x, const sdf_dev* sdfs, contact_hash_dev h,
                          int block)
{
    if (c.S) k_detect_hashed<<<grid_for(c.S, block), block, 0, s>>>(c, x, sdfs, h);
}

}  // namespace sbs
