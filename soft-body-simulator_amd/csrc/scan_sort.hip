// Exclusive prefix sum and stable LSD radix sort on gfx950, built from wave64 prefix sums
// (__shfl_up over the 64 lanes of a wave) and ballots.
//
// Scan: each 256-thread block scans a tile of 1024 words (4 per thread) with a wave prefix sum of
// the thread totals and an LDS pass over the 4 wave totals; the block totals are scanned
// recursively and added back.
//
// Radix sort: 8 bits per pass. A pass counts digits per 1024-word tile (LDS histogram), scans the
// digit-major histogram (offset of digit d in tile b = all smaller digits + digit d in earlier
// tiles), and scatters each tile in input order: within a wave, the lanes holding the same digit
// are found with 8 ballots (one per digit bit), a lane's rank is the number of such lanes below it,
// and the per-wave digit counts are combined across the tile's waves and chunks in order, so the
// sort is stable.
#include "scan_sort.h"

namespace sbs {

namespace {

constexpr int kBlock = 256, kItems = 4, kTile = kBlock * kItems;
static_assert(kTile == kScanTile, "tile size of the workspace formulas");

unsigned blocks_of(std::size_t n) { return unsigned((n + kTile - 1) / kTile); }

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive(uint32_t v)
{
    int const lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t const u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// Exclusive prefix of v over the block (256 threads); *total gets the block sum.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t wsum[kBlock / 64];
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t const inc = wave_inclusive(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t base = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wave) base += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return base + inc - v;
}

__global__ void __launch_bounds__(kBlock) k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                       std::size_t n, uint32_t* __restrict__ tile_sums)
{
    std::size_t const base = std::size_t(blockIdx.x) * kTile + std::size_t(threadIdx.x) * kItems;
    uint32_t v[kItems];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        v[i] = base + i < n ? in[base + i] : 0u;
        sum += v[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive(sum, &total);
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBlock) k_add_tiles(uint32_t* __restrict__ out, std::size_t n,
                                                      const uint32_t* __restrict__ tile_offsets)
{
    std::size_t const base = std::size_t(blockIdx.x) * kTile + std::size_t(threadIdx.x) * kItems;
    uint32_t const add = tile_offsets[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kItems; ++i)
        if (base + i < n) out[base + i] += add;
}

// out[n] = tile_offsets[tiles]: the offset of the (virtual) tile after the last one, or with
// tiles == 0 the single tile's sum.
__global__ void k_set_total(uint32_t* out, std::size_t n, const uint32_t* tile_offsets, unsigned tiles)
{
    out[n] = tile_offsets[tiles];
}

__global__ void k_zero(uint32_t* out) { out[0] = 0; }

__global__ void __launch_bounds__(kBlock) k_copy2(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                                  uint32_t* __restrict__ oa, uint32_t* __restrict__ ob, std::size_t n)
{
    std::size_t const base = std::size_t(blockIdx.x) * kTile + threadIdx.x;
#pragma unroll
    for (int c = 0; c < kItems; ++c) {
        std::size_t const i = base + std::size_t(c) * kBlock;
        if (i < n) {
            oa[i] = a[i];
            ob[i] = b[i];
        }
    }
}

// Digit histogram of one tile, digit-major: hist[d * tiles + tile].
__global__ void __launch_bounds__(kBlock) k_digit_hist(const uint32_t* __restrict__ keys, std::size_t n, int shift,
                                                       uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    std::size_t const base = std::size_t(blockIdx.x) * kTile;
#pragma unroll
    for (int c = 0; c < kItems; ++c) {
        std::size_t const i = base + std::size_t(c) * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[std::size_t(threadIdx.x) * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kBlock) k_digit_scatter(const uint32_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ vals, std::size_t n, int shift,
                                                          const uint32_t* __restrict__ offsets,
                                                          uint32_t* __restrict__ okeys, uint32_t* __restrict__ ovals)
{
    __shared__ uint32_t run[256];           // next output position of each digit for this tile
    __shared__ uint32_t wcount[kBlock / 64][256];
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    run[threadIdx.x] = offsets[std::size_t(threadIdx.x) * gridDim.x + blockIdx.x];
    uint64_t const below = (uint64_t(1) << lane) - 1;
    std::size_t const base = std::size_t(blockIdx.x) * kTile;
    for (int c = 0; c < kItems; ++c) {
        for (int w = 0; w < kBlock / 64; ++w) wcount[w][threadIdx.x] = 0;
        __syncthreads();
        std::size_t const i = base + std::size_t(c) * kBlock + threadIdx.x;
        bool const valid = i < n;
        uint32_t const key = valid ? keys[i] : 0u;
        uint32_t const d = (key >> shift) & 0xffu;
        // lanes of this wave holding a valid item with the same digit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t const ones = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? ones : ~ones;
        }
        uint32_t const rank = uint32_t(__popcll(peers & below));
        if (valid && rank == 0) wcount[wave][d] = uint32_t(__popcll(peers));
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcount[w][d];
            okeys[pos] = key;
            ovals[pos] = vals[i];
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kBlock / 64; ++w) add += wcount[w][threadIdx.x];
        run[threadIdx.x] += add;
        __syncthreads();
    }
}

}  // namespace

void exclusive_scan(hipStream_t s, const uint32_t* in, uint32_t* out, std::size_t n, uint32_t* ws)
{
    if (n == 0) {
        k_zero<<<1, 1, 0, s>>>(out);
        return;
    }
    unsigned const tiles = blocks_of(n);
    uint32_t* sums = ws;
    uint32_t* offs = ws + tiles + 1;
    k_scan_tiles<<<tiles, kBlock, 0, s>>>(in, out, n, sums);
    if (tiles == 1) {  // one tile: no offsets to add, the total is its sum
        k_set_total<<<1, 1, 0, s>>>(out, n, sums, 0);
        return;
    }
    exclusive_scan(s, sums, offs, tiles, offs + tiles + 1);
    k_add_tiles<<<tiles, kBlock, 0, s>>>(out, n, offs);
    k_set_total<<<1, 1, 0, s>>>(out, n, offs, tiles);
}

void radix_sort_pairs(hipStream_t s, uint32_t* keys, uint32_t* vals, std::size_t n, int bits, uint32_t* ws)
{
    if (n <= 1 || bits <= 0) return;
    unsigned const tiles = blocks_of(n);
    std::size_t const hn = std::size_t(256) * tiles;
    uint32_t* tk = ws;
    uint32_t* tv = ws + n;
    uint32_t* hist = ws + 2 * n;
    uint32_t* offs = hist + hn + 1;
    uint32_t* sws = offs + hn + 1;
    uint32_t *ik = keys, *iv = vals, *ok = tk, *ov = tv;
    for (int shift = 0; shift < bits; shift += 8) {
        k_digit_hist<<<tiles, kBlock, 0, s>>>(ik, n, shift, hist);
        exclusive_scan(s, hist, offs, hn, sws);
        k_digit_scatter<<<tiles, kBlock, 0, s>>>(ik, iv, n, shift, offs, ok, ov);
        uint32_t* t;
        t = ik, ik = ok, ok = t;
        t = iv, iv = ov, ov = t;
    }
    if (ik != keys) {
        k_copy2<<<blocks_of(n), kBlock, 0, s>>>(ik, iv, keys, vals, n);
    }
}

}  // namespace sbs
