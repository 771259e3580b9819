// Device primitives of the contact hash: exclusive prefix sum and stable LSD radix sort of
// (key, value) pairs (implemented in scan_sort.hip). Launches go to the given stream and use only
// the workspace the caller passes.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace sbs {

// Words a 256-thread block scans or sorts per tile.
constexpr std::size_t kScanTile = 1024;

// Workspace (in uint32 words) of exclusive_scan for n elements: per level the tile sums and their
// scan (tiles + 1 words each), recursively.
constexpr std::size_t scan_workspace(std::size_t n)
{
    std::size_t ws = 0;
    while (n > 1) {
        std::size_t const t = (n + kScanTile - 1) / kScanTile;
        ws += 2 * (t + 1);
        n = t;
    }
    return ws + 2;
}
// out[i] = sum of in[0..i) for i < n, and out[n] = the total (out has n + 1 words). in may equal
// out only when n == 0. ws: scan_workspace(n) words.
void exclusive_scan(hipStream_t s, const uint32_t* in, uint32_t* out, std::size_t n, uint32_t* ws);

// Workspace (in uint32 words) of radix_sort_pairs for n pairs: the other key and value buffers,
// the digit histogram and its scan (256 words per tile, + 1 each), the histogram scan's workspace.
constexpr std::size_t sort_workspace(std::size_t n)
{
    std::size_t const hist = std::size_t(256) * ((n + kScanTile - 1) / kScanTile);
    return 2 * n + 2 * (hist + 1) + scan_workspace(hist);
}
// Stable sort of (keys, vals) by the low `bits` bits of keys (bits <= 32), in place; ws:
// sort_workspace(n) words.
void radix_sort_pairs(hipStream_t s, uint32_t* keys, uint32_t* vals, std::size_t n, int bits, uint32_t* ws);

}  // namespace sbs
