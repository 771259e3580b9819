// Device primitives of the contact hash: exclusive prefix sum and stable LSD radix sort of
// (key, value) pairs (implemented in scan_sort.hip). Launches go to the given stream and use only
// the workspace the caller passes.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace sbs {

// Workspace (in uint32 words) of exclusive_scan for n elements.
std::size_t scan_workspace(std::size_t n);
// out[i] = sum of in[0..i) for i < n, and out[n] = the total (out has n + 1 words). in may equal
// out only when n == 0. ws: scan_workspace(n) words.
void exclusive_scan(hipStream_t s, const uint32_t* in, uint32_t* out, std::size_t n, uint32_t* ws);

// Workspace (in uint32 words) of radix_sort_pairs for n pairs.
std::size_t sort_workspace(std::size_t n);
// Stable sort of (keys, vals) by the low `bits` bits of keys (bits <= 32), in place; ws:
// sort_workspace(n) words.
void radix_sort_pairs(hipStream_t s, uint32_t* keys, uint32_t* vals, std::size_t n, int bits, uint32_t* ws);

}  // namespace sbs
