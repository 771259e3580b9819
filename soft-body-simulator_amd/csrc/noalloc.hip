// The radix sort of device-resident pairs without allocation (sbs_util_sort_pairs_noalloc).
//
// Its work buffers are the file-scope device array `scratch` below, sized for
// SBS_SORT_NOALLOC_MAX_PAIRS pairs; every GPU the library runs on holds its own copy of the array
// until the process ends. Host code never passes the array's name to a kernel (in host code that
// name is a stand-in, not a device address): the first call on a device looks up the device address
// of that device's copy with hipGetSymbolAddress, with the device current, and keeps it per device;
// every launch takes that pointer. The entry point calls no hipMalloc, hipFree or allocating routine.
// All calls on one GPU share its array, so calls that overlap on one GPU race; the sort writes
// every word of the work buffers it reads, so nothing in the array needs clearing between calls.
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>

#include "abi_util.h"
#include "sbs_gpu.h"
#include "scan_sort.h"

namespace sbs {
namespace {

constexpr std::size_t kScratchWords = sort_workspace(SBS_SORT_NOALLOC_MAX_PAIRS);

__device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
std::array<uint32_t*, kMaxDevices> g_scratch_ptr{};  // per device: the device address of its `scratch`

// The device address of `scratch` on device `device` (current), looked up once per device.
uint32_t* scratch_of(int device)
{
    require(device >= 0 && device < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        require(p != nullptr, "no device address for the sort's scratch array", SBS_ERR_INTERNAL);
        g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    return g_scratch_ptr[device];
}

}  // namespace
}  // namespace sbs

using namespace sbs;

extern "C" int sbs_util_sort_pairs_noalloc(int device, uint32_t* keys, uint32_t* vals, uint64_t n, int bits,
                                           void* stream)
{
    return guarded([&] {
        require(n == 0 || (keys && vals), "null argument");
        require(bits >= 0 && bits <= 32, "bits must be 0..32");
        require(n <= uint64_t(SBS_SORT_NOALLOC_MAX_PAIRS),
                "too many pairs for the allocation-free sort (at most SBS_SORT_NOALLOC_MAX_PAIRS)");
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        require(device >= 0 && device < count, "no such device", SBS_ERR_NO_DEVICE);
        hip_check(hipSetDevice(device), "hipSetDevice");
        uint32_t* const ws = scratch_of(device);
        if (n <= 1 || bits == 0) return;
        radix_sort_pairs(static_cast<hipStream_t>(stream), keys, vals, std::size_t(n), bits, ws);
        hip_check(hipGetLastError(), "sort launch");
    });
}
