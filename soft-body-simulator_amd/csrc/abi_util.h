// Error plumbing of the C ABI: C++ exceptions and HIP failures become status codes and a
// thread-local message; nothing propagates across the extern "C" boundary.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>

#include "sbs_gpu.h"

namespace sbs {

struct api_error : std::runtime_error {
    int code;
    api_error(int c, std::string const& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(std::string const& msg);

inline void require(bool ok, char const* msg, int code = SBS_ERR_INVALID_ARGUMENT)
{
    if (!ok) throw api_error(code, msg);
}

inline void hip_check(hipError_t e, char const* what)
{
    if (e != hipSuccess) {
        int const code = (e == hipErrorOutOfMemory) ? SBS_ERR_OUT_OF_MEMORY
                         : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? SBS_ERR_NO_DEVICE
                                                                                 : SBS_ERR_HIP;
        throw api_error(code, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// Owning device array (hipMalloc / hipFree).
template <class T>
struct dev_buf {
    T* p = nullptr;
    std::size_t n = 0;
    dev_buf() = default;
    dev_buf(dev_buf const&) = delete;
    dev_buf& operator=(dev_buf const&) = delete;
    dev_buf(dev_buf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    dev_buf& operator=(dev_buf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~dev_buf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(std::size_t count)
    {
        release();
        if (count == 0) return;
        hip_check(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)), "hipMalloc");
        n = count;
    }
    void upload(T const* h, std::size_t count, hipStream_t s)
    {
        alloc(count);
        if (count) hip_check(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s), "upload");
    }
};

template <class F>
int guarded(F&& f)
{
    try {
        f();
        return SBS_OK;
    } catch (api_error const& e) {
        set_last_error(e.what());
        return e.code;
    } catch (std::bad_alloc const&) {
        set_last_error("host allocation failed");
        return SBS_ERR_OUT_OF_MEMORY;
    } catch (std::exception const& e) {
        set_last_error(e.what());
        return SBS_ERR_INTERNAL;
    } catch (...) {
        set_last_error("unknown error");
        return SBS_ERR_INTERNAL;
    }
}

}  // namespace sbs
