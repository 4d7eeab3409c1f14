// device.hpp -- HIP runtime helpers shared by the device-side translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "core.hpp"

namespace wx {

#define WX_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            ::wx::fail(WX_ERR_DEVICE, std::string(#call " failed: ") + hipGetErrorString(e_));         \
    } while (0)

// Owning device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 16;
        WX_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
    template <typename T> void upload(const T *src, size_t count, hipStream_t s = nullptr) {
        alloc(count * sizeof(T));
        if (count) WX_HIP(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
        if (count) WX_HIP(hipStreamSynchronize(s));
    }
};

inline uint32_t pow2_shift(uint32_t k) {
    uint32_t s = 0;
    while ((1u << s) < k) ++s;
    return s;
}

} // namespace wx
