// Host helpers around the HIP runtime shared by the engines.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "internal.hpp"

namespace hgqp {

// Throws Error with an HGQP_* code for a failed HIP call.
void hip_check(hipError_t e, const char* what);
#define HIPCHECK(call) ::hgqp::hip_check((call), #call)

// Checks that a HIP device is present and is gfx950.
void require_device();

// One hipMalloc per handle; sub-buffers are 256-byte aligned slices.
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0;
    void reserve(size_t bytes) {
        size = bytes;
        HIPCHECK(hipMalloc(reinterpret_cast<void**>(&base), bytes));
    }
    template <class T>
    T* take(size_t count) {
        size_t off = (used + 255) & ~size_t(255);
        size_t bytes = count * sizeof(T);
        if (off + bytes > size) throw Error(HGQP_DEVICE_ERROR, "internal: arena overflow");
        used = off + bytes;
        return reinterpret_cast<T*>(base + off);
    }
    static size_t span(size_t count, size_t elem) { return ((count * elem + 255) & ~size_t(255)) + 256; }
    ~Arena() {
        if (base) (void)hipFree(base);
    }
};

// Host-to-device copy of caller memory through two pinned staging buffers
// (process-wide, 2 x 64 MiB, allocated on first use and held until exit):
// host threads fill one buffer while the DMA engine drains the other. Small
// copies go straight through hipMemcpyAsync. Returns once `src` has been read;
// the copy completes in stream order on `stream`.
void upload(void* dst, const void* src, size_t bytes, hipStream_t stream);

// A device allocation owned by a scope (temporaries of handle creation).
template <class T>
struct DeviceTemp {
    T* p = nullptr;
    DeviceTemp() = default;
    explicit DeviceTemp(size_t count) { alloc(count); }
    void alloc(size_t count) {
        if (count) HIPCHECK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    }
    ~DeviceTemp() {
        if (p) (void)hipFree(p);
    }
    DeviceTemp(const DeviceTemp&) = delete;
    DeviceTemp& operator=(const DeviceTemp&) = delete;
};

}  // namespace hgqp
