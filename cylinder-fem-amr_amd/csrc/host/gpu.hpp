// Host-side access to the gfx950 kernel library: error checks with the reference's abort
// behaviour (CUDA_CHECK, src/cfa/common.cuh:33-41), growable device arrays, the solver stream.
// This layer talks to the device only through the C ABI of include/cfa_hip.h.
#pragma once

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../../include/cfa_hip.h"
#include "cfa_common.hpp"

namespace cfa {

// Aborts with the library's message on a non-zero status (reference: CUDA_CHECK -> exit).
inline void gpu_check(int rc, const char* what, const char* file, int line) {
    if (rc != CFA_OK) {
        std::fprintf(stderr, "HIP Error at %s:%d in %s: %s\n", file, line, what, cfa_error_string(rc));
        std::exit(EXIT_FAILURE);
    }
}
#define GPU_CHECK(call) ::cfa::gpu_check((call), #call, __FILE__, __LINE__)

// Stream used by every host object (the default stream unless the caller sets one)
inline cfa_stream_t& solver_stream() {
    static cfa_stream_t s = nullptr;
    return s;
}

// Device array with capacity growth; memory is kept when a smaller size is requested.
template <typename T>
struct DevArray {
    T* ptr = nullptr;
    size_t cap = 0;
    DevArray() = default;
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    DevArray(DevArray&& o) noexcept : ptr(o.ptr), cap(o.cap) { o.ptr = nullptr; o.cap = 0; }
    DevArray& operator=(DevArray&& o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; cap = o.cap; o.ptr = nullptr; o.cap = 0; }
        return *this;
    }
    ~DevArray() { release(); }
    void release() {
        if (ptr) cfa_free(ptr);
        ptr = nullptr;
        cap = 0;
    }
    // at least n elements; new capacity = n + n/2 + pad when growing (src/cfa/device.cuh:71-115)
    void reserve(size_t n, size_t pad = 64) {
        if (n <= cap && ptr) return;
        release();
        const size_t c = n + n / 2 + pad;
        void* p = nullptr;
        GPU_CHECK(cfa_malloc(&p, c * sizeof(T)));
        ptr = static_cast<T*>(p);
        cap = c;
    }
    void upload(const T* h, size_t n) {
        reserve(n);
        if (n) GPU_CHECK(cfa_memcpy(ptr, h, n * sizeof(T), 0, solver_stream()));
    }
    void upload(const std::vector<T>& h) { upload(h.data(), h.size()); }
    void download(T* h, size_t n) const {
        if (n) GPU_CHECK(cfa_memcpy(h, ptr, n * sizeof(T), 1, solver_stream()));
    }
    std::vector<T> to_host(size_t n) const {
        std::vector<T> h(n);
        download(h.data(), n);
        return h;
    }
    operator T*() const { return ptr; }
};

inline void dev_copy(double* dst, const double* src, size_t n) {
    if (n) GPU_CHECK(cfa_memcpy(dst, src, n * sizeof(double), 2, solver_stream()));
}
inline void dev_zero(void* p, size_t bytes) {
    if (bytes) GPU_CHECK(cfa_zero(p, bytes, solver_stream()));
}
inline void dev_sync() { GPU_CHECK(cfa_stream_sync(solver_stream())); }

}  // namespace cfa
