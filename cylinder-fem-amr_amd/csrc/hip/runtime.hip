// Runtime plumbing of the C ABI: device selection, memory, copies, kernel zero-fill, timers.
// Stands in for the CUDA runtime calls wrapped by CUDA_CHECK in the reference
// (src/cfa/common.cuh:33-43, src/cfa/device.cuh:8-19).
#include <cstdio>

#include "cfa_device.hpp"

using namespace cfa_dev;

namespace {
__global__ void k_zero_words(unsigned long long* __restrict__ p, size_t words) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}
__global__ void k_zero_bytes(unsigned char* __restrict__ p, size_t bytes) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < bytes; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}
}  // namespace

extern "C" {

int cfa_init(int device) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    e = hipFree(nullptr);   // create the context now
    if (e != hipSuccess) return (int)e;
    return scratch_setup();   // the no-allocation entry point's workspace address on this device
}

const char* cfa_error_string(int code) {
    switch (code) {
        case CFA_OK: return "success";
        case CFA_E_ARG: return "invalid argument";
        case CFA_E_ROWLEN: return "CSR row too long for the kernel";
        case CFA_E_NOTREADY: return "object not set up";
        case CFA_E_IO: return "file missing or not in the expected layout";
        case CFA_E_SIZE: return "input larger than the fixed workspace of a no-allocation entry point";
        default: return code > 0 ? hipGetErrorString((hipError_t)code) : "unknown error";
    }
}

int cfa_malloc(void** d_ptr, size_t bytes) {
    if (!d_ptr) return CFA_E_ARG;
    *d_ptr = nullptr;
    const hipError_t e = hipMalloc(d_ptr, bytes > 0 ? bytes : 8);
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_free(void* d_ptr) {
    if (!d_ptr) return CFA_OK;
    const hipError_t e = hipFree(d_ptr);
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_host_malloc(void** h_ptr, void** d_alias, size_t bytes) {
    if (!h_ptr) return CFA_E_ARG;
    hipError_t e = hipHostMalloc(h_ptr, bytes > 0 ? bytes : 8, hipHostMallocMapped);
    if (e != hipSuccess) return (int)e;
    if (d_alias) {
        e = hipHostGetDevicePointer(d_alias, *h_ptr, 0);
        if (e != hipSuccess) return (int)e;
    }
    return CFA_OK;
}

int cfa_host_free(void* h_ptr) {
    if (!h_ptr) return CFA_OK;
    const hipError_t e = hipHostFree(h_ptr);
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_memcpy(void* dst, const void* src, size_t bytes, int kind, cfa_stream_t stream) {
    if (bytes == 0) return CFA_OK;
    if (!dst || !src) return CFA_E_ARG;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, k, as_stream(stream));
    if (e == hipSuccess && kind == 1) e = hipStreamSynchronize(as_stream(stream));
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_zero(void* d_ptr, size_t bytes, cfa_stream_t stream) {
    if (bytes == 0) return CFA_OK;
    if (!d_ptr) return CFA_E_ARG;
    if ((reinterpret_cast<size_t>(d_ptr) % 8) == 0 && bytes % 8 == 0) {
        const size_t words = bytes / 8;
        const int blocks = (int)std::min<size_t>((words + BLOCK - 1) / BLOCK, 4096);
        hipLaunchKernelGGL(k_zero_words, dim3(blocks), dim3(BLOCK), 0, as_stream(stream),
                           static_cast<unsigned long long*>(d_ptr), words);
    } else {
        const int blocks = (int)std::min<size_t>((bytes + BLOCK - 1) / BLOCK, 4096);
        hipLaunchKernelGGL(k_zero_bytes, dim3(blocks), dim3(BLOCK), 0, as_stream(stream),
                           static_cast<unsigned char*>(d_ptr), bytes);
    }
    return launch_status();
}

int cfa_stream_sync(cfa_stream_t stream) {
    const hipError_t e = hipStreamSynchronize(as_stream(stream));
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_stream_create(cfa_stream_t* stream) {
    if (!stream) return CFA_E_ARG;
    hipStream_t s = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    *stream = s;
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_stream_destroy(cfa_stream_t stream) {
    if (!stream) return CFA_OK;
    const hipError_t e = hipStreamDestroy(as_stream(stream));
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_device_sync(void) {
    const hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_timer_create(void** timer) {
    if (!timer) return CFA_E_ARG;
    hipEvent_t ev;
    const hipError_t e = hipEventCreate(&ev);
    *timer = ev;
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_timer_destroy(void* timer) {
    if (!timer) return CFA_OK;
    const hipError_t e = hipEventDestroy(static_cast<hipEvent_t>(timer));
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_timer_record(void* timer, cfa_stream_t stream) {
    const hipError_t e = hipEventRecord(static_cast<hipEvent_t>(timer), as_stream(stream));
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_stream_wait(cfa_stream_t stream, void* event) {
    if (!event) return CFA_E_ARG;
    const hipError_t e = hipStreamWaitEvent(as_stream(stream), static_cast<hipEvent_t>(event), 0);
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_timer_elapsed_ms(void* t_begin, void* t_end, float* ms) {
    hipError_t e = hipEventSynchronize(static_cast<hipEvent_t>(t_end));
    if (e != hipSuccess) return (int)e;
    e = hipEventElapsedTime(ms, static_cast<hipEvent_t>(t_begin), static_cast<hipEvent_t>(t_end));
    return e == hipSuccess ? CFA_OK : (int)e;
}

}  // extern "C"
