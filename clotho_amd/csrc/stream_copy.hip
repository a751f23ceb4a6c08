// stream_copy.hip — STREAM-style device copy: measures the HBM bandwidth the
// roofline is quoted against on this box, and calibrates rocprofv3's
// FETCH_SIZE / WRITE_SIZE counters against a kernel of known traffic.
#include <hip/hip_runtime.h>

#include "../../include/clotho_amd.h"

namespace {

// one pass: each thread issues four 16-byte loads before its four stores
__global__ void __launch_bounds__(256) k_stream_copy(const uint4* __restrict__ src,
                                                     uint4* __restrict__ dst, size_t n) {
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    uint4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = base + 256 * (size_t)k;
        v[k] = i < n ? src[i] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = base + 256 * (size_t)k;
        if (i < n) dst[i] = v[k];
    }
}

}  // namespace

extern "C" int clotho_stream_copy(size_t bytes, int repeats, double* gbps, int device) {
    if (bytes < 4096 || repeats < 1 || !gbps) return CLOTHO_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return CLOTHO_ERR_HIP;
    const size_t n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CLOTHO_ERR_HIP;
    if (hipMalloc((void**)&a, n * sizeof(uint4)) == hipSuccess &&
        hipMalloc((void**)&b, n * sizeof(uint4)) == hipSuccess &&
        hipMemset(a, 1, n * sizeof(uint4)) == hipSuccess &&
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
        hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        const unsigned blocks = (unsigned)((n + 1023) / 1024);
        hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, s, a, b, n);
        bool ok = hipGetLastError() == hipSuccess && hipEventRecord(e0, s) == hipSuccess;
        for (int r = 0; ok && r < repeats; ++r) {
            hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, s, a, b, n);
            ok = hipGetLastError() == hipSuccess;
        }
        float ms = 0.f;
        ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (ok && ms > 0.f) {
            *gbps = 2.0 * (double)(n * sizeof(uint4)) * repeats / (ms * 1e-3) / 1e9;
            rc = CLOTHO_OK;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

extern "C" int clotho_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!free_bytes || !total_bytes) return CLOTHO_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return CLOTHO_ERR_HIP;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return CLOTHO_ERR_HIP;
    *free_bytes = f;
    *total_bytes = t;
    return CLOTHO_OK;
}
