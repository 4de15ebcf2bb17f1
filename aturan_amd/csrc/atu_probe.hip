// A device-to-device copy of a known number of bytes: the benchmark's measure of the HBM bandwidth
// this box reaches with a plain streaming kernel, and the calibration of the byte counters
// (FETCH_SIZE, WRITE_SIZE) the roofline's traffic figure is read from. No counterpart in the
// reference.

#include "atu_kernels.h"

#include <cstdio>

namespace {

typedef double dvec2 __attribute__((ext_vector_type(2)));

// four 16-byte values per thread, each wave's loads and stores covering whole 128-byte lines
__global__ void __launch_bounds__(256) k_copy_probe(const dvec2* __restrict__ a, dvec2* __restrict__ b,
                                                    size_t n) {
    const size_t base = (size_t)blockIdx.x * (4 * 256) + threadIdx.x;
    dvec2 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const size_t q = base + (size_t)u * 256;
        if (q < n) v[u] = __builtin_nontemporal_load(a + q);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const size_t q = base + (size_t)u * 256;
        if (q < n) __builtin_nontemporal_store(v[u], b + q);
    }
}

}  // namespace

extern "C" int atu_copy_probe(size_t bytes, int launches, double* ms_per_launch) {
    if (bytes < 16 || bytes % 16 != 0 || launches < 1 || !ms_per_launch) return ATU_EINVAL;
    const size_t n = bytes / 16;
    dvec2 *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ATU_OK;
    if (hipMalloc((void**)&a, bytes) != hipSuccess || hipMalloc((void**)&b, bytes) != hipSuccess) rc = ATU_ENOMEM;
    if (!rc && (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 0, bytes, s) != hipSuccess))
        rc = ATU_EHIP;
    const dim3 grid((unsigned)((n + 4 * 256 - 1) / (4 * 256)));
    if (!rc) {
        hipLaunchKernelGGL(k_copy_probe, grid, dim3(256), 0, s, a, b, n);   // warm-up
        if (hipEventRecord(e0, s) != hipSuccess) rc = ATU_EHIP;
        for (int q = 0; q < launches && !rc; q++) hipLaunchKernelGGL(k_copy_probe, grid, dim3(256), 0, s, a, b, n);
        float ms = 0.f;
        if (!rc && (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
                    hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = ATU_EHIP;
        if (!rc) *ms_per_launch = ms / launches;
    }
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
