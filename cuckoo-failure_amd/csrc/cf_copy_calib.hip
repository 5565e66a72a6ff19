// cf_copy_calib -- a plain device-to-device copy of a known size, for calibrating the HBM traffic
// counters (FETCH_SIZE / WRITE_SIZE) that profile the engine's kernels: the counter values of
// this kernel divided by the bytes it must move give the scale factor of each counter.
//
//   cf_copy_calib [MiB]      (default 1024 MiB; prints the bytes read and written per launch)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const double4* __restrict__ src, double4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? (size_t)atol(argv[1]) : 1024;
    const size_t bytes = mib << 20, n = bytes / sizeof(double4);
    double4 *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess) {
        fprintf(stderr, "allocation failed\n");
        return 1;
    }
    if (hipMemset(src, 1, bytes) != hipSuccess) return 1;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1;
    for (int rep = 0; rep < 2; ++rep) {
        if (hipEventRecord(a) != hipSuccess) return 1;
        k_copy<<<2048, 256>>>(src, dst, n);
        if (hipEventRecord(b) != hipSuccess || hipEventSynchronize(b) != hipSuccess) return 1;
        float ms = 0;
        if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 1;
        printf("copy %zu bytes read, %zu bytes written: %.3f ms, %.1f GB/s\n", bytes, bytes, ms, 2.0 * bytes / ms / 1e6);
    }
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
