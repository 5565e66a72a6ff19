// cf_modmul_rate -- peak rate of the engine's two exact modular products on this GPU, for the
// roofline of the suffix-table kernels:
//   * the FP64 product of cf_core.hpp (residues below 2^50: split by FMA, quotient by rint,
//     6 FP64 vector instructions), and
//   * the 64-bit Montgomery product of cf_p62.hip (primes up to 2^63, R = 2^64).
// Every thread runs 8 independent product chains (x <- x y mod p) for `iters` rounds, on a grid
// of 2048 workgroups of 256 threads; after a warm-up launch, the best of five timed launches (each
// about 12 ms at the default 16384 rounds: long enough for the clock to settle, where a 3 ms run
// reads 20% low) gives the products per second, printed as one JSON line.
//
//   cf_modmul_rate [iters]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "cf_core.hpp"

namespace {

constexpr int kChains = 8, kBlocks = 2048, kThreads = 256;

__global__ void k_fp64(double p, double pinv, int iters, double* out) {
    const cf::Mod m{p, pinv};
    double x[kChains], y[kChains];
    for (int i = 0; i < kChains; ++i) {
        x[i] = (double)(threadIdx.x + 17 * i + 3);
        y[i] = (double)(blockIdx.x * 977 + 31 * i + 5);
    }
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int i = 0; i < kChains; ++i) x[i] = cf::mulred(x[i], y[i], m);
    double s = 0;
    for (int i = 0; i < kChains; ++i) s += x[i];
    out[blockIdx.x * kThreads + threadIdx.x] = s;
}

__device__ __forceinline__ uint64_t mont(uint64_t a, uint64_t b, uint64_t p, uint64_t pinv) {
    const uint64_t lo = a * b, hi = __umul64hi(a, b);
    const uint64_t q = lo * pinv, h2 = __umul64hi(q, p);
    return hi >= h2 ? hi - h2 : hi - h2 + p;
}

__global__ void k_mont64(uint64_t p, uint64_t pinv, int iters, uint64_t* out) {
    uint64_t x[kChains], y[kChains];
    for (int i = 0; i < kChains; ++i) {
        x[i] = (uint64_t)(threadIdx.x + 17 * i + 3);
        y[i] = (uint64_t)(blockIdx.x * 977 + 31 * i + 5);
    }
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int i = 0; i < kChains; ++i) x[i] = mont(x[i], y[i], p, pinv);
    uint64_t s = 0;
    for (int i = 0; i < kChains; ++i) s ^= x[i];
    out[blockIdx.x * kThreads + threadIdx.x] = s;
}

#define CHECK(x)                                                                       \
    do {                                                                               \
        hipError_t st_ = (x);                                                          \
        if (st_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(st_));              \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

template <class F>
double rate(F launch, int iters) {
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    launch();  // warm-up
    CHECK(hipDeviceSynchronize());
    float best = 0;
    for (int rep = 0; rep < 5; ++rep) {
        CHECK(hipEventRecord(a, 0));
        launch();
        CHECK(hipEventRecord(b, 0));
        CHECK(hipEventSynchronize(b));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, a, b));
        if (rep == 0 || ms < best) best = ms;
    }
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
    return (double)kBlocks * kThreads * kChains * iters / (best * 1e-3);
}

}  // namespace

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 16384;
    const uint64_t p50 = 1125899865948161ull;  // prime 0 of the 50-bit list
    const uint64_t p62 = 4611686018405367809ull;  // prime 0 of the 62-bit list
    uint64_t pinv = p62;
    for (int i = 0; i < 6; ++i) pinv *= 2 - p62 * pinv;
    void* out = nullptr;
    CHECK(hipMalloc(&out, sizeof(double) * kBlocks * kThreads));
    const double f = rate([&] { k_fp64<<<kBlocks, kThreads>>>((double)p50, 1.0 / (double)p50, iters, (double*)out); }, iters);
    const double g = rate([&] { k_mont64<<<kBlocks, kThreads>>>(p62, pinv, iters, (uint64_t*)out); }, iters);
    CHECK(hipFree(out));
    std::printf("{\"fp64_exact_modmul_per_s\": %.4e, \"fp64_fma_equivalent_tflops\": %.3f, "
                "\"mont64_modmul_per_s\": %.4e, \"chains_per_thread\": %d, \"threads\": %d, \"iters\": %d}\n",
                f, f * 12 / 1e12, g, kChains, kBlocks * kThreads, iters);
    return 0;
}
