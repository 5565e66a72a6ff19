// cf_check.hip -- the reference's two direct one-table algorithms on the device, one workgroup
// per prime, as independent cross-checks of the suffix-table engine (`cuckoo_fail single
// --algo forward|cubic`):
//   forward: Algorithm 1, the O(n^4) recurrence over block boundaries
//            (include/cuckoo/single_table.hpp:44-78)
//   cubic:   the O(n^3) suffix-table build with shared inner sums (single_table.hpp:135-166)
//            followed by the table query nbad(n, M) (single_table.hpp:104-120)
// Both run the reference's loops as written (no convolution, no rescaling), with the exact
// FP64 modular product of cf_core.hpp; the factorial tables are built on the device too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "cf_core.hpp"
#include "cf_engine.hpp"

namespace cf {
namespace {

constexpr int kCheckThreads = 256;

struct CheckShape {
    int algo, d, model, n, M;
    int64_t stride;  // doubles of arena per prime
};

__device__ __forceinline__ double wsum(double x, const Mod& m) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x = reduce(x + __shfl_xor(x, o, 64), m);
    return x;
}

// Sum of x over the workgroup, returned to thread 0.
__device__ double block_sum(double x, double* red, const Mod& m) {
    x = wsum(x, m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kCheckThreads / 64; ++w) s = reduce(s + red[w], m);
    __syncthreads();
    return s;
}

struct Fac {
    const double *f, *fi;
    Mod m;
    // C(a, b), 0 unless 0 <= b <= a (the reference's Factorials::binom, modint.hpp:119-123)
    __device__ __forceinline__ double binom(int a, int b) const {
        if (a < 0 || b < 0 || b > a) return 0.0;
        return mulred(mulred(f[a], fi[b], m), fi[a - b], m);
    }
};

// fact[0..F), ifact[0..F) by one thread (F <= M + 1 < p, so every factorial is invertible)
__device__ void factorials(double* fact, double* ifact, int F, const Mod& m, uint64_t p) {
    if (threadIdx.x == 0) {
        fact[0] = 1.0;
        for (int i = 1; i < F; ++i) fact[i] = canon(mulred(fact[i - 1], (double)i, m), m);
        ifact[F - 1] = canon(powmod(fact[F - 1], p - 2, m), m);
        for (int i = F - 1; i > 0; --i) ifact[i - 1] = canon(mulred(ifact[i], (double)i, m), m);
    }
    __syncthreads();
}

// g(v): C(v, d) for d-subsets, v^d with replacement (single_table.hpp:16-23)
__device__ __forceinline__ double gval(const Fac& fc, int v, int d, int model) {
    if (v < 0) return 0.0;
    return model ? powmod((double)v, (uint64_t)d, fc.m) : fc.binom(v, d);
}

__device__ void cubic(const CheckShape s, double* a, const Mod& m, uint64_t p, double* red, uint64_t* out) {
    const int N = s.n, M = s.M, F = M + 1, W = N + 1;
    double *fact = a, *ifact = a + F, *gp = ifact + F, *V = gp + (int64_t)W * W, *S = V + (int64_t)W * W;
    factorials(fact, ifact, F, m, p);
    const Fac fc{fact, ifact, m};
    const int t = (int)threadIdx.x;
    // gp[v][e] = g(v)^e
    for (int v = t; v <= N; v += kCheckThreads) {
        const double g = gval(fc, v, s.d, s.model);
        double x = 1.0;
        gp[(int64_t)v * W] = 1.0;
        for (int e = 1; e <= N; ++e) gp[(int64_t)v * W + e] = x = canon(mulred(x, g, m), m);
    }
    __syncthreads();
    for (int v = 0; v < N; ++v) {
        // stage 1 with q = v: S[v'][v] for v' >= v from the finished row V[v][.]
        if (v >= 1)
            for (int vp = v + t; vp < N; vp += kCheckThreads) {
                double acc = 0;
                for (int w = 0; w < v; ++w)
                    acc = reduce(acc + mulred(mulred(fc.binom(vp - w - 1, vp - v), fc.binom(vp, w), m),
                                              V[(int64_t)v * W + w], m), m);
                S[(int64_t)vp * N + v] = canon(acc, m);
            }
        __syncthreads();
        // stage 2: column v of V from S[v][1..v]
        for (int pp = v + 1 + t; pp <= N; pp += kCheckThreads) {
            double val = mulred(fc.binom(pp - 1, v), gp[(int64_t)v * W + pp], m);
            for (int q = 1; q <= v; ++q) {
                const double c = mulred(mulred(fc.binom(pp - q - 1, pp - v - 1), fc.binom(pp, q), m),
                                        gp[(int64_t)v * W + pp - q], m);
                val = reduce(val - mulred(c, S[(int64_t)v * N + q], m), m);
            }
            V[(int64_t)pp * W + v] = canon(val, m);
        }
        __syncthreads();
    }
    // nbad(n, M) = sum_q C(n, q) Q^(n-q) sum_{w<q} C(M-w-1, M-q) C(M, w) V[q][w]
    const double Q = gval(fc, M, s.d, s.model);
    double acc = 0;
    for (int q = 1 + t; q <= N; q += kCheckThreads) {
        double inner = 0;
        for (int w = 0; w < q; ++w)
            inner = reduce(inner + mulred(mulred(fc.binom(M - w - 1, M - q), fc.binom(M, w), m), V[(int64_t)q * W + w], m), m);
        acc = reduce(acc + mulred(mulred(fc.binom(N, q), powmod(Q, (uint64_t)(N - q), m), m), inner, m), m);
    }
    const double total = block_sum(acc, red, m);
    if (t == 0) *out = (uint64_t)canon(total, m);
}

// Only the boundaries (0, 0) and (l, b) with z <= l < b <= M carry weight, so the tables are
// kept in shifted coordinates: Fs[l - z][b - z] = F[l][b], and hs[j][u] = h(b)^u for j = b - z
// (j >= 1) and j = 0 for b = 0. Memory O(M + n^2) per prime.
__device__ void forward(const CheckShape s, double* a, const Mod& m, uint64_t p, double* red, uint64_t* out) {
    const int n = s.n, M = s.M, F = M + 1, z = M - n, H = n + 1;
    double *fact = a, *ifact = a + F, *hs = ifact + F, *Fs = hs + (int64_t)H * H;
    factorials(fact, ifact, F, m, p);
    const Fac fc{fact, ifact, m};
    const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6, nw = kCheckThreads / 64;
    // h(b) = g(M - b), eq. (6)
    for (int j = t; j <= n; j += kCheckThreads) {
        const double h = gval(fc, M - (j ? z + j : 0), s.d, s.model);
        double x = 1.0;
        hs[(int64_t)j * H] = 1.0;
        for (int u = 1; u <= n; ++u) hs[(int64_t)j * H + u] = x = canon(mulred(x, h, m), m);
    }
    __syncthreads();
    // F[r][c] = F[0][0] tau(0, 0, r, c) + sum_{z <= l < r, l < b < c} F[l][b] tau(l, b, r, c)
    // (eq. (14)); each wave takes its own c, its lanes split l
    for (int r = z; r < M; ++r) {
        for (int c = r + 1 + wave; c <= M; c += nw) {
            double acc = 0;
            if (lane == 0)  // the start boundary (0, 0): phi = C(c - 1, r), nu = 0, u = r - z
                acc = mulred(mulred(mulred(fc.binom(c - 1, r), fc.binom(M, c), m), fc.binom(n, r - z), m), hs[r - z], m);
            for (int l = z + lane; l < r; l += 64) {
                const int u = r - l, lz = l - z;
                const double cl = fc.binom(n - lz, u);  // C(n - nu(l), u), nu(l) = l - z
                if (cl == 0.0) continue;
                double part = 0;
                for (int b = l + 1; b < c; ++b) {
                    // phi(c - l, b - l, r - l) = -C(c - b - 1, r - b) C(r - l - 1, b - l - 1), eq. (4)
                    const double phi = mulred(fc.binom(c - b - 1, r - b), fc.binom(r - l - 1, b - l - 1), m);
                    const double tau = mulred(mulred(phi, fc.binom(M - b, c - b), m), hs[(int64_t)(b - z) * H + u], m);
                    part = reduce(part - mulred(Fs[(int64_t)lz * H + (b - z)], tau, m), m);
                }
                acc = reduce(acc + mulred(part, cl, m), m);
            }
            acc = wsum(acc, m);
            if (lane == 0) Fs[(int64_t)(r - z) * H + (c - z)] = canon(acc, m);
        }
        __syncthreads();
    }
    // eq. (15): total = sum of F[l][b] C(M-l-1, M-b) h(b)^(n - nu(l)) over the weighted
    // boundaries; at (0, 0) the binomial C(M - 1, M) is 0
    double acc = 0;
    for (int l = z + t; l < M; l += kCheckThreads)
        for (int b = l + 1; b <= M; ++b)
            acc = reduce(acc + mulred(mulred(Fs[(int64_t)(l - z) * H + (b - z)], fc.binom(M - l - 1, M - b), m),
                                      hs[(int64_t)(b - z) * H + n - (l - z)], m), m);
    const double total = block_sum(acc, red, m);
    if (t == 0) *out = (uint64_t)canon(total, m);
}

__global__ void __launch_bounds__(kCheckThreads) k_check(CheckShape s, const uint64_t* __restrict__ primes,
                                                         double* __restrict__ mem, uint64_t* __restrict__ res) {
    __shared__ double red[kCheckThreads / 64];
    const uint64_t p = primes[blockIdx.x];
    Mod m;
    m.p = (double)p;
    m.pinv = 1.0 / m.p;
    double* a = mem + s.stride * blockIdx.x;
    if (s.algo == CF_ALGO_CUBIC) cubic(s, a, m, p, red, res + blockIdx.x);
    else forward(s, a, m, p, red, res + blockIdx.x);
}

int64_t check_doubles(int algo, int n, int M) {
    const int64_t F = M + 1;
    if (algo == CF_ALGO_CUBIC) return 2 * F + 2 * (int64_t)(n + 1) * (n + 1) + (int64_t)n * n;
    return 2 * F + 2 * (int64_t)(n + 1) * (n + 1);
}

}  // namespace
}  // namespace cf

using namespace cf_int;

extern "C" int cf_run_algo(cf_engine* e, int algo, int d, int model, int n, int M, const uint64_t* primes, size_t np,
                           uint64_t* res) {
    if (!e) return CF_E_ARGS;
    e->err.clear();
    e->stats = cf_stats{};
    if ((algo != CF_ALGO_FORWARD && algo != CF_ALGO_CUBIC) || d < 1 || (model != 0 && model != 1) || n < 1 || M <= n ||
        (np && (!primes || !res)))
        return fail(e, CF_E_ARGS, "bad arguments");
    // the O(n^4) / O(n^3) loops run on one workgroup per prime: keep a launch to seconds
    if ((algo == CF_ALGO_FORWARD && n > 256) || (algo == CF_ALGO_CUBIC && n > 1024) || M > (1 << 20))
        return fail(e, CF_E_UNSUPPORTED, "n = " + std::to_string(n) + " is beyond this algorithm's cross-check size");
    for (size_t b = 0; b < np; ++b)
        if (primes[b] < 3 || !(primes[b] & 1) || primes[b] >= (uint64_t(1) << 50) || primes[b] <= (uint64_t)M)
            return fail(e, CF_E_ARGS, "prime " + std::to_string(primes[b]) + " is not an odd modulus in (M, 2^50)");
    if (!np) return CF_OK;
    HIP_TRY(e, hipSetDevice(e->device), "hipSetDevice");
    const cf::CheckShape s{algo, d, model, n, M, (cf::check_doubles(algo, n, M) + 31) & ~(int64_t)31};
    const size_t per = sizeof(double) * (size_t)s.stride;
    const size_t fit = std::min<size_t>(usable_bytes(e) / per, 65535);
    if (!fit) return fail(e, CF_E_NOMEM, "one prime needs " + std::to_string(per) + " bytes of device memory");
    float ms_total = 0;
    for (size_t at = 0; at < np; at += fit) {
        const size_t cnt = std::min(fit, np - at);
        int rc = ensure(e, &e->arena, &e->arena_cap, cnt * per, "the cross-check tables");
        if (rc) return rc;
        rc = ensure(e, &e->small, &e->small_cap, 2 * sizeof(uint64_t) * cnt, "the primes");
        if (rc) return rc;
        uint64_t* dp = (uint64_t*)e->small;
        uint64_t* dr = dp + cnt;
        HIP_TRY(e, hipMemcpyAsync(dp, primes + at, sizeof(uint64_t) * cnt, hipMemcpyHostToDevice, e->stream), "upload");
        HIP_TRY(e, hipEventRecord(e->ev0, e->stream), "hipEventRecord");
        cf::k_check<<<dim3((unsigned)cnt), cf::kCheckThreads, 0, e->stream>>>(s, dp, (double*)e->arena, dr);
        HIP_TRY(e, hipGetLastError(), "kernel launch");
        HIP_TRY(e, hipEventRecord(e->ev1, e->stream), "hipEventRecord");
        HIP_TRY(e, hipStreamSynchronize(e->stream), "kernel execution");
        float ms = 0;
        HIP_TRY(e, hipEventElapsedTime(&ms, e->ev0, e->ev1), "hipEventElapsedTime");
        ms_total += ms;
        HIP_TRY(e, hipMemcpy(res + at, dr, sizeof(uint64_t) * cnt, hipMemcpyDeviceToHost), "download");
        e->stats.launches++;
        e->stats.tasks += (int)cnt;
        e->stats.device_bytes = std::max(e->stats.device_bytes, cnt * per);
    }
    e->stats.batches = e->stats.launches;
    e->stats.kernel_ms = ms_total;
    return CF_OK;
}
