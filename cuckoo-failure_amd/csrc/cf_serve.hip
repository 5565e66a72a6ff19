// cf_serve.hip -- queries on resident suffix tables: N_bad(n_j, M_j) mod p for a batch of
// queries and every prime of a table set, reading each table row once for the whole batch.
//
// For one prime, with the stored rows a_v[w] = V[v][w] / (w! v! (v-1-w)!) (the stage-1 inputs of
// the build), the query is (fast/cuckoo_fast.cu:306-323 with the Horner finish of :877-890)
//     N_bad(n, M) = M! n! sum_{v = v0+1}^{n} H(v) s_v(M),   s_v(M) = sum_{w<v} a_v[w] / (M - w),
//     H(v) = Q^(n-v) / ((n-v)! (M-v)!),                     Q = g(M).
// The batch is cut into passes of kServeJ queries (sorted by n, largest first); for each pass
// the rows 1 .. max n are split into ranges of about equal area, one per wave. A wave walks its
// range from the top row down, keeping H_j(v) of query j in lane j (H(v-1) = H(v) Q /
// ((n-v+1)(M-v+1)), so no inverse of Q is needed). A row is read in chunks of 512 entries (8 per
// lane), the next chunk loaded while the current one is used for every query of the pass; a
// chunk that the row fills only partly runs a variant with 1, 2 or 4 entries per lane;
// 1/(M - w) comes from the prime's inverse table, copied into LDS by the workgroup. Each wave
// writes one partial sum per query; the finish kernel adds the partials and multiplies by M! n!.
#include <hip/hip_runtime.h>

#include <mutex>

#include <type_traits>

#include "cf_core.hpp"
#include "cf_serve.hpp"

namespace cf {

namespace {

__device__ __forceinline__ double wave_sum(double x, const Mod& m) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x = reduce(x + __shfl_xor(x, o, 64), m);
    return x;
}

constexpr int kWaves = kServeThreads / 64;
constexpr int kChunk = 8;  // row entries per lane per chunk (512 per wave)
constexpr int kPad = 64 * kChunk;
static_assert(kPad == kServeLdsPad, "the host's LDS bound assumes this padding");

__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const uint64_t u = __double_as_longlong(x);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane), hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Workgroup -> (prime b, query pass z, range group y). Workgroups are dealt to the XCDs round
// robin; the passes and range groups of one prime get numbers 8 apart, so they run on one XCD
// at about the same time and share its L2 for the prime's rows.
__device__ __forceinline__ void serve_block(const ServeArgs& a, int& b, int& z, int& y) {
    const int L = (int)blockIdx.x, xcd = L & 7, s = L >> 3, per = a.passes * a.groups;
    b = (s / per) * 8 + xcd;
    z = (s % per) / a.groups;
    y = (s % per) % a.groups;
}

__global__ void __launch_bounds__(kServeThreads, 2)
    k_serve_rows(ServeArgs a) {
    extern __shared__ double inv_lds[];
    int b, z, y;
    serve_block(a, b, z, y);
    if (b >= a.np) return;  // padding of the last group of 8 primes
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int range = y * kWaves + wave;  // row range of this wave within the pass
    const int j0 = z * kServeJ;           // first query of this pass
    const int jn = min(kServeJ, a.J - j0);
    const Mod m{a.primes[b].p, a.primes[b].pinv};
    const double* keep = a.store + a.stride * b;
    const double* fact = keep;
    const double* ifact = keep + a.F1;
    const double* inv = keep + 2 * a.F1;
    const ServeQuery* q = a.q + j0;

    // any query of this pass for this prime? (queries are sorted by n, largest first)
    int live = 0;
    for (int j = 0; j < jn; ++j) live |= b < q[j].need;
    if (!live) return;  // uniform over the workgroup
    // 1/i at inv_lds[kPad + i]; kPad zeros in front, so 1/(M - w) may be read without a select
    // for the masked entries w >= v of a chunk (M - w > -kPad there)
    double* invp = inv_lds + kPad;
    for (int i = threadIdx.x; i < a.F1 + kPad; i += kServeThreads) inv_lds[i] = i < kPad ? 0.0 : inv[i - kPad];
    __syncthreads();

    const int* bd = a.bounds + z * (a.ranges + 1);
    const int lo = bd[range], hi = bd[range + 1];  // rows lo+1 .. hi
    // lane j < jn keeps query j: n, M, whether it runs on this prime, and H_j at the current row
    int qn = 0, qM = 0;
    double H = 0.0, Q = 0.0;
    bool qlive = false;
    if (lane < jn) {
        const ServeQuery qq = q[lane];
        qn = qq.n;
        qM = qq.M;
        qlive = b < qq.need;
        if (a.model) Q = powmod((double)qM, (uint64_t)a.dd, m);
        else if (qM >= a.dd) Q = reduce(mulred(mulred(fact[qM], ifact[a.dd], m), ifact[qM - a.dd], m), m);
        const int top = min(hi, qn);  // the first row this query uses in the range
        if (top > lo && top > a.v0)
            H = reduce(mulred(mulred(powmod(Q, (uint64_t)(qn - top), m), ifact[qn - top], m), ifact[qM - top], m), m);
    }
    double acc[kServeJ];
#pragma unroll
    for (int j = 0; j < kServeJ; ++j) acc[j] = 0.0;

    const int vmin = max(lo + 1, a.v0 + 1);
    double x[kChunk];
    auto load = [&](int v, int w0) {
        const double* row = keep + keep_row(a.F1, v);
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
            const int w = w0 + lane + 64 * k;
            x[k] = row[w < v ? w : 0];
            if (w >= v) x[k] = 0.0;
        }
    };
    if (hi >= vmin) load(hi, 0);
    for (int v = hi; v >= vmin; --v) {
        // the queries that use row v, one bit each (uniform)
        const uint64_t on = __ballot(lane < jn && qlive && v <= qn);
        for (int w0 = 0; w0 < v; w0 += 64 * kChunk) {
            double cur[kChunk];
#pragma unroll
            for (int k = 0; k < kChunk; ++k) cur[k] = x[k];
            // prefetch the next chunk of this row, or the first of the next row
            if (w0 + 64 * kChunk < v) load(v, w0 + 64 * kChunk);
            else if (v - 1 >= vmin) load(v - 1, 0);
            const int wl = w0 + lane + 64 * (kChunk - 1);
            // the chunk's products with every query of the pass that uses row v, over its first KK
            // groups of 64 entries (the groups past the row's end hold zeros and are skipped)
            auto products = [&](auto KKc) {
                constexpr int KK = decltype(KKc)::value;
#pragma unroll
                for (int j = 0; j < kServeJ; ++j) {
                    if ((on >> j) & 1) {
                        const int M = __builtin_amdgcn_readlane(qM, j);
                        const double h = readlane_f64(H, j);
                        const double* ip = invp + (M - wl);  // entry k at ip[64 (kChunk-1-k)]
                        double pr[KK];
#pragma unroll
                        for (int k = 0; k < KK; ++k) pr[k] = mulred(cur[k], ip[64 * (kChunk - 1 - k)], m);
                        // tree sum: up to 8 terms below 0.62 p each stay below 5 p
#pragma unroll
                        for (int st = 1; st < KK; st <<= 1)
#pragma unroll
                            for (int k = 0; k + st < KK; k += 2 * st) pr[k] += pr[k + st];
                        acc[j] = reduce(acc[j] + mulred(reduce(pr[0], m), h, m), m);
                    }
                }
            };
            // groups of 64 entries that reach into row v, rounded up to a power of two (uniform)
            const int kk = (v - w0 + 63) >> 6;
            if (kk > 4) products(std::integral_constant<int, 8>{});
            else if (kk > 2) products(std::integral_constant<int, 4>{});
            else if (kk > 1) products(std::integral_constant<int, 2>{});
            else products(std::integral_constant<int, 1>{});
        }
        // H_j(v-1) = H_j(v) Q / ((n-v+1)(M-v+1)) for the queries that used row v (a query whose
        // last row n is below v keeps H_j(n), set at the start)
        if (lane < jn && qlive && v <= qn)
            H = reduce(mulred(mulred(H, Q, m), mulred(invp[qn - v + 1], invp[qM - v + 1], m), m), m);
    }
#pragma unroll
    for (int j = 0; j < kServeJ; ++j) {
        if (j < jn) {
            const double t = wave_sum(acc[j], m);
            if (lane == 0) a.part[((int64_t)b * a.ranges + range) * a.J + j0 + j] = t;
        }
    }
}

__global__ void k_serve_finish(ServeArgs a, uint64_t* __restrict__ res) {
    const int b = blockIdx.x;
    const Mod m{a.primes[b].p, a.primes[b].pinv};
    const double* fact = a.store + a.stride * b;
    for (int j = threadIdx.x; j < a.J; j += blockDim.x) {
        const ServeQuery q = a.q[j];
        uint64_t r = 0;
        if (b < q.need) {
            double s = 0.0;
            const double* p = a.part + (int64_t)b * a.ranges * a.J + j;
            for (int w = 0; w < a.ranges; ++w) {
                s += p[(int64_t)w * a.J];
                if ((w & 3) == 3) s = reduce(s, m);
            }
            r = (uint64_t)canon(mulred(mulred(reduce(s, m), fact[q.M], m), fact[q.n], m), m);
        }
        res[(int64_t)b * a.J + j] = r;
    }
}

}  // namespace

// The array itself. In host code its name only stands for the symbol: device addresses come from
// hipGetSymbolAddress, per device, and only those are handed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[kQueryScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kQueryScratchDevices] = {};
}  // namespace

hipError_t attach_query_scratch(int device) {
    if (device < 0 || device >= kQueryScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t st = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (st != hipSuccess) return st;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void* query_scratch(int device) {
    if (device < 0 || device >= kQueryScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

hipError_t launch_serve(const ServeArgs& a, uint64_t* res, hipStream_t stream) {
    if (a.J <= 0 || a.np <= 0) return hipSuccess;
    if (a.passes != (a.J + kServeJ - 1) / kServeJ || a.groups * kWaves != a.ranges) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(((a.np + 7) / 8) * 8 * a.passes * a.groups));
    const size_t lds = sizeof(double) * (size_t)(a.F1 + kPad);
    hipError_t st = hipFuncSetAttribute((const void*)k_serve_rows, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (st != hipSuccess) return st;
    k_serve_rows<<<grid, kServeThreads, lds, stream>>>(a);
    st = hipGetLastError();
    if (st != hipSuccess) return st;
    k_serve_finish<<<dim3((unsigned)a.np), 64, 0, stream>>>(a, res);
    return hipGetLastError();
}

}  // namespace cf
