// cf_separate.hip -- the separate-tables model (d tables of lengths m_1..m_d, each item one slot
// per table; the paper's Section 4, include/cuckoo/separate_tables.hpp) on the device, one
// workgroup per prime, behind `cuckoo_fail separate -n N -m m1,..,md [--algo suffix|forward]`:
//   suffix:  Algorithm 5 (separate_tables.hpp:175-291): the suffix table V[p][v] over the states
//            v >= 1, |v| < p <= n, with the inner sums evaluated as the generating function
//            -log(1 - sigma) by Horner's rule in sigma = x_1 + .. + x_d, then N_bad for m;
//   forward: Algorithm 4 (separate_tables.hpp:27-108), the recurrence over the boundary states.
// The state sets (integer vectors, independent of the prime) are enumerated on the host and
// uploaded once per call; everything modular runs on the device with the exact FP64 product.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "cf_core.hpp"
#include "cf_engine.hpp"

namespace cf {
namespace {

constexpr int kSepThreads = 256;
constexpr int kSepMaxD = 8;

struct SepShape {
    int algo, n, d, M, N, F;  // N states; F factorial entries (max(M, n) + 1)
    int m[kSepMaxD];
    int64_t T;       // suffix: entries of V; forward: boundary states
    int64_t stride;  // doubles of arena per prime
};

// Host-built state sets, read by every workgroup.
struct SepStates {
    const int* coord;  // suffix: v (d per state); forward: b (d per state)
    const int* sum;    // |v| / |b|
    const int* prev;   // suffix: index of v - e_t or -1 (d per state)
    const int* cnt;    // suffix: states with |v| <= s, s = 0 .. n-1
    const int64_t* voff;  // suffix: start of row p of V, p = 1 .. n
    const int* lev;    // forward: l of each state
    const int* lstart; // forward: first state with the same l
};

__device__ __forceinline__ double wsum(double x, const Mod& m) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x = reduce(x + __shfl_xor(x, o, 64), m);
    return x;
}

__device__ double block_sum(double x, double* red, const Mod& m) {
    x = wsum(x, m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    double s = 0;
    for (int w = 0; w < kSepThreads / 64; ++w) s = reduce(s + red[w], m);
    __syncthreads();
    return s;  // every thread
}

struct Fac {
    const double *f, *fi;
    Mod m;
    __device__ __forceinline__ double binom(int a, int b) const {
        if (a < 0 || b < 0 || b > a) return 0.0;
        return mulred(mulred(f[a], fi[b], m), fi[a - b], m);
    }
};

__device__ void factorials(double* fact, double* ifact, int F, const Mod& m, uint64_t p) {
    if (threadIdx.x == 0) {
        fact[0] = 1.0;
        for (int i = 1; i < F; ++i) fact[i] = canon(mulred(fact[i - 1], (double)i, m), m);
        ifact[F - 1] = canon(powmod(fact[F - 1], p - 2, m), m);
        for (int i = F - 1; i > 0; --i) ifact[i - 1] = canon(mulred(ifact[i], (double)i, m), m);
    }
    __syncthreads();
}

// phi(a, b, r) of the coefficient formula (single_table.hpp:25-31)
__device__ __forceinline__ double phi(const Fac& fc, int a, int b, int r) {
    if (b <= 0) return fc.binom(a - 1, r);
    return -mulred(fc.binom(a - b - 1, r - b), fc.binom(r - 1, b - 1), fc.m);
}

__device__ void suffix(const SepShape& s, const SepStates& st, double* a, const Mod& m, uint64_t p, double* red,
                       uint64_t* out) {
    const int n = s.n, d = s.d, N = s.N, t = (int)threadIdx.x;
    double *fact = a, *ifact = fact + s.F, *inv = ifact + s.F, *g = inv + (n + 1), *vf = g + N, *vfi = vf + N;
    double *V = vfi + N, *X = V + s.T, *R0 = X + N, *R1 = R0 + N;
    factorials(fact, ifact, s.F, m, p);
    const Fac fc{fact, ifact, m};
    for (int k = 1 + t; k <= n; k += kSepThreads) inv[k] = canon(mulred(ifact[k], fact[k - 1], m), m);  // 1/k
    for (int i = t; i < N; i += kSepThreads) {
        double gi = 1.0, f = 1.0, fi = 1.0;
        for (int u = 0; u < d; ++u) {
            const int v = st.coord[(int64_t)i * d + u];
            gi = mulred(gi, (double)v, m);
            f = mulred(f, fact[v], m);
            fi = mulred(fi, ifact[v], m);
        }
        g[i] = canon(gi, m), vf[i] = canon(f, m), vfi[i] = canon(fi, m);
    }
    __syncthreads();
    // V[p][v] = C(p-1, |v|) g(v)^p before the inner sums are taken off
    for (int pp = 1; pp <= n; ++pp)
        for (int i = t; i < st.cnt[pp - 1]; i += kSepThreads)
            V[st.voff[pp] + i] = canon(mulred(fc.binom(pp - 1, st.sum[i]), powmod(g[i], (uint64_t)pp, m), m), m);
    __syncthreads();
    for (int q = 1; q < n; ++q) {
        // row q is final: X_q[w] = V[q][w] / (prod w_t! (q - |w| - 1)!) on |w| <= q - 1
        const int nq = st.cnt[q - 1];
        for (int i = t; i < nq; i += kSepThreads)
            X[i] = mulred(mulred(V[st.voff[q] + i], vfi[i], m), ifact[q - 1 - st.sum[i]], m);
        __syncthreads();
        // Horner in sigma: R_s = X_q / s + sigma R_{s+1} on max(d, q-s) <= |v| <= n-1-s. Each
        // step reads only what the step before wrote (its range shifted by one in |v|), so two
        // buffers alternate; the first step reads nothing (no state has |v| = d - 1).
        double *Ro = R0, *Rn = R1;
        for (int sg = n - 1 - d; sg >= 1; --sg) {
            const int lo = max(d, q - sg), hi = n - 1 - sg;
            const int ib = st.cnt[lo - 1], ie = st.cnt[hi];
            const double is = inv[sg];
            for (int i = ib + t; i < ie; i += kSepThreads) {
                double acc = 0;
                for (int u = 0; u < d; ++u) {
                    const int j = st.prev[(int64_t)i * d + u];
                    if (j >= 0) acc += Ro[j];
                }
                acc = reduce(acc, m);
                if (i < nq) acc = reduce(acc + mulred(X[i], is, m), m);
                Rn[i] = acc;
            }
            __syncthreads();
            double* tmp = Ro;
            Ro = Rn, Rn = tmp;
        }
        // Y = sigma R_1 on |v| >= q; S[v][q] = prod v_t! / (|v| - q)! Y[v], taken off V[p][v]
        for (int i = nq + t; i < N; i += kSepThreads) {
            double acc = 0;
            for (int u = 0; u < d; ++u) {
                const int j = st.prev[(int64_t)i * d + u];
                if (j >= 0) acc += Ro[j];
            }
            acc = reduce(acc, m);
            if (acc == 0.0) continue;
            const int sv = st.sum[i];
            const double S = mulred(mulred(acc, vf[i], m), ifact[sv - q], m);
            double gp = powmod(g[i], (uint64_t)(sv + 1 - q), m);
            for (int pp = sv + 1; pp <= n; ++pp) {
                const double c = mulred(mulred(fc.binom(pp - q - 1, pp - sv - 1), fc.binom(pp, q), m), gp, m);
                double& x = V[st.voff[pp] + i];
                x = reduce(x - mulred(c, S, m), m);
                gp = mulred(gp, g[i], m);
            }
        }
        __syncthreads();
    }
    // N_bad = sum_q C(n, q) Q^(n-q) sum_{|w| < q} C(M-|w|-1, M-q) prod_t C(m_t, w_t) V[q][w]
    double Q = 1.0;
    for (int u = 0; u < d; ++u) Q = mulred(Q, (double)s.m[u], m);
    double acc = 0;
    for (int q = 1; q <= n; ++q) {
        double inner = 0;
        for (int i = t; i < st.cnt[q - 1]; i += kSepThreads) {
            double c = fc.binom(s.M - st.sum[i] - 1, s.M - q);
            for (int u = 0; u < d && c != 0.0; ++u) c = mulred(c, fc.binom(s.m[u], st.coord[(int64_t)i * d + u]), m);
            inner = reduce(inner + mulred(c, V[st.voff[q] + i], m), m);
        }
        acc = reduce(acc + mulred(mulred(fc.binom(n, q), powmod(Q, (uint64_t)(n - q), m), m), inner, m), m);
    }
    const double total = block_sum(acc, red, m);
    if (t == 0) *out = (uint64_t)canon(total, m);
}

// Algorithm 4: F over the boundary states (0, 0) and (l, b), z <= l < M, in order of l; each
// state sums over the earlier levels' states with b' <= b and |b'| < |b|.
__device__ void forward(const SepShape& s, const SepStates& st, double* a, const Mod& m, uint64_t p, double* red,
                        uint64_t* out) {
    const int n = s.n, d = s.d, M = s.M, z = M - n, t = (int)threadIdx.x;
    const int64_t K = s.T;
    double *fact = a, *ifact = fact + s.F, *h = ifact + s.F, *Fw = h + K;
    factorials(fact, ifact, s.F, m, p);
    const Fac fc{fact, ifact, m};
    for (int64_t i = t; i < K; i += kSepThreads) {
        double x = 1.0;
        for (int u = 0; u < d; ++u) x = mulred(x, (double)(s.m[u] - st.coord[i * d + u]), m);
        h[i] = canon(x, m);  // h(b) = prod_t (m_t - b_t), eq. (16)
    }
    if (t == 0) Fw[0] = 1.0;
    __syncthreads();
    for (int64_t i = 1; i < K; ++i) {
        const int li = st.lev[i], bi = st.sum[i];
        const int* b = st.coord + i * d;
        double acc = 0;
        for (int64_t j = t; j < st.lstart[i]; j += kSepThreads) {
            const double fj = Fw[j];
            if (fj == 0.0 || st.sum[j] >= bi) continue;
            const int* bj = st.coord + j * d;
            bool le = true;
            for (int u = 0; u < d && le; ++u) le = bj[u] <= b[u];
            if (!le) continue;
            const int lj = st.lev[j];
            double tau = phi(fc, bi - lj, st.sum[j] - lj, li - lj);
            for (int u = 0; u < d && tau != 0.0; ++u) tau = mulred(tau, fc.binom(s.m[u] - bj[u], b[u] - bj[u]), m);
            if (tau == 0.0) continue;
            const int u = li - max(lj, z), nu = lj > z ? lj - z : 0;
            tau = mulred(mulred(tau, fc.binom(n - nu, u), m), powmod(h[j], (uint64_t)u, m), m);
            acc = reduce(acc + mulred(fj, tau, m), m);
        }
        const double f = block_sum(acc, red, m);
        if (t == 0) Fw[i] = canon(f, m);
        __syncthreads();
    }
    double acc = 0;
    for (int64_t i = t; i < K; i += kSepThreads) {
        if (Fw[i] == 0.0) continue;
        const int li = st.lev[i], nu = li > z ? li - z : 0;
        acc = reduce(acc + mulred(mulred(Fw[i], fc.binom(M - li - 1, M - st.sum[i]), m),
                                  powmod(h[i], (uint64_t)(n - nu), m), m), m);
    }
    const double total = block_sum(acc, red, m);
    if (t == 0) *out = (uint64_t)canon(total, m);
}

__global__ void __launch_bounds__(kSepThreads) k_separate(SepShape s, SepStates st, const uint64_t* __restrict__ primes,
                                                          double* __restrict__ mem, uint64_t* __restrict__ res) {
    __shared__ double red[kSepThreads / 64];
    const uint64_t p = primes[blockIdx.x];
    Mod m;
    m.p = (double)p;
    m.pinv = 1.0 / m.p;
    double* a = mem + s.stride * blockIdx.x;
    if (s.algo == CF_ALGO_SUFFIX) suffix(s, st, a, m, p, red, res + blockIdx.x);
    else forward(s, st, a, m, p, red, res + blockIdx.x);
}

// ---- host: the state sets ----

// Every v >= 1 with |v| <= smax in order of increasing |v| (lexicographic inside one |v|).
void compositions(int d, int smax, std::vector<int>& coord, std::vector<int>& sum) {
    std::vector<int> v(d);
    for (int s = d; s <= smax; ++s) {
        std::fill(v.begin(), v.end(), 1);
        v[d - 1] = s - (d - 1);
        for (;;) {
            coord.insert(coord.end(), v.begin(), v.end());
            sum.push_back(s);
            int rest = v[d - 1], t = d - 2;
            while (t >= 0 && rest <= d - 1 - t) rest += v[t--];
            if (t < 0) break;
            ++v[t];
            --rest;
            for (int u = t + 1; u < d - 1; ++u) v[u] = 1, --rest;
            v[d - 1] = rest;
        }
    }
}

}  // namespace
}  // namespace cf

using namespace cf_int;

extern "C" int cf_run_separate(cf_engine* e, int algo, int n, const int32_t* m, int d, const uint64_t* primes, size_t np,
                               uint64_t* res) {
    if (!e) return CF_E_ARGS;
    e->err.clear();
    e->stats = cf_stats{};
    if ((algo != CF_ALGO_SUFFIX && algo != CF_ALGO_FORWARD) || d < 1 || !m || n < 1 || (np && (!primes || !res)))
        return fail(e, CF_E_ARGS, "bad arguments");
    if (d > cf::kSepMaxD) return fail(e, CF_E_UNSUPPORTED, "more than 8 tables");
    int64_t M = 0;
    for (int u = 0; u < d; ++u) {
        if (m[u] < 1) return fail(e, CF_E_ARGS, "separate tables: lengths must be positive");
        M += m[u];
    }
    if (M <= n || M > (1 << 20)) return fail(e, CF_E_ARGS, "separate tables: need 1 <= n < M");
    for (size_t b = 0; b < np; ++b)
        if (primes[b] < 3 || !(primes[b] & 1) || primes[b] >= (uint64_t(1) << 50) || primes[b] <= (uint64_t)M)
            return fail(e, CF_E_ARGS, "prime " + std::to_string(primes[b]) + " is not an odd modulus in (M, 2^50)");
    cf::SepShape s{};
    s.algo = algo, s.n = n, s.d = d, s.M = (int)M, s.F = (int)std::max<int64_t>(M, n) + 1;
    for (int u = 0; u < d; ++u) s.m[u] = m[u];
    // the state sets, as ints in one upload
    std::vector<int> coord, sum, prev, cnt, lev, lstart;
    std::vector<int64_t> voff;
    if (algo == CF_ALGO_SUFFIX) {
        // |v| <= n - 1: C(n - 1, d) states; the table rows p = 1 .. n hold cnt[p-1] of them
        double states = 1;
        for (int u = 0; u < d; ++u) states = states * (n - 1 - u) / (u + 1);
        if (states > 4e6 || (double)n * states > 2e9)
            return fail(e, CF_E_UNSUPPORTED, "separate suffix table beyond this build's cross-check size");
        cf::compositions(d, n - 1, coord, sum);
        s.N = (int)sum.size();
        cnt.assign(n, 0);
        for (int x = 0, i = 0; x < n; ++x) {
            while (i < s.N && sum[i] <= x) ++i;
            cnt[x] = i;
        }
        voff.assign(n + 1, 0);
        int64_t T = 0;
        for (int pp = 1; pp <= n; ++pp) voff[pp] = T, T += cnt[pp - 1];
        s.T = T;
        // v - e_t by a key of the vector (radix n + 1)
        std::vector<int> index;
        double cells = 1;
        for (int u = 0; u < d; ++u) cells *= n + 1;
        const bool dense = cells < 6e7;
        std::map<std::vector<int>, int> sparse;
        if (dense) index.assign((size_t)cells, -1);
        auto key = [&](const int* v) {
            int64_t k = 0;
            for (int u = 0; u < d; ++u) k = k * (n + 1) + v[u];
            return k;
        };
        for (int i = 0; i < s.N; ++i) {
            if (dense) index[(size_t)key(&coord[(size_t)i * d])] = i;
            else sparse[std::vector<int>(coord.begin() + (size_t)i * d, coord.begin() + (size_t)(i + 1) * d)] = i;
        }
        prev.assign((size_t)s.N * d, -1);
        std::vector<int> w(d);
        for (int i = 0; i < s.N; ++i)
            for (int u = 0; u < d; ++u) {
                if (coord[(size_t)i * d + u] < 2) continue;
                std::copy(coord.begin() + (size_t)i * d, coord.begin() + (size_t)(i + 1) * d, w.begin());
                --w[u];
                prev[(size_t)i * d + u] = dense ? index[(size_t)key(w.data())] : sparse.at(w);
            }
        s.stride = 2 * (int64_t)s.F + (n + 1) + 3 * (int64_t)s.N + s.T + 3 * (int64_t)s.N;
    } else {
        // boundary states: (0, 0), then for l = z .. M-1 every v = m - b with 1 <= v_t <= m_t, |v| < M - l
        const int z = (int)M - n;
        std::vector<int> vc, vs;
        cf::compositions(d, (int)M - 1 - z, vc, vs);
        coord.assign(d, 0), sum.push_back(0), lev.push_back(0), lstart.push_back(0);
        for (int l = z; l < M; ++l) {
            const int first = (int)sum.size();
            for (size_t i = 0; i < vs.size() && vs[i] < M - l; ++i) {
                bool ok = true;
                for (int u = 0; u < d && ok; ++u) ok = vc[i * d + u] <= m[u];
                if (!ok) continue;
                int bs = 0;
                for (int u = 0; u < d; ++u) coord.push_back(m[u] - vc[i * d + u]), bs += m[u] - vc[i * d + u];
                sum.push_back(bs), lev.push_back(l), lstart.push_back(first);
            }
            if (sum.size() > (1u << 20))
                return fail(e, CF_E_UNSUPPORTED, "separate forward recurrence beyond this build's cross-check size");
        }
        s.T = (int64_t)sum.size();
        s.N = 0;
        s.stride = 2 * (int64_t)s.F + 2 * s.T;
    }
    s.stride = (s.stride + 31) & ~(int64_t)31;
    if (!np) return CF_OK;
    HIP_TRY(e, hipSetDevice(e->device), "hipSetDevice");
    // one small buffer: primes, results, then the state arrays
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_p = 0, o_r = up(8 * np), o_c = o_r + up(8 * np), o_s = o_c + up(4 * coord.size());
    const size_t o_pv = o_s + up(4 * sum.size()), o_cn = o_pv + up(4 * prev.size()), o_vo = o_cn + up(4 * cnt.size());
    const size_t o_lv = o_vo + up(8 * voff.size()), o_ls = o_lv + up(4 * lev.size()), total = o_ls + up(4 * lstart.size());
    int rc = ensure(e, &e->small, &e->small_cap, total, "the state sets");
    if (rc) return rc;
    char* sm = (char*)e->small;
    auto put = [&](size_t off, const void* src, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpyAsync(sm + off, src, bytes, hipMemcpyHostToDevice, e->stream) : hipSuccess;
    };
    HIP_TRY(e, put(o_c, coord.data(), 4 * coord.size()), "upload");
    HIP_TRY(e, put(o_s, sum.data(), 4 * sum.size()), "upload");
    HIP_TRY(e, put(o_pv, prev.data(), 4 * prev.size()), "upload");
    HIP_TRY(e, put(o_cn, cnt.data(), 4 * cnt.size()), "upload");
    HIP_TRY(e, put(o_vo, voff.data(), 8 * voff.size()), "upload");
    HIP_TRY(e, put(o_lv, lev.data(), 4 * lev.size()), "upload");
    HIP_TRY(e, put(o_ls, lstart.data(), 4 * lstart.size()), "upload");
    const cf::SepStates st{(const int*)(sm + o_c), (const int*)(sm + o_s), (const int*)(sm + o_pv), (const int*)(sm + o_cn),
                           (const int64_t*)(sm + o_vo), (const int*)(sm + o_lv), (const int*)(sm + o_ls)};
    const size_t per = sizeof(double) * (size_t)s.stride;
    const size_t fit = std::min<size_t>(usable_bytes(e) / per, 65535);
    if (!fit) return fail(e, CF_E_NOMEM, "one prime needs " + std::to_string(per) + " bytes of device memory");
    float ms_total = 0;
    for (size_t at = 0; at < np; at += fit) {
        const size_t k = std::min(fit, np - at);
        rc = ensure(e, &e->arena, &e->arena_cap, k * per, "the separate-table arrays");
        if (rc) return rc;
        HIP_TRY(e, put(o_p, primes + at, 8 * k), "upload");
        HIP_TRY(e, hipEventRecord(e->ev0, e->stream), "hipEventRecord");
        cf::k_separate<<<dim3((unsigned)k), cf::kSepThreads, 0, e->stream>>>(s, st, (const uint64_t*)(sm + o_p),
                                                                            (double*)e->arena, (uint64_t*)(sm + o_r));
        HIP_TRY(e, hipGetLastError(), "kernel launch");
        HIP_TRY(e, hipEventRecord(e->ev1, e->stream), "hipEventRecord");
        HIP_TRY(e, hipStreamSynchronize(e->stream), "kernel execution");
        float ms = 0;
        HIP_TRY(e, hipEventElapsedTime(&ms, e->ev0, e->ev1), "hipEventElapsedTime");
        ms_total += ms;
        HIP_TRY(e, hipMemcpy(res + at, sm + o_r, 8 * k, hipMemcpyDeviceToHost), "download");
        e->stats.launches++;
        e->stats.tasks += (int)k;
        e->stats.device_bytes = std::max(e->stats.device_bytes, k * per);
    }
    e->stats.batches = e->stats.launches;
    e->stats.kernel_ms = ms_total;
    return CF_OK;
}
