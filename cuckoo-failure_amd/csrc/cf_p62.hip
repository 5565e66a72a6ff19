// cf_p62.hip -- the one-table suffix-table recurrence modulo primes of up to 62 bits: the residues
// of the reference's -DFP_RB=64 build (fast/cuckoo_fast.cu:48-60, 86-99; its prime list of
// c 2^20 + 1 below 2^62, fast/README.md:26-27), which the FP64 product of the main engine cannot
// hold (its exact range ends at 2^50).
//
// Arithmetic: residues are 64-bit integers in Montgomery form x R mod p, R = 2^64, p < 2^63 odd.
// The product a b R^-1 mod p is Montgomery's reduction of the 128-bit a b: with q = (a b mod R)
// p^-1 mod R, a b - q p is divisible by R and (a b - q p) / R = hi(a b) - hi(q p) lies in (-p, p)
// (the low words cancel exactly), so one conditional add of p gives [0, p). Sums are reduced by
// one conditional subtraction (p < 2^63, so a + b does not wrap).
//
// The recurrence is the engine's rescaled form (cf_core.hpp): stored tables P2[p][w] = p V[p][w] /
// (w! p! (p-1-w)!) and S2[v][q] = S[v][q] / (q! v! (v-q)!), step v:
//   stage 1 (v > v0):  a[w] = P2[v][w] / v (w < v);  c = a (*) k;  S2[e][v] = c[e] / (e-v)!^2, v <= e < nm
//   stage 2 (v >= v0): b[q] = S2[v][q] v! g^(v-q) (1 <= q <= v, none at v = v0);  c = b (*) k;
//                      P2[p][v] = (g^v/v! - p c[p mod L]) g^(p-v)/v! / (p-v-1)!^2,  v < p <= nm
//   queries:           tot_j <- tot_j Q_j + n_j!/((n_j-v)! (M_j-v)!) sum_w a[w] / (M_j - w)   (n_j >= v)
// with k[t] = 1/t (1 <= t < nm) and the result N_bad = M_j! tot_j. One workgroup runs one prime:
// the cyclic convolution of length L = max(16, next_pow2(nm)) <= 2^14 in LDS (DIF forward, natural
// in and bit-reversed out; pointwise product with the kernel's spectrum, made by the same forward
// and scaled by 1/L; DIT inverse back to natural order; two radix-2 levels per LDS pass). The tables are packed
// triangles in HBM (P2 row p at p(p-1)/2, S2 row e at e(e-1)/2 with entry q at q-1); the 16
// columns of the current column block go to column-major staging first and are copied into the
// rows when the block ends, so the column writes of every step are contiguous.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "cf_engine.hpp"
#include "cf_p62.hpp"

namespace cf {

namespace {

using u64 = uint64_t;

__device__ __forceinline__ u64 mul_hi(u64 a, u64 b) { return __umul64hi(a, b); }

__device__ __forceinline__ u64 mmul(u64 a, u64 b, const Mod62& m) {
    const u64 lo = a * b, hi = mul_hi(a, b);
    const u64 q = lo * m.pinv;
    const u64 h2 = mul_hi(q, m.p);
    return hi >= h2 ? hi - h2 : hi - h2 + m.p;
}
__device__ __forceinline__ u64 madd(u64 a, u64 b, const Mod62& m) {
    const u64 s = a + b;
    return s >= m.p ? s - m.p : s;
}
__device__ __forceinline__ u64 msub(u64 a, u64 b, const Mod62& m) { return a >= b ? a - b : a - b + m.p; }
__device__ __forceinline__ u64 mpow(u64 b, u64 e, const Mod62& m) {
    u64 r = m.one;
    for (; e; e >>= 1) {
        if (e & 1) r = mmul(r, b, m);
        b = mmul(b, b, m);
    }
    return r;
}

__device__ __forceinline__ u64 shfl_xor_u64(u64 x, int lane_mask) {
    const unsigned lo = __shfl_xor((unsigned)x, lane_mask, 64), hi = __shfl_xor((unsigned)(x >> 32), lane_mask, 64);
    return ((u64)hi << 32) | lo;
}

constexpr int kQG = 8;  // queries reduced together

template <int LG>
struct P62 {
    static constexpr int L = 1 << LG;
    static constexpr int T = L / 2 < 64 ? 64 : (L / 2 < kP62Threads ? L / 2 : kP62Threads);  // whole waves
    static constexpr int NW = T / 64;
    static constexpr int HI = L / 64 + 2;
};

// The arena of one task (64-bit words), in this order.
struct Arena62 {
    u64 *fact, *ifact, *inv, *ifact2, *mont;  // F1 each (Montgomery form; mont[i] = i R mod p)
    u64 *tw, *twi;                            // L each: tw[h + j] = w_{2h}^j (and inverses)
    u64* hat;                                 // L: spectrum of the kernel times 1/L (bit-reversed order)
    u64 *Vt, *St;                             // packed triangles
    u64 *Vs, *Ss;                             // staging: 16 columns x NS rows, column-major
    u64* qtot;                                // 2 per query: Horner accumulator, Q = g(M)
};
__device__ __forceinline__ Arena62 arena62(u64* mem, const Task62& d) {
    Arena62 a;
    const int64_t F1 = d.F1, L = d.L, NS = d.nm + 1, tri = (int64_t)d.nm * (d.nm + 1) / 2;
    a.fact = mem + d.arena;
    a.ifact = a.fact + F1;
    a.inv = a.ifact + F1;
    a.ifact2 = a.inv + F1;
    a.mont = a.ifact2 + F1;
    a.tw = a.mont + F1;
    a.twi = a.tw + L;
    a.hat = a.twi + L;
    a.Vt = a.hat + L;
    a.St = a.Vt + tri;
    a.Vs = a.St + tri;
    a.Ss = a.Vs + 16 * NS;
    a.qtot = a.Ss + 16 * NS;
    return a;
}

template <int LG>
struct Work62 {
    using PP = P62<LG>;
    static constexpr int L = PP::L, T = PP::T, NW = PP::NW;
    const Task62& d;
    const Mod62 m;
    const Arena62 a;
    const QueryDesc* q;
    u64* buf;
    u64 *glo, *ghi, *red;
    int t;

    __device__ void sync() const { __syncthreads(); }
    __device__ u64 g_of(int v) const {
        if (d.model) return mpow(a.mont[v], (u64)d.dd, m);
        if (v < d.dd) return 0;
        return mmul(mmul(a.fact[v], a.ifact[d.dd], m), a.ifact[v - d.dd], m);
    }
    __device__ u64 gpow(int e) const { return mmul(ghi[e >> 6], glo[e & 63], m); }
    // g^i (i < 64) and g^(64 j) (64 j <= nm + 64)
    __device__ void power_tables(u64 g) {
        const int nhi = (d.nm >> 6) + 2;
        for (int i = t; i < 64 + nhi; i += T) {
            if (i < 64) glo[i] = mpow(g, (u64)i, m);
            else ghi[i - 64] = mpow(mpow(g, 64, m), (u64)(i - 64), m);
        }
        sync();
    }
    // Forward DIF, natural order in, bit-reversed out: the radix-2 levels h = L/2, L/4, .., 1 (pairs
    // (i, i + h) in blocks of 2h, twiddle tw[h + j] on the difference), taken two at a time as
    // radix-4 passes over groups {i, i + h/2, i + h, i + 3h/2} (one LDS round trip and one barrier
    // per two levels), and a last radix-2 pass when log2 L is odd.
    __device__ void fwd() {
        int h = L / 2;
        for (; h >= 2; h >>= 2) {
            const int q = h >> 1;
            for (int k = t; k < L / 4; k += T) {
                const int j = k & (q - 1), i = ((k - j) << 2) + j;
                const u64 x0 = buf[i], x1 = buf[i + q], x2 = buf[i + h], x3 = buf[i + h + q];
                const u64 y0 = madd(x0, x2, m), y2 = mmul(msub(x0, x2, m), a.tw[h + j], m);
                const u64 y1 = madd(x1, x3, m), y3 = mmul(msub(x1, x3, m), a.tw[h + j + q], m);
                const u64 w = a.tw[q + j];
                buf[i] = madd(y0, y1, m);
                buf[i + q] = mmul(msub(y0, y1, m), w, m);
                buf[i + h] = madd(y2, y3, m);
                buf[i + h + q] = mmul(msub(y2, y3, m), w, m);
            }
            sync();
        }
        if (h == 1) {
            for (int k = t; k < L / 2; k += T) {
                const u64 x = buf[2 * k], y = buf[2 * k + 1];
                buf[2 * k] = madd(x, y, m);
                buf[2 * k + 1] = msub(x, y, m);  // tw[1] = 1
            }
            sync();
        }
    }
    // Inverse DIT, bit-reversed in, natural out: the levels h = 1, 2, .., L/2 (pairs (i, i + h),
    // the odd entry times twi[h + j]), a first radix-2 pass when log2 L is odd, then two levels
    // (h, 2h) per radix-4 pass over groups {i, i + h, i + 2h, i + 3h}.
    __device__ void inv() {
        int h = 1;
        if (LG & 1) {
            for (int k = t; k < L / 2; k += T) {
                const u64 x = buf[2 * k], y = buf[2 * k + 1];
                buf[2 * k] = madd(x, y, m);
                buf[2 * k + 1] = msub(x, y, m);  // twi[1] = 1
            }
            sync();
            h = 2;
        }
        for (; h < L; h <<= 2) {
            for (int k = t; k < L / 4; k += T) {
                const int j = k & (h - 1), i = ((k - j) << 2) + j;
                const u64 wh = a.twi[h + j];
                const u64 x0 = buf[i], x1 = mmul(buf[i + h], wh, m), x2 = buf[i + 2 * h], x3 = mmul(buf[i + 3 * h], wh, m);
                const u64 y0 = madd(x0, x1, m), y1 = msub(x0, x1, m), y2 = madd(x2, x3, m), y3 = msub(x2, x3, m);
                const u64 t2 = mmul(y2, a.twi[2 * h + j], m), t3 = mmul(y3, a.twi[2 * h + j + h], m);
                buf[i] = madd(y0, t2, m);
                buf[i + 2 * h] = msub(y0, t2, m);
                buf[i + h] = madd(y1, t3, m);
                buf[i + 3 * h] = msub(y1, t3, m);
            }
            sync();
        }
    }
    // buf = buf (*) k (cyclic, length L), natural order in and out
    __device__ void convolve() {
        fwd();
        for (int i = t; i < L; i += T) buf[i] = mmul(buf[i], a.hat[i], m);
        sync();
        inv();
    }
    // entry (r, c), c < r, of a table: the triangle row when column block c/16 has been copied
    // into the rows (it ends before row block r/16), else the staging column c % 16
    __device__ u64 v_at(int r, int c) const {
        return (c >> 4) < (r >> 4) ? a.Vt[(int64_t)r * (r - 1) / 2 + c] : a.Vs[(int64_t)(c & 15) * (d.nm + 1) + r];
    }
    __device__ u64 s_at(int r, int c) const {  // 1 <= c <= r
        return (c >> 4) < (r >> 4) ? a.St[(int64_t)r * (r - 1) / 2 + c - 1] : a.Ss[(int64_t)(c & 15) * (d.nm + 1) + r];
    }
    // stage-1 input of row v into buf
    __device__ void stage1_in(int v) {
        const u64 iv = a.inv[v];
        for (int w = t; w < L; w += T) buf[w] = w < v ? mmul(v_at(v, w), iv, m) : 0;
        sync();
    }
    // query terms of row v (in buf) for the queries with n >= v
    __device__ void queries(int v) {
        int nact = 0;
        while (nact < d.q_cnt && q[d.q_off + nact].n >= v) ++nact;
        for (int j0 = 0; j0 < nact; j0 += kQG) {
            const int cnt = nact - j0 < kQG ? nact - j0 : kQG;
            u64 acc[kQG];
#pragma unroll
            for (int jj = 0; jj < kQG; ++jj) acc[jj] = 0;
            for (int w = t; w < v; w += T) {
                const u64 x = buf[w];
#pragma unroll
                for (int jj = 0; jj < kQG; ++jj)
                    if (jj < cnt) acc[jj] = madd(acc[jj], mmul(x, a.inv[q[d.q_off + j0 + jj].M - w], m), m);
            }
#pragma unroll
            for (int jj = 0; jj < kQG; ++jj) {
                if (jj < cnt) {  // cnt is uniform: whole waves take the same branches
                    u64 s = acc[jj];
                    for (int o = 32; o >= 1; o >>= 1) s = madd(s, shfl_xor_u64(s, o), m);
                    if ((t & 63) == 0) red[(t >> 6) * kQG + jj] = s;
                }
            }
            sync();
            if (t < cnt) {
                u64 s = 0;
                for (int wv = 0; wv < NW; ++wv) s = madd(s, red[wv * kQG + t], m);
                const int j = j0 + t;
                const QueryDesc& qq = q[d.q_off + j];
                const u64 coef = mmul(mmul(a.fact[qq.n], a.ifact[qq.n - v], m), a.ifact[qq.M - v], m);
                a.qtot[2 * j] = madd(mmul(a.qtot[2 * j], a.qtot[2 * j + 1], m), mmul(coef, s, m), m);
            }
            sync();
        }
    }
    // column block X (columns 16 X .. 16 X + 15) from staging into the triangles' rows
    __device__ void flush(int X) {
        const int nm = d.nm, NS = nm + 1, c0 = 16 * X;
        for (int r = c0 + 1 + t; r <= nm; r += T) {  // P2 rows r > c
            u64* row = a.Vt + (int64_t)r * (r - 1) / 2;
            for (int k = 0; k < 16 && c0 + k < r; ++k) row[c0 + k] = a.Vs[(int64_t)k * NS + r];
        }
        for (int r = c0 + t; r < nm; r += T) {  // S2 rows r >= q, q >= 1
            u64* row = a.St + (int64_t)r * (r - 1) / 2;
            for (int k = 0; k < 16 && c0 + k <= r; ++k)
                if (c0 + k >= 1) row[c0 + k - 1] = a.Ss[(int64_t)k * NS + r];
        }
        sync();
    }

    __device__ void run() {
        const int nm = d.nm, NS = nm + 1, v0 = d.model ? 1 : d.dd;
        // staging to zero (column 0 of P2 is never computed: V[p][0] = 0), query state, spectrum
        for (int i = t; i < 32 * NS; i += T) a.Vs[i] = 0;
        for (int j = t; j < d.q_cnt; j += T) {
            a.qtot[2 * j] = 0;
            a.qtot[2 * j + 1] = g_of(q[d.q_off + j].M);
        }
        for (int i = t; i < L; i += T) buf[i] = (i >= 1 && i < nm) ? a.inv[i] : 0;
        sync();
        fwd();
        for (int i = t; i < L; i += T) a.hat[i] = mmul(buf[i], d.linv, m);
        sync();
        for (int v = 1; v < nm; ++v) {
            const int col = v & 15;
            u64* scol = a.Ss + (int64_t)col * NS;
            u64* vcol = a.Vs + (int64_t)col * NS;
            if (v > v0) {
                stage1_in(v);
                queries(v);
                convolve();
                for (int e = v + t; e < nm; e += T) scol[e] = mmul(buf[e], a.ifact2[e - v], m);
            } else {
                for (int e = v + t; e < nm; e += T) scol[e] = 0;
            }
            sync();
            if (v >= v0) {
                power_tables(g_of(v));
                const bool conv = v > v0;
                if (conv) {
                    const u64 fv = a.fact[v];
                    for (int i = t; i < L; i += T) buf[i] = (i >= 1 && i <= v) ? mmul(mmul(s_at(v, i), fv, m), gpow(v - i), m) : 0;
                    sync();
                    convolve();
                }
                const u64 iv = a.ifact[v], s1 = mmul(gpow(v), iv, m);
                for (int p = v + 1 + t; p <= nm; p += T) {
                    const u64 pc = conv ? mmul(a.mont[p], buf[p & (L - 1)], m) : 0;
                    const u64 f = mmul(mmul(gpow(p - v), iv, m), a.ifact2[p - v - 1], m);
                    vcol[p] = mmul(msub(s1, pc, m), f, m);
                }
            } else {
                for (int p = v + 1 + t; p <= nm; p += T) vcol[p] = 0;
            }
            sync();
            if (col == 15) flush(v >> 4);
        }
        // row nm of P2 feeds only the queries with n = nm
        if (nm > v0) {
            stage1_in(nm);
            queries(nm);
        }
    }
};

template <int LG>
__global__ void __launch_bounds__(kP62Threads)
    k_p62(const Task62* __restrict__ tasks, const QueryDesc* __restrict__ qs, u64* __restrict__ mem, u64* __restrict__ res) {
    using PP = P62<LG>;
    __shared__ u64 buf[PP::L];
    __shared__ u64 glo[64], ghi[PP::HI], red[PP::NW * kQG];
    const Task62& d = tasks[blockIdx.x];
    Work62<LG> w{d, d.m, arena62(mem, d), qs, buf, glo, ghi, red, (int)threadIdx.x};
    w.run();
    for (int j = threadIdx.x; j < d.q_cnt; j += PP::T) {
        const QueryDesc& qq = qs[d.q_off + j];
        res[qq.out] = mmul(mmul(w.a.qtot[2 * j], w.a.fact[qq.M], d.m), 1, d.m);  // out of Montgomery form
    }
}

template <int LG>
hipError_t launch_lg(int ntasks, const Task62* tasks, const QueryDesc* qs, u64* mem, u64* res, hipStream_t stream) {
    k_p62<LG><<<dim3((unsigned)ntasks), P62<LG>::T, 0, stream>>>(tasks, qs, mem, res);
    return hipGetLastError();
}

}  // namespace

int64_t arena62_words(int nm, int F1, int L, int q_cnt) {
    return 5 * (int64_t)F1 + 3 * (int64_t)L + (int64_t)nm * (nm + 1) + 32 * (int64_t)(nm + 1) + 2 * (int64_t)q_cnt;
}

hipError_t launch_p62(int log2L, int ntasks, const Task62* tasks, const QueryDesc* qs, uint64_t* mem, uint64_t* res,
                      hipStream_t stream) {
    if (ntasks <= 0) return hipSuccess;
    switch (log2L) {
        case 4: return launch_lg<4>(ntasks, tasks, qs, mem, res, stream);
        case 5: return launch_lg<5>(ntasks, tasks, qs, mem, res, stream);
        case 6: return launch_lg<6>(ntasks, tasks, qs, mem, res, stream);
        case 7: return launch_lg<7>(ntasks, tasks, qs, mem, res, stream);
        case 8: return launch_lg<8>(ntasks, tasks, qs, mem, res, stream);
        case 9: return launch_lg<9>(ntasks, tasks, qs, mem, res, stream);
        case 10: return launch_lg<10>(ntasks, tasks, qs, mem, res, stream);
        case 11: return launch_lg<11>(ntasks, tasks, qs, mem, res, stream);
        case 12: return launch_lg<12>(ntasks, tasks, qs, mem, res, stream);
        case 13: return launch_lg<13>(ntasks, tasks, qs, mem, res, stream);
        case 14: return launch_lg<14>(ntasks, tasks, qs, mem, res, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cf
