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
// the cyclic convolution of length L = max(16, next_pow2(nm)) <= 2^17, in LDS up to 2^14 and in the
// task's HBM arena beyond (every pass still loads and stores whole groups; DIF forward, natural
// in and bit-reversed out; pointwise product with the kernel's spectrum, made by the same forward
// and scaled by 1/L; DIT inverse back to natural order; radix-16 register passes). The tables are
// kept as the main engine keeps them (cf_core.hpp tab_off, cf_host.cpp plan_slots): 16 x 16 tiles
// in a pool whose 32 KB segments are reused once their row block has been read (about nm^2/4
// entries per table), with the columns of the current column block staged column-major (two
// blocks, by column-block parity) and copied into their tiles when the block ends, so the column
// writes of every step are contiguous.
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
// a v mod p for a constant v < p with q' = floor(v 2^64 / p) (Shoup's product): with q =
// hi(a q'), a v - q p lies in [0, 2p) and is exact mod 2^64 (2p < 2^64), one conditional
// subtraction. With a in Montgomery form and v plain, the result is the Montgomery form of the
// product, the residue mmul gives for v's Montgomery form, in about two thirds of its instructions.
__device__ __forceinline__ u64 smul(u64 a, u64 v, u64 vq, const Mod62& m) {
    const u64 r = a * v - mul_hi(a, vq) * m.p;
    return r >= m.p ? r - m.p : r;
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

// The transform of length L = 2^LG as register passes: F16 = LG / 4 passes of radix 16 and, when
// LG % 4 = r > 0, one pass of radix RP = 2^r (the last forward pass, the first inverse pass).
// A radix-16 pass has G1 = L / 16 groups; thread t holds groups t, t + T, .. (at most 512 threads,
// so each has up to 256 registers: two groups each at L = 2^14, eight at 2^17).
template <int LG>
struct P62 {
    static constexpr int L = 1 << LG;
    static constexpr int G1 = L / 16;
    static constexpr int T = G1 < 64 ? 64 : (G1 < 512 ? G1 : 512);  // whole waves
    static constexpr bool HBM = LG > 14;  // the transform's buffer in the arena, not in LDS
    static constexpr int NW = T / 64;
    static constexpr int HI = L / 64 + 2;
    static constexpr int F16 = LG / 4, RP = 1 << (LG % 4), NF = F16 + (LG % 4 ? 1 : 0);
    static constexpr int RM = LG % 4 ? RP : 16;  // radix of the last forward / first inverse pass
    static constexpr int NSEG = ((L + 1 + 15) / 16 + 15) / 16;  // segments per row block, nm <= L
    static_assert(T <= kP62Threads, "workgroup size");
};

// The arena of one task (64-bit words), in this order.
struct Arena62 {
    u64 *fact, *ifact, *inv, *ifact2, *mont;  // F1 each (Montgomery form; mont[i] = i R mod p)
    u64 *tw, *twi;                            // L each: tw[h + j] = w_{2h}^j (and inverses)
    u64* shc;                                 // 32: w_16^r, their Shoup quotients, w_16^-r, quotients
    u64* hat;                                 // L: spectrum of the kernel times 1/L (bit-reversed order)
    u64 *Vpool, *Spool;                       // tiles of 16 x 16 entries, row-major
    u64 *Vs, *Ss;                             // staging: 2 blocks x 16 columns x NS rows, column-major
    u64* qtot;                                // 2 per query: Horner accumulator, Q = g(M)
    u64* tbuf;                                // L: the transform's buffer when it is not in LDS
};
__device__ __forceinline__ Arena62 arena62(u64* mem, const Task62& d) {
    Arena62 a;
    const int64_t F1 = d.F1, L = d.L, NS = d.NS, tiles = 256 * d.pool;
    a.fact = mem + d.arena;
    a.ifact = a.fact + F1;
    a.inv = a.ifact + F1;
    a.ifact2 = a.inv + F1;
    a.mont = a.ifact2 + F1;
    a.tw = a.mont + F1;
    a.twi = a.tw + L;
    a.shc = a.twi + L;
    a.hat = a.shc + 32;
    a.Vpool = a.hat + L;
    a.Spool = a.Vpool + tiles;
    a.Vs = a.Spool + tiles;
    a.Ss = a.Vs + 32 * NS;
    a.qtot = a.Ss + 32 * NS;
    a.tbuf = a.qtot + 2 * (int64_t)d.q_cnt;
    return a;
}

template <int LG>
struct Work62 {
    using PP = P62<LG>;
    static constexpr int L = PP::L, T = PP::T, NW = PP::NW, G1 = PP::G1, F16 = PP::F16, NF = PP::NF, RM = PP::RM;
    const Task62& d;
    const Mod62 m;
    const Arena62 a;
    const QueryDesc* q;
    u64* buf;
    u64 *glo, *ghi, *red;
    u64 *W, *Wi, *wsm, *wsmi, *c16;  // LDS: w^(16x), w^-(16x) (x < G1), w^b, w^-b (b < 16) for w of order
                                     // L; the Shoup constants of w_16^r and w_16^-r (r < 8), as in shc
    const int* slots;            // segment slots of the plan, Y * NSEG + s
    int* segrow;                 // LDS: the segment slots of the current row block
    int t;

    __device__ __forceinline__ void sync() const { __syncthreads(); }
    __device__ __forceinline__ u64 g_of(int v) const {
        if (d.model) return mpow(a.mont[v], (u64)d.dd, m);
        if (v < d.dd) return 0;
        return mmul(mmul(a.fact[v], a.ifact[d.dd], m), a.ifact[v - d.dd], m);
    }
    // g^e v! and g^e / v! (e <= nm): the low table carries the step's factor
    __device__ __forceinline__ u64 gpow_f(int e) const { return mmul(ghi[e >> 6], glo[64 + (e & 63)], m); }
    __device__ __forceinline__ u64 gpow_i(int e) const { return mmul(ghi[e >> 6], glo[128 + (e & 63)], m); }
    // g^i, g^i v!, g^i / v! (i < 64) and g^(64 j) (64 j <= nm + 64) for g = g(v); the caller's next
    // barrier publishes them
    __device__ __forceinline__ void power_tables(u64 g, int v) {
        const int nhi = (d.nm >> 6) + 2;
        for (int i = t; i < 64 + nhi; i += T) {
            if (i < 64) {
                const u64 gi = mpow(g, (u64)i, m);
                glo[i] = gi;
                glo[64 + i] = mmul(gi, a.fact[v], m);
                glo[128 + i] = mmul(gi, a.ifact[v], m);
            } else {
                ghi[i - 64] = mpow(mpow(g, 64, m), (u64)(i - 64), m);
            }
        }
    }
    // Forward DIF, natural order in, bit-reversed out, by radix-2 levels h = L/2, .., 1 (pairs (i, i + h)
    // in blocks of 2h, twiddle tw[h + j] on the difference), two levels per LDS pass. Used once per
    // task, for the kernel's spectrum; the v-loop runs the same butterflies as the register passes below.
    __device__ __forceinline__ void fwd() {
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

    // w^e (tw) or w^-e (twi) for w of order L, e < L: tw[L/2 + e] = w^e below L/2, and w^(L/2) = -1
    __device__ __forceinline__ u64 wexp(const u64* tab, int e) const {
        return e < L / 2 ? tab[L / 2 + e] : msub(0, tab[e], m);
    }
    // ---- the v-loop's transforms: groups of R entries in registers, one round trip through the
    //      buffer (LDS, or HBM beyond 2^14) per pass ----
    // Buffer word of transform entry i: bits 0-4 XORed with bits 4-8 (a bijection of [0, L)), so the
    // strided groups of every pass spread over the LDS banks.
    __device__ __forceinline__ static int sw(int i) { return i ^ ((i >> 4) & 31); }
    template <int R>
    __device__ __forceinline__ void ld(u64 (&x)[16], int i0, int s) const {
#pragma unroll
        for (int k = 0; k < R; ++k) x[k] = buf[sw(i0 + k * s)];
    }
    template <int R>
    __device__ __forceinline__ void st(const u64 (&x)[16], int i0, int s) const {
#pragma unroll
        for (int k = 0; k < R; ++k) buf[sw(i0 + k * s)] = x[k];
    }
    // A radix-R pass over the group x[k] = entry i0 + k S (i0 = j mod S within its block) does the
    // butterflies of log2 R radix-2 levels. Their twiddles factor into constants (roots of order
    // 16) and one power of w_L per entry: a DIF pass is the constant DIF network of R entries, then
    // entry k times w_2h^(j rev(k)) (rev: the log2 R bits of k reversed); a DIT pass is entry k
    // times w_(Rh)^-(j rev(k)), then the constant DIT network. Same products as the radix-2 levels,
    // word for word.
    // Constant DIF network: level l pairs (k, k + e), e = R / 2^(l+1), the difference times
    // w_16^((16/R) 2^l r) for r = k mod e (none for r = 0), by Shoup's product.
    template <int R>
    __device__ __forceinline__ void dif_c(u64 (&x)[16]) const {
#pragma unroll
        for (int l = 0, e = R / 2; e >= 1; ++l, e >>= 1)
#pragma unroll
            for (int k0 = 0; k0 < R; k0 += 2 * e)
#pragma unroll
                for (int r = 0; r < e; ++r) {
                    const u64 u = x[k0 + r], w = x[k0 + r + e], df = msub(u, w, m);
                    x[k0 + r] = madd(u, w, m);
                    const int i = ((16 / R) << l) * r;
                    x[k0 + r + e] = r ? smul(df, c16[i], c16[8 + i], m) : df;
                }
    }
    // Constant DIT network: level l pairs (k, k + e), e = 2^l, the second entry times
    // w_16^-((8 >> l) r) for r = k mod e (none for r = 0), by Shoup's product.
    template <int R>
    __device__ __forceinline__ void dit_c(u64 (&x)[16]) const {
#pragma unroll
        for (int l = 0, e = 1; e < R; ++l, e <<= 1)
#pragma unroll
            for (int k0 = 0; k0 < R; k0 += 2 * e)
#pragma unroll
                for (int r = 0; r < e; ++r) {
                    const int i = (8 >> l) * r;
                    const u64 u = x[k0 + r], o = r ? smul(x[k0 + r + e], c16[16 + i], c16[24 + i], m) : x[k0 + r + e];
                    x[k0 + r] = madd(u, o, m);
                    x[k0 + r + e] = msub(u, o, m);
                }
    }
    static constexpr int rev4(int k) { return ((k & 1) << 3) | ((k & 2) << 1) | ((k & 4) >> 1) | ((k & 8) >> 3); }
    // Entry k times b^rev4(k), k = 1..15: b^2, b^4, b^8 by squaring, the other powers by one
    // product each (14 products); each power is applied as soon as it is made (the power e goes to
    // entry rev4(e)), so at most eight are live.
    __device__ __forceinline__ void twiddle_chain(u64 (&x)[16], u64 b) const {
        u64 pw[8];
        pw[1] = b;
        pw[2] = mmul(b, b, m);
        pw[3] = mmul(pw[2], b, m);
        pw[4] = mmul(pw[2], pw[2], m);
#pragma unroll
        for (int e = 5; e < 8; ++e) pw[e] = mmul(pw[4], pw[e - 4], m);
#pragma unroll
        for (int e = 1; e < 8; ++e) x[rev4(e)] = mmul(x[rev4(e)], pw[e], m);
        const u64 p8 = mmul(pw[4], pw[4], m);
        x[rev4(8)] = mmul(x[rev4(8)], p8, m);
#pragma unroll
        for (int e = 9; e < 16; ++e) x[rev4(e)] = mmul(x[rev4(e)], mmul(p8, pw[e - 8], m), m);
    }
    // Entry k times tab[step rev4(k)], k = 1..15 (w_L^(+-16 step rev4(k)) from W or Wi).
    __device__ __forceinline__ void twiddle_tab(u64 (&x)[16], const u64* tab, int step) const {
#pragma unroll
        for (int k = 1; k < 16; ++k) x[k] = mmul(x[k], tab[step * rev4(k)], m);
    }
    // w_L^j and w_L^-j for j < L/16: the roots of order L for j mod 16 times W[j/16]
    __device__ __forceinline__ u64 wpow(int j) const { return mmul(wsm[j & 15], W[j >> 4], m); }
    __device__ __forceinline__ u64 wpowi(int j) const { return mmul(wsmi[j & 15], Wi[j >> 4], m); }
    // Forward passes P .. NF-2 (radix 16, h = L/2 / 16^P, twiddles w_L^(16^P j rev4(k)) from W), one
    // phase each.
    template <int P>
    __device__ __forceinline__ void fwd_mid() {
        if constexpr (P <= NF - 2) {
            constexpr int H = (L / 2) >> (4 * P), S = 2 * H / 16, F = 1 << (4 * (P - 1));
#pragma unroll 1
            for (int k = t; k < G1; k += T) {
                const int j = k & (S - 1), i0 = (k - j) * 16 + j;  // block (k / S) of 2H = 16 S entries
                u64 x[16];
                ld<16>(x, i0, S);
                dif_c<16>(x);
                twiddle_tab(x, W, F * j);
                st<16>(x, i0, S);
            }
            sync();
            fwd_mid<P + 1>();
        }
    }
    // Inverse passes P .. NF-2 (radix 16; h = RP 16^(P-1) after a partial first pass, else 16^P;
    // twiddles w_L^-((L / 16h) j rev4(k)) from Wi), one phase each.
    template <int P>
    __device__ __forceinline__ void inv_mid() {
        if constexpr (P <= NF - 2) {
            constexpr int H = (LG % 4 ? PP::RP : 16) << (4 * (P - 1)), F = L / (256 * H);
#pragma unroll 1
            for (int k = t; k < G1; k += T) {
                const int j = k & (H - 1), i0 = (k - j) * 16 + j;  // block (k / H) of 16 H entries
                u64 x[16];
                ld<16>(x, i0, H);
                twiddle_tab(x, Wi, F * j);
                dit_c<16>(x);
                st<16>(x, i0, H);
            }
            sync();
            inv_mid<P + 1>();
        }
    }
    // The middle of a cyclic convolution with the kernel's spectrum, from the phase after the first
    // forward pass (or, when L = 16, after the input alone) to the phase before the last inverse pass.
    __device__ __forceinline__ void conv_mid() {
        fwd_mid<1>();
        // the last forward pass, the product with the spectrum and the first inverse pass share
        // their groups (RM consecutive entries, j = 0)
        for (int k = t; k < L / RM; k += T) {
            u64 y[16];
            ld<RM>(y, k * RM, 1);
            dif_c<RM>(y);  // j = 0: no twiddles beyond the constants
#pragma unroll
            for (int i = 0; i < RM; ++i) y[i] = mmul(y[i], a.hat[k * RM + i], m);
            dit_c<RM>(y);
            st<RM>(y, k * RM, 1);
        }
        sync();
        inv_mid<1>();
    }
    // The first forward pass (h = L/2, j = k) of group k (entries k + i G1) from registers into the buffer.
    __device__ __forceinline__ void first_pass(u64 (&x)[16], int k) {
        if constexpr (NF >= 2) {
            dif_c<16>(x);
            twiddle_chain(x, wpow(k));
        }
        st<16>(x, k, G1);
    }
    // The last inverse pass (h = L/16, j = k) of group k: x[i] = result entry k + i G1.
    __device__ __forceinline__ void last_pass(u64 (&x)[16], int k) {
        ld<16>(x, k, G1);
        if constexpr (NF >= 2) {
            twiddle_chain(x, wpowi(k));
            dit_c<16>(x);
        }
    }
    // Entry (r, c) of a table, row block r/16 the one whose segment slots are in segrow: its tile
    // when column block c/16 has been flushed (it ends before row block r/16), else the staging
    // column of c. Branch-free, so a phase's loads are all in flight together.
    __device__ __forceinline__ int64_t tab_ix(int r, int c) const {
        const int Y = r >> 4, X = c >> 4;
        const int64_t tile = (int64_t)segrow[X >> 4] * 16 + (X & 15);
        return X < Y ? tile * 256 + ((r & 15) << 4) + (c & 15) : st_col(c, d.NS) + r;
    }
    __device__ __forceinline__ u64 v_at(int r, int c) const {
        return ((c >> 4) < (r >> 4) ? a.Vpool : a.Vs)[tab_ix(r, c)];
    }
    __device__ __forceinline__ u64 s_at(int r, int c) const {  // 1 <= c <= r
        return ((c >> 4) < (r >> 4) ? a.Spool : a.Ss)[tab_ix(r, c)];
    }
    // stage-1 input a[w] = P2[v][w] / v, w < v
    __device__ __forceinline__ u64 in1(int v, int w, u64 iv) const { return mmul(v_at(v, w), iv, m); }
    // stage-1 input of group k into x (entries w = k + i G1)
    __device__ __forceinline__ void in1_grp(u64 (&x)[16], int v, int k, u64 iv) const {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int w = k + i * G1;
            const u64 y = v_at(v, w < v ? w : 0);
            x[i] = w < v ? mmul(y, iv, m) : 0;
        }
    }
    // stage-1 input of row v into buf (the last row's queries)
    __device__ __forceinline__ void stage1_in(int v) {
        const u64 iv = a.inv[v];
        for (int w = t; w < L; w += T) buf[sw(w)] = w < v ? in1(v, w, iv) : 0;
        sync();
    }
    // Adds the terms of the first cnt <= 8 queries over the entries x[i] (w = k + i G1) of row v.
    __device__ __forceinline__ void query_add(u64 (&acc)[kQG], const u64 (&x)[16], int v, int k, int cnt) const {
#pragma unroll
        for (int jj = 0; jj < kQG; ++jj) {
            if (jj < cnt) {  // cnt is uniform: whole waves take the same branches
                const int M = q[d.q_off + jj].M;
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int w = k + i * G1;
                    acc[jj] = madd(acc[jj], mmul(x[i], a.inv[w < v ? M - w : 1], m), m);  // x[i] = 0 for w >= v
                }
            }
        }
    }
    // The wave sums of acc into red[wave][jj].
    __device__ __forceinline__ void query_wave(const u64 (&acc)[kQG], int cnt) {
#pragma unroll
        for (int jj = 0; jj < kQG; ++jj) {
            if (jj < cnt) {
                u64 s = acc[jj];
                for (int o = 32; o >= 1; o >>= 1) s = madd(s, shfl_xor_u64(s, o), m);
                if ((t & 63) == 0) red[(t >> 6) * kQG + jj] = s;
            }
        }
    }
    // Thread t < cnt adds the wave sums of query j0 + t and updates its Horner accumulator.
    __device__ __forceinline__ void query_finish(int v, int j0, int cnt) {
        if (t < cnt) {
            u64 s = 0;
            for (int wv = 0; wv < NW; ++wv) s = madd(s, red[wv * kQG + t], m);
            const int j = j0 + t;
            const QueryDesc& qq = q[d.q_off + j];
            const u64 coef = mmul(mmul(a.fact[qq.n], a.ifact[qq.n - v], m), a.ifact[qq.M - v], m);
            a.qtot[2 * j] = madd(mmul(a.qtot[2 * j], a.qtot[2 * j + 1], m), mmul(coef, s, m), m);
        }
    }
    // query terms of row v (in buf) for the queries with n >= v
    __device__ __forceinline__ void queries(int v) {
        int nact = 0;
        while (nact < d.q_cnt && q[d.q_off + nact].n >= v) ++nact;
        for (int j0 = 0; j0 < nact; j0 += kQG) {
            const int cnt = nact - j0 < kQG ? nact - j0 : kQG;
            u64 acc[kQG];
#pragma unroll
            for (int jj = 0; jj < kQG; ++jj) acc[jj] = 0;
            for (int w = t; w < v; w += T) {
                const u64 x = buf[sw(w)];
#pragma unroll
                for (int jj = 0; jj < kQG; ++jj)
                    if (jj < cnt) acc[jj] = madd(acc[jj], mmul(x, a.inv[q[d.q_off + j0 + jj].M - w], m), m);
            }
#pragma unroll
            for (int jj = 0; jj < kQG; ++jj) {
                if (jj < cnt) {
                    u64 s = acc[jj];
                    for (int o = 32; o >= 1; o >>= 1) s = madd(s, shfl_xor_u64(s, o), m);
                    if ((t & 63) == 0) red[(t >> 6) * kQG + jj] = s;
                }
            }
            sync();
            query_finish(v, j0, cnt);
            sync();
        }
    }
    // Column block X from staging into its tiles (Y, X) of both tables, for the row blocks Y > X,
    // one row of 16 entries per thread and table; then the segment slots of row block X + 1 (the
    // rows of the next 16 steps) into segrow.
    __device__ __forceinline__ void flush(int X) {
        const int NS = d.NS, r0 = 16 * (X + 1);
        const int64_t st = st_col(16 * X, NS);
        for (int i = t; i < 2 * (NS - r0); i += T) {
            const int tab = i >= NS - r0, r = r0 + (tab ? i - (NS - r0) : i);
            const u64* src = (tab ? a.Ss : a.Vs) + st + r;
            const int64_t tile = (int64_t)slots[d.slot_off + (r >> 4) * d.NSEG + (X >> 4)] * 16 + (X & 15);
            u64* dst = (tab ? a.Spool : a.Vpool) + tile * 256 + ((r & 15) << 4);
            u64 y[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) y[k] = src[(int64_t)k * NS];
#pragma unroll
            for (int k = 0; k < 16; ++k) dst[k] = y[k];
        }
        for (int s = t; s < d.NSEG; s += T) segrow[s] = slots[d.slot_off + (X + 1) * d.NSEG + s];
        sync();
    }

    __device__ __forceinline__ void run() {
        const int nm = d.nm, NS = d.NS, v0 = d.model ? 1 : d.dd;
        // staging to zero (column 0 of P2 is never computed: V[p][0] = 0), query state, spectrum,
        // the twiddle tables of the register passes
        for (int i = t; i < 64 * NS; i += T) a.Vs[i] = 0;
        for (int j = t; j < d.q_cnt; j += T) {
            a.qtot[2 * j] = 0;
            a.qtot[2 * j + 1] = g_of(q[d.q_off + j].M);
        }
        for (int x = t; x < G1; x += T) W[x] = wexp(a.tw, 16 * x), Wi[x] = wexp(a.twi, 16 * x);
        if (t < 16) wsm[t] = wexp(a.tw, t), wsmi[t] = wexp(a.twi, t);
        if (t < 32) c16[t] = a.shc[t];
        for (int i = t; i < L; i += T) buf[i] = (i >= 1 && i < nm) ? a.inv[i] : 0;
        sync();
        fwd();
        for (int i = t; i < L; i += T) a.hat[i] = mmul(buf[i], d.linv, m);
        sync();
        for (int v = 1; v < nm; ++v) {
            const int col = v & 15;
            u64* scol = a.Ss + st_col(v, NS);
            u64* vcol = a.Vs + st_col(v, NS);
            int nact = 0;
            while (nact < d.q_cnt && q[d.q_off + nact].n >= v) ++nact;
            if (v > v0) {
                // stage-1 input of row v, the query sums, the first forward pass; g's power tables
                const u64 iv = a.inv[v];
                power_tables(g_of(v), v);
                if (nact > kQG) {  // many queries: their sums from the input in LDS first
                    for (int k = t; k < G1; k += T) {
                        u64 x[16];
                        in1_grp(x, v, k, iv);
                        st<16>(x, k, G1);
                    }
                    sync();
                    queries(v);
                    for (int k = t; k < G1; k += T) {
                        u64 x[16];
                        ld<16>(x, k, G1);
                        first_pass(x, k);
                    }
                } else {
                    u64 acc[kQG];
#pragma unroll
                    for (int jj = 0; jj < kQG; ++jj) acc[jj] = 0;
                    for (int k = t; k < G1; k += T) {
                        u64 x[16];
                        in1_grp(x, v, k, iv);
                        query_add(acc, x, v, k, nact);
                        first_pass(x, k);
                    }
                    query_wave(acc, nact);  // finished after the convolution
                }
                sync();
                conv_mid();
                // last inverse pass; stage-1 output S2[e][v] = c[e] / (e - v)!^2, v <= e < nm
                for (int k = t; k < G1; k += T) {
                    u64 x[16];
                    last_pass(x, k);
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const int e = k + i * G1;
                        if (e >= v && e < nm) scol[e] = mmul(x[i], a.ifact2[e - v], m);
                    }
                }
                if (nact > 0 && nact <= kQG) query_finish(v, 0, nact);
            } else {
                for (int e = v + t; e < nm; e += T) scol[e] = 0;
                if (v == v0) power_tables(g_of(v), v);
            }
            sync();
            if (v >= v0) {
                const u64 s1 = gpow_i(v);  // g^v / v!
                if (v > v0) {
                    // stage-2 input b[i] = S2[v][i] v! g^(v-i), 1 <= i <= v; the first forward pass
                    for (int k = t; k < G1; k += T) {
                        u64 x[16];
#pragma unroll
                        for (int i = 0; i < 16; ++i) {
                            const int e = k + i * G1;
                            const bool ok = e >= 1 && e <= v;
                            const int ce = ok ? e : 1;
                            const u64 y = s_at(v, ce);
                            x[i] = ok ? mmul(y, gpow_f(v - ce), m) : 0;
                        }
                        first_pass(x, k);
                    }
                    sync();
                    conv_mid();
                    // last inverse pass; stage-2 output P2[p][v], v < p <= nm, from entry p mod L
                    for (int k = t; k < G1; k += T) {
                        u64 x[16];
                        last_pass(x, k);
#pragma unroll
                        for (int i = 0; i < 16; ++i) {
                            const int e = k + i * G1, p = e == 0 ? L : e;
                            if (p > v && p <= nm) {
                                const u64 pc = mmul(a.mont[p], x[i], m);
                                const u64 f = mmul(gpow_i(p - v), a.ifact2[p - v - 1], m);
                                vcol[p] = mmul(msub(s1, pc, m), f, m);
                            }
                        }
                    }
                } else {
                    for (int p = v + 1 + t; p <= nm; p += T) {
                        const u64 f = mmul(gpow_i(p - v), a.ifact2[p - v - 1], m);
                        vcol[p] = mmul(s1, f, m);
                    }
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
__global__ void __launch_bounds__(P62<LG>::T)
    k_p62(const Task62* __restrict__ tasks, const QueryDesc* __restrict__ qs, const int* __restrict__ slots,
          u64* __restrict__ mem, u64* __restrict__ res) {
    using PP = P62<LG>;
    __shared__ u64 lbuf[PP::HBM ? 1 : PP::L];
    __shared__ u64 glo[3 * 64], ghi[PP::HI], red[PP::NW * kQG];
    __shared__ u64 W[PP::G1], Wi[PP::G1], wsm[16], wsmi[16], c16[32];
    __shared__ int segrow[PP::NSEG];
    const Task62& d = tasks[blockIdx.x];
    const Arena62 ar = arena62(mem, d);
    u64* buf = PP::HBM ? ar.tbuf : lbuf;
    Work62<LG> w{d, d.m, ar, qs, buf, glo, ghi, red, W, Wi, wsm, wsmi, c16, slots, segrow, (int)threadIdx.x};
    w.run();
    for (int j = threadIdx.x; j < d.q_cnt; j += PP::T) {
        const QueryDesc& qq = qs[d.q_off + j];
        res[qq.out] = mmul(mmul(w.a.qtot[2 * j], w.a.fact[qq.M], d.m), 1, d.m);  // out of Montgomery form
    }
}

template <int LG>
hipError_t launch_lg(int ntasks, const Task62* tasks, const QueryDesc* qs, const int* slots, u64* mem, u64* res,
                     hipStream_t stream) {
    k_p62<LG><<<dim3((unsigned)ntasks), P62<LG>::T, 0, stream>>>(tasks, qs, slots, mem, res);
    return hipGetLastError();
}

}  // namespace

int64_t arena62_words(int NS, int64_t pool, int F1, int L, int q_cnt) {
    return 5 * (int64_t)F1 + 3 * (int64_t)L + 32 + 512 * pool + 64 * (int64_t)NS + 2 * (int64_t)q_cnt +
           (L > (1 << 14) ? (int64_t)L : 0);
}

hipError_t launch_p62(int log2L, int ntasks, const Task62* tasks, const QueryDesc* qs, const int* slots, uint64_t* mem,
                      uint64_t* res, hipStream_t stream) {
    if (ntasks <= 0) return hipSuccess;
    switch (log2L) {
        case 4: return launch_lg<4>(ntasks, tasks, qs, slots, mem, res, stream);
        case 5: return launch_lg<5>(ntasks, tasks, qs, slots, mem, res, stream);
        case 6: return launch_lg<6>(ntasks, tasks, qs, slots, mem, res, stream);
        case 7: return launch_lg<7>(ntasks, tasks, qs, slots, mem, res, stream);
        case 8: return launch_lg<8>(ntasks, tasks, qs, slots, mem, res, stream);
        case 9: return launch_lg<9>(ntasks, tasks, qs, slots, mem, res, stream);
        case 10: return launch_lg<10>(ntasks, tasks, qs, slots, mem, res, stream);
        case 11: return launch_lg<11>(ntasks, tasks, qs, slots, mem, res, stream);
        case 12: return launch_lg<12>(ntasks, tasks, qs, slots, mem, res, stream);
        case 13: return launch_lg<13>(ntasks, tasks, qs, slots, mem, res, stream);
        case 14: return launch_lg<14>(ntasks, tasks, qs, slots, mem, res, stream);
        case 15: return launch_lg<15>(ntasks, tasks, qs, slots, mem, res, stream);
        case 16: return launch_lg<16>(ntasks, tasks, qs, slots, mem, res, stream);
        case 17: return launch_lg<17>(ntasks, tasks, qs, slots, mem, res, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cf
