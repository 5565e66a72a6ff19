// cf_core.hpp -- the per-thread building blocks of the suffix-table engine.
//
// Everything here is written once and compiled twice: by hipcc for gfx950 (the product kernels
// in cf_kernels.hip call these functions with the thread index and put a workgroup barrier
// between phases) and by a host compiler for a CPU emulation that runs the same phases thread by
// thread (a debugging build, never a product path).
//
// Arithmetic. Residues are doubles holding integers. A prime p < 2^50 leaves three spare bits of
// the 53-bit mantissa, so sums up to 8p are exact. The product a*b is split exactly into
// h + l with one FMA (h = fl(ab), l = fma(a, b, -h)); q = rint(h / p) is the nearest integer to
// ab/p up to a small error, and r = fma(-q, p, h) + l = ab - qp is computed without rounding.
// With |a|, |b| bounded the result is bounded (see mulred), so values are kept "balanced"
// (roughly in (-p, p)) and only brought to [0, p) at the very end.
//
// The algorithm is the one-table suffix-table recurrence of the reference
// (include/cuckoo/single_table.hpp:80-96, 178-229; fast/cuckoo_fast.cu:145-243) in a rescaled
// form in which every factorial that depends on both indices of an entry is applied once, when
// the entry is written, and never looked up when it is read:
//   V2[p][w] = V[p][w] / (w! p! (p-1-w)!)     (stored suffix table)
//   S2[v][q] = S[v][q] / (q! v! (v-q)!)       (stored shared inner sums)
// Step v, stage 1 (column v of S2 from row v of V2):
//   a[w]  = V2[v][w]                               = U[v][w] / v!,   w < v
//   c     = a (*) k,  k[t] = 1/t (1 <= t < nm)     cyclic, length L >= nm
//   S2[v'][v] = c[v'] / (v'-v)!^2                  v <= v' < nm
// Step v, stage 2 (column v of V2 from row v of S2), g = g(v):
//   b[q]  = S2[v][q] * v! g^(v-q)                  1 <= q <= v
//   c     = b (*) k
//   V2[p][v] = (g^v/v! * 1/p - c[p mod L]) * g^(p-v)/v! / (p-v-1)!^2     v < p <= nm
// and the query sum of row v (fast/cuckoo_fast.cu:306-310) is sum_w a[w] / (M - w) = out[v] / v!.
// The table actually stored is P2[p][w] = p V2[p][w], so
//   P2[p][v] = (g^v/v! - p c[p mod L]) * g^(p-v)/v! / (p-v-1)!^2
// needs no per-entry 1/p (p c[p] is one product by the integer p), and the row read at step v is
// scaled back by the one number 1/v.
#pragma once
#include <stdint.h>

#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CF_HD __host__ __device__ __forceinline__
#else
#define CF_HD inline
#endif

namespace cf {

// Compile-time loops: f(std::integral_constant<int, i>) for i = A .. B-1 (static_for) or
// i = B-1 down to A (static_rfor), so a loop index can be a template argument.
template <int A, int B, class F>
CF_HD void static_for(F&& f) {
    if constexpr (A < B) {
        f(std::integral_constant<int, A>{});
        static_for<A + 1, B>(f);
    }
}
template <int A, int B, class F>
CF_HD void static_rfor(F&& f) {
    if constexpr (A < B) {
        f(std::integral_constant<int, B - 1>{});
        static_rfor<A, B - 1>(f);
    }
}

// ------------------------------------------------------------------------------------------
// Exact modular arithmetic on doubles
// ------------------------------------------------------------------------------------------
struct Mod {
    double p;     // the prime, p < 2^50
    double pinv;  // fl(1 / p)
};

// a*b reduced: returns ab - qp with q = rint(fl(ab) * pinv). For |a| <= A p, |b| <= B p the result
// satisfies |r| <= p/2 + 3 * 2^-53 |ab| <= (1/2 + 3 A B / 8) p, and it is exact whenever that bound
// is below 2^53 (the partial h - qp is an integer of about the same size).
#ifdef CF_EMU_CHECK
// CPU emulation only: verify every product exactly (a b - r divisible by p, |r| below 2^52).
void emu_check_mulred(double a, double b, double r, const Mod& m);
#define CF_CHECK_MULRED(a, b, r, m) emu_check_mulred(a, b, r, m)
#else
#define CF_CHECK_MULRED(a, b, r, m)
#endif

CF_HD double mulred(double a, double b, const Mod& m) {
    const double h = a * b;
    const double l = __builtin_fma(a, b, -h);
    const double q = __builtin_rint(h * m.pinv);
    const double r = __builtin_fma(-q, m.p, h);
    CF_CHECK_MULRED(a, b, r + l, m);
    return r + l;
}
// x - rint(x/p) p: for |x| <= 16p the result lies in [-p/2 - 4, p/2 + 4].
CF_HD double reduce(double x, const Mod& m) {
    const double q = __builtin_rint(x * m.pinv);
    return __builtin_fma(-q, m.p, x);
}
// The canonical residue in [0, p).
CF_HD double canon(double x, const Mod& m) {
    double r = reduce(x, m);
    if (r < 0) r += m.p;
    if (r >= m.p) r -= m.p;
    return r;
}
CF_HD double powmod(double b, uint64_t e, const Mod& m) {
    double r = 1.0;
    b = reduce(b, m);
    while (e) {
        if (e & 1) r = mulred(r, b, m);
        b = mulred(b, b, m);
        e >>= 1;
    }
    return reduce(r, m);
}

// ------------------------------------------------------------------------------------------
// Small DFTs in registers (natural order in and out). W[i] = w16^i for i < 8 (w16 a primitive
// 16th root of unity, W[0] unused). Radix-2 decimation in frequency; the bit-reversed result is
// renamed back to natural order at compile time. Inputs must satisfy |x| <= p/2 + 4: after four
// levels every entry is below 16 (p/2 + 4) < 2^53.
// ------------------------------------------------------------------------------------------
template <int H>
CF_HD void dif_level(double (&x)[16], int base, const double* W, const Mod& m) {
    // pairs (i, i + H) inside the block [base, base + 2H); twiddle w_{2H}^i = W[(8 / H) i]
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const double a = x[base + i], b = x[base + i + H];
        x[base + i] = a + b;
        const double d = a - b;
        x[base + i + H] = i ? mulred(d, W[(8 / H) * i], m) : d;
    }
}

CF_HD int bitrev4(int k) { return ((k & 1) << 3) | ((k & 2) << 1) | ((k & 4) >> 1) | ((k & 8) >> 3); }

// Radix-R DFT on the 16 / R consecutive groups of R entries (R = 2, 4, 8, 16).
template <int R>
CF_HD void dft_groups(double (&x)[16], const double* W, const Mod& m) {
#pragma unroll
    for (int g = 0; g < 16; g += R) {
        if (R >= 16) dif_level<8>(x, g, W, m);
        if (R >= 8)
#pragma unroll
            for (int b = 0; b < R; b += 8) dif_level<4>(x, g + b, W, m);
        if (R >= 4)
#pragma unroll
            for (int b = 0; b < R; b += 4) dif_level<2>(x, g + b, W, m);
#pragma unroll
        for (int b = 0; b < R; b += 2) dif_level<1>(x, g + b, W, m);
    }
    // the entry at offset j of a group holds frequency bitrev_R(j)
    double y[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) y[i] = x[i];
    constexpr int sh = R == 16 ? 0 : R == 8 ? 1 : R == 4 ? 2 : 3;
#pragma unroll
    for (int g = 0; g < 16; g += R)
#pragma unroll
        for (int k = 0; k < R; ++k) x[g + k] = y[g + (bitrev4(k) >> sh)];
}

CF_HD void reduce16(double (&x)[16], const Mod& m) {
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = reduce(x[i], m);
}

// x[k] *= tau^k for k = 1..15 with tau, tau^2, tau^4, tau^8 given (reduced): 11 products build the
// other powers, 15 apply them.
CF_HD void twiddle_chain(double (&x)[16], const double (&b)[4], const Mod& m) {
    const double t1 = b[0], t2 = b[1], t4 = b[2], t8 = b[3];
    const double t3 = mulred(t1, t2, m), t5 = mulred(t4, t1, m), t6 = mulred(t4, t2, m);
    const double t7 = mulred(t4, t3, m);
    const double t9 = mulred(t8, t1, m), t10 = mulred(t8, t2, m), t11 = mulred(t8, t3, m);
    const double t12 = mulred(t8, t4, m), t13 = mulred(t8, t5, m), t14 = mulred(t8, t6, m);
    const double t15 = mulred(t8, t7, m);
    x[1] = mulred(x[1], t1, m), x[2] = mulred(x[2], t2, m), x[3] = mulred(x[3], t3, m);
    x[4] = mulred(x[4], t4, m), x[5] = mulred(x[5], t5, m), x[6] = mulred(x[6], t6, m);
    x[7] = mulred(x[7], t7, m), x[8] = mulred(x[8], t8, m), x[9] = mulred(x[9], t9, m);
    x[10] = mulred(x[10], t10, m), x[11] = mulred(x[11], t11, m), x[12] = mulred(x[12], t12, m);
    x[13] = mulred(x[13], t13, m), x[14] = mulred(x[14], t14, m), x[15] = mulred(x[15], t15, m);
}

// LDS position with one pad word per 16: spreads the strided accesses of the later rounds over
// the banks.
CF_HD int lpos(int P) { return P + (P >> 4); }

// ------------------------------------------------------------------------------------------
// Per-task context: pointers into the task's slice of device memory and its constants.
// ------------------------------------------------------------------------------------------
struct TaskDesc {
    double p, pinv;
    double root, rooti;  // primitive L-th root of unity and its inverse (balanced)
    double linv;         // 1 / L
    double W[8], Wi[8];  // w16^i and w16^-i, w16 = root^(L/16)
    int64_t arena;       // offset (in doubles) of the task's memory
    int nm;              // largest n among the task's queries: table rows 0..nm
    int F1;              // factorial tables cover 0..F1-1
    int q_off, q_cnt;    // the task's queries (sorted by n, largest first)
    int slot_off;        // offset (in ints) of the segment table for this nm
    int NSEG;            // segments per row block
    int pool;            // tiles in each of the two pools
    int NB, NS;          // 16-row blocks of the tables; staging rows (16 NB)
    int NT;              // threads of the task's workgroup (L / 16)
    int lg;              // log2 L
    int dd;              // d, the number of candidate slots per item
    int model;           // 0: uniform d-subset, g(v) = C(v, d); 1: with replacement, g(v) = v^d
    int pad_;
    double wroot, wrooti;  // a task spread over workgroups: root of order 2^16 and its inverse
    double* keep;        // or null: resident table store of this prime (factorials, then the rows
                         // a_v[w] = V[v][w] / (w! v! (v-1-w)!) of v = v0+1 .. nm, packed by row)
};

// The resident store of one prime: fact, ifact, inv (F1 each), then row v (v entries) at
// 3 F1 + v (v - 1) / 2 for v = 1 .. nm (rows v <= v0 are zero and never written).
CF_HD int64_t keep_doubles(int F1, int nm) { return 3 * (int64_t)F1 + (int64_t)nm * (nm + 1) / 2; }
CF_HD int64_t keep_row(int F1, int v) { return 3 * (int64_t)F1 + (int64_t)v * (v - 1) / 2; }

struct QueryDesc {
    int n, M;
    int64_t out;  // index into the result array
};

// Layout of a task's arena (doubles), in this order.
struct Arena {
    double *Vpool, *Spool;  // pool tiles of 16 x 16 entries, row-major
    double *Vst, *Sst;      // staging: 2 blocks (column block parity) x 16 columns x NS rows, column-major
    double *fact, *ifact, *inv;
    double* ifact2;  // 1/i!^2: the stage outputs' factorial factors
    double* qtot;  // per query: Horner accumulator, then Q = g(M)
    double* hat;   // kernel spectrum times 1/L, in the final round's order (L entries)
};

CF_HD int64_t arena_doubles(int pool, int NS, int F1, int q_cnt, int L) {
    return (int64_t)pool * 512 + 64 * (int64_t)NS + 4 * (int64_t)F1 + 2 * (int64_t)q_cnt + L;
}

// A table of more than 16384 rows (16 workgroups per prime, transforms of L = 16 N2 = 2^16 or
// 2^17) holds after that arena its exchange rows (16 N2 doubles), the per-workgroup query sums
// (16 x the query slots, at least 8) and the hand-over entry (16).
CF_HD int wide_qslots(int q_cnt) { return q_cnt > 8 ? q_cnt : 8; }
CF_HD int64_t wide_extra_doubles(int q_cnt, int log2L) {
    return 16 * (int64_t(1) << (log2L - 4)) + 16 * (int64_t)wide_qslots(q_cnt) + 16;
}

CF_HD Arena arena_of(double* base, const TaskDesc& t) {
    Arena a;
    a.Vpool = base + t.arena;
    a.Spool = a.Vpool + (int64_t)t.pool * 256;
    a.Vst = a.Spool + (int64_t)t.pool * 256;
    a.Sst = a.Vst + 32 * (int64_t)t.NS;
    a.fact = a.Sst + 32 * (int64_t)t.NS;
    a.ifact = a.fact + t.F1;
    a.inv = a.ifact + t.F1;
    a.ifact2 = a.inv + t.F1;
    a.qtot = a.ifact2 + t.F1;
    a.hat = a.qtot + 2 * (int64_t)t.q_cnt;
    return a;
}

// A double at a byte offset from a base pointer that is the same for the whole workgroup: with a
// 32-bit offset (below 4 GiB) the per-entry address stays one register.
template <class O>
CF_HD double ldo(const double* base, O off) { return *(const double*)((const char*)base + off); }
CF_HD uint32_t ix8(int i) { return (uint32_t)i << 3; }

// Staging column of column v: block parity (v/16 mod 2), then column v%16 (in doubles).
CF_HD int64_t st_col(int v, int NS) { return (int64_t)((((v >> 4) & 1) << 4) + (v & 15)) * NS; }

// Byte offset of entry (r, c) (c < r) of a table from its pool base: tile (r/16, c/16) (its tile
// number given) when the column block has been flushed, else its staging column, the staging
// starting st_off bytes after the pool base. O is uint32_t while the pools and staging of a table
// span less than 4 GiB (tables of up to 16384 rows), uint64_t beyond.
template <class O>
CF_HD O tab_off(O st_off, int slot, int NS, int r, int c) {
    const int Y = r >> 4, X = c >> 4;
    return X < Y ? ((O)slot << 11) + ((O)(r & 15) << 7) + ((O)(c & 15) << 3)
                 : st_off + (((((O)(X & 1) << 4) + (O)(c & 15)) * (O)NS + (O)r) << 3);
}

// Address of entry (r, c) of a table with rows r and columns c < r: tile (r/16, c/16) of the pool
// (slot given) when the column block has been flushed (c/16 < r/16), else the staging column c%16.
// Branch-free, so a phase can issue all its loads before waiting for any.
CF_HD const double* tab_addr(const double* pool, const double* st, int slot, int NS, int r, int c) {
    const int Y = r >> 4, X = c >> 4;
    return X < Y ? pool + (int64_t)slot * 256 + ((r & 15) << 4) + (c & 15) : st + st_col(c, NS) + r;
}

// ------------------------------------------------------------------------------------------
// The NTT plan for L = 2^LOG2L with T = L/16 threads, 16 entries per thread.
// Round r < NTW is a radix-16 round on groups of stride S_r = L / 16^(r+1): DFT16 over the group
// then the twiddle w_{16 S_r}^(n2 k) (four-step decomposition, in place). The final round is a
// radix-RF DFT on consecutive entries (RF = 16 when LOG2L is a multiple of 4). The forward output
// is in digit-reversed order, which the pointwise product and the mirrored inverse never undo.
// ------------------------------------------------------------------------------------------
template <int LOG2L>
struct Plan {
    static constexpr int L = 1 << LOG2L;
    static constexpr int T = L >> 4;
    static constexpr int REM = LOG2L & 3;
    static constexpr int NTW = REM ? (LOG2L >> 2) : (LOG2L >> 2) - 1;  // twiddled radix-16 rounds
    static constexpr int RF = REM ? (1 << REM) : 16;                   // radix of the final round
    static constexpr int LDSW = L + (L >> 4);                          // padded buffer (doubles)
    static constexpr int lgS(int r) { return LOG2L - 4 * (r + 1); }
    // twiddle tables of rounds 1..NTW-1 (16 S_r entries each), forward then inverse
    static constexpr int tw_off(int r) {
        int o = 0;
        for (int k = 1; k < r; ++k) o += 16 << lgS(k);
        return o;
    }
    static constexpr int TWN = tw_off(NTW);  // entries per direction
    static constexpr int HI = (L >> 6) + 2;  // power-table entries g^(64 j), j <= nm/64
    static constexpr int NSEG = ((L + 1 + 15) / 16 + 15) / 16;  // segments per row block, nm <= L
    static constexpr int NBMAX = (L + 1 + 15) / 16;               // row blocks, nm <= L
};

// Shared (LDS) state of a workgroup running transforms of length 2^LOG2L on tables of up to
// 2^CAP rows (CAP > LOG2L for a task spread over several workgroups).
template <int LOG2L, int CAP = LOG2L>
struct Shared {
    using P = Plan<LOG2L>;
    using PC = Plan<CAP>;
    double buf[P::LDSW];
    double tw[2 * P::TWN + 2];
    double wsmall[2][16];  // w^b and w^-b for b < 16 (w the L-th root): the round-0 twiddles with tw
    double lo_pre[64], lo_post[64], hi[PC::HI];
    double red[P::T / 64 * 8 + 8];  // query partial sums: [wave][8]
    double scal[8];                 // broadcast scalars
    int segrow[PC::NSEG + 1];       // segment slots of the current row block
    unsigned short dseg[P::NBMAX];  // segments of the tiles whose flush is deferred (one block)
};

// Per-thread state.
struct Thread {
    double x[16];   // the thread's 16 transform entries
    double pa[16];  // data fetched one phase ahead of its use: a table row, the factorials of a
                    // stage output, or the kernel spectrum for the final round (never two at once)
};

template <int LOG2L>
struct Engine {
    using P = Plan<LOG2L>;
    static constexpr int L = P::L, T = P::T;

    // ---- forward round 0 on registers (entries n1 T + t), result stored to LDS ----
    // tau^(1, 2, 4, 8) for tau = w^t (dir 0) or w^-t (dir 1), t = 16 a + b: w^(16 a) is the
    // round-1 table's entry (k = 1, n2 = a), w^b comes from wsmall, and three squarings follow.
    template <class SH>
    static CF_HD void round0_bases(double (&b4)[4], const SH& sh, int t, int dir, const Mod& m) {
        constexpr int S1 = 1 << P::lgS(1);
        static_assert(T == 16 * S1, "t = 16 a + b with a < S1");
        const double* t1 = sh.tw + dir * P::TWN + P::tw_off(1);
        // unreduced products: each below 0.67p, which keeps twiddle_chain's products exact
        b4[0] = mulred(t1[S1 + (t >> 4)], sh.wsmall[dir][t & 15], m);
#pragma unroll
        for (int i = 1; i < 4; ++i) b4[i] = mulred(b4[i - 1], b4[i - 1], m);
    }
    template <class SH>
    static CF_HD void fwd0(Thread& th, SH& sh, int t, const TaskDesc& d, const Mod& m) {
        double b4[4];
        round0_bases(b4, sh, t, 0, m);
        reduce16(th.x, m);
        dft_groups<16>(th.x, d.W, m);
        twiddle_chain(th.x, b4, m);
#pragma unroll
        for (int k = 0; k < 16; ++k) sh.buf[lpos(k * T + t)] = th.x[k];
    }
    // ---- forward round r (1 <= r < NTW) ----
    // Every round reads its 16 entries (and its 15 twiddles) in a loop of their own before any
    // arithmetic uses them, so the loads are issued several at a time instead of one load, one
    // wait, one reduction after another.
    template <int r, class SH>
    static CF_HD void fwdr(Thread& th, SH& sh, int t, const TaskDesc& d, const Mod& m) {
        constexpr int lg = P::lgS(r), S = 1 << lg;
        const int b = t >> lg, n2 = t & (S - 1), base = (b << (lg + 4)) + n2;
        double(&x)[16] = th.x;
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = sh.buf[lpos(base + k * S)];
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = reduce(x[k], m);
        dft_groups<16>(x, d.W, m);
        const double* tw = sh.tw + P::tw_off(r);
        double wt[16];
#pragma unroll
        for (int k = 1; k < 16; ++k) wt[k] = tw[k * S + n2];
#pragma unroll
        for (int k = 1; k < 16; ++k) x[k] = mulred(x[k], wt[k], m);
#pragma unroll
        for (int k = 0; k < 16; ++k) sh.buf[lpos(base + k * S)] = x[k];
    }
    // ---- final round: forward radix-RF, then either keep (kernel spectrum) or multiply by the
    //      spectrum and run the inverse radix-RF ----
    // make_hat: store the spectrum (times 1/L) to hat; otherwise multiply by the spectrum entries
    // already fetched into th.pa.
    template <class SH>
    static CF_HD void final_round(bool make_hat, Thread& th, SH& sh, int t, const TaskDesc& d, const Mod& m,
                                  double* hat) {
        double(&x)[16] = th.x;
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = sh.buf[lpos(16 * t + i)];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = reduce(x[i], m);
        dft_groups<P::RF>(x, d.W, m);
        if (make_hat) {
            double* h = hat + 16 * t;
#pragma unroll
            for (int i = 0; i < 16; ++i) h[i] = reduce(mulred(x[i], d.linv, m), m);
            return;
        }
        // with a radix-2 or radix-4 final round the products (below p: the spectrum entries are
        // below 2p + 16) enter the inverse DFT unreduced, and two levels keep them below 4p
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            x[i] = mulred(x[i], th.pa[i], m);
            if (P::RF > 4) x[i] = reduce(x[i], m);
        }
        dft_groups<P::RF>(x, d.Wi, m);
#pragma unroll
        for (int i = 0; i < 16; ++i) sh.buf[lpos(16 * t + i)] = x[i];
    }
    // ---- inverse round r (1 <= r < NTW): twiddle w^-(n2 k), then inverse DFT16 ----
    template <int r, class SH>
    static CF_HD void invr(Thread& th, SH& sh, int t, const TaskDesc& d, const Mod& m) {
        constexpr int lg = P::lgS(r), S = 1 << lg;
        const int b = t >> lg, n2 = t & (S - 1), base = (b << (lg + 4)) + n2;
        double(&x)[16] = th.x;
        const double* tw = sh.tw + P::TWN + P::tw_off(r);
double wt[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = sh.buf[lpos(base + k * S)];
#pragma unroll
        for (int k = 1; k < 16; ++k) wt[k] = tw[k * S + n2];
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = reduce(k ? mulred(x[k], wt[k], m) : x[k], m);
        dft_groups<16>(x, d.Wi, m);
#pragma unroll
        for (int k = 0; k < 16; ++k) sh.buf[lpos(base + k * S)] = x[k];
    }
    // ---- inverse round 0: LDS -> registers (entries n1 T + t, natural order) ----
    template <class SH>
    static CF_HD void inv0(Thread& th, SH& sh, int t, const TaskDesc& d, const Mod& m) {
        double(&x)[16] = th.x;
        double b4[4];
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = sh.buf[lpos(k * T + t)];
        round0_bases(b4, sh, t, 1, m);
        twiddle_chain(x, b4, m);
        reduce16(x, m);
        dft_groups<16>(x, d.Wi, m);
    }

    // ---- prologue pieces ----
    // w^(+-b) for b < 16 and the LDS tables of rounds >= 1 (entries idx = t, t + T, ...)
    template <class SH>
    static CF_HD void twiddles(SH& sh, int t, const TaskDesc& d, const Mod& m) {
        if (t < 16) {
            sh.wsmall[0][t] = powmod(d.root, (uint64_t)t, m);
            sh.wsmall[1][t] = powmod(d.rooti, (uint64_t)t, m);
        }
        for (int r = 1; r < P::NTW; ++r) {
            const int lg = P::lgS(r), S = 1 << lg;
            for (int idx = t; idx < 16 * S; idx += T) {
                const int k = idx >> lg, n2 = idx & (S - 1);
                const uint64_t e = ((uint64_t)n2 * (uint64_t)k << (4 * r)) & (L - 1);
                sh.tw[P::tw_off(r) + idx] = powmod(d.root, e, m);
                sh.tw[P::TWN + P::tw_off(r) + idx] = powmod(d.rooti, e, m);
            }
        }
    }
};

}  // namespace cf
