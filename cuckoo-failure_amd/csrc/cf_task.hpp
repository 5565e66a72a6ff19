// cf_task.hpp -- one prime ("task") from start to finish: prologue tables, the v-loop of the
// suffix-table recurrence with streaming queries, and the canonical results.
//
// The whole task is one sequence of phases; within a phase every thread works on its own
// entries, and consecutive phases are separated by a workgroup barrier. The sequence is written
// once (run_task) over a Runner: the HIP kernel's runner executes a phase for its own thread and
// then waits at the barrier, the CPU emulation's runner executes the phase for every thread in
// turn.
//
// Reference correspondence: the v-loop is fast/cuckoo_fast.cu:841-855 (stage 1 for v > d,
// stage 2 for v >= d); Stage1Pre/Post, PowTab, Stage2Pre/Post are :146-224 (here in the rescaled
// form described in cf_core.hpp); the query reduction is :306-323 with the host Horner finish of
// :877-890 done on the device, row by row as the rows become final.
#pragma once
#include <type_traits>

#include "cf_core.hpp"

namespace cf {

// Everything a phase needs about the task.
struct Ctx {
    const TaskDesc* d;  // in device memory: fields are read when needed
    Mod m;
    Arena a;
    const int* slots;
    const QueryDesc* q;
    int model;  // 0: g(v) = C(v, dd); 1: g(v) = v^dd
    int dd;     // d, the number of candidate slots per item
    int v0;     // first v with g(v) != 0
    // Element map: register k of thread t holds transform entry k es + eb + t (a one-workgroup
    // task: es = T, eb = 0; workgroup g of a wide task: es = 4096, eb = 256 g); L the task's
    // transform length.
    int es, eb, L;
};

// DEFER: -1 the built-in choice for the length (deferred tile flush where the transform has the
// carrier rounds), 0 never; a task-kernel variant the autotuner can pick.
template <int LOG2L, int CAP = LOG2L, int DEFER = -1>
struct Task {
    using P = Plan<LOG2L>;
    using E = Engine<LOG2L>;
    using Sh = Shared<LOG2L, CAP>;
    static constexpr int L = P::L, T = P::T, NW = T / 64;  // NW waves

    // ---------------- prologue ----------------
    // Factorials by a blocked prefix product: each thread owns a contiguous chunk.
    static CF_HD int chunk(const Ctx& c) { return (c.d->F1 + T - 1) / T; }

    static CF_HD void fact_local(Sh& sh, int t, const Ctx& c) {
        const int ch = chunk(c), lo = t * ch, hi = lo + ch < c.d->F1 ? lo + ch : c.d->F1;
        double run = 1.0;
        for (int i = lo; i < hi; ++i) {
            if (i > 0) run = mulred(run, (double)i, c.m);
            c.a.fact[i] = run;
        }
        sh.buf[t] = reduce(run, c.m);
    }
    // Inclusive scan (products) of buf[0..T) in log2 T rounds, ping-ponging between the halves.
    static CF_HD void scan_step(Sh& sh, int t, int s, int src, const Ctx& c) {
        const double* in = sh.buf + src * T;
        double* out = sh.buf + (1 - src) * T;
        out[t] = t >= s ? reduce(mulred(in[t - s], in[t], c.m), c.m) : in[t];
    }
    static CF_HD void fact_fix(Sh& sh, int t, int src, const Ctx& c) {
        const int ch = chunk(c), lo = t * ch, hi = lo + ch < c.d->F1 ? lo + ch : c.d->F1;
        const double pre = t ? sh.buf[src * T + t - 1] : 1.0;
        for (int i = lo; i < hi; ++i) c.a.fact[i] = reduce(mulred(c.a.fact[i], pre, c.m), c.m);
    }
    // ifact[i] = (F1-1)!^-1 * prod_{j=i+1}^{F1-1} j, by the mirrored blocked scan.
    static CF_HD void ifact_local(Sh& sh, int t, const Ctx& c) {
        const int ch = chunk(c), lo = t * ch, hi = lo + ch < c.d->F1 ? lo + ch : c.d->F1;
        double run = 1.0;
        for (int i = hi - 1; i >= lo; --i) {
            c.a.ifact[i] = run;  // prod_{j=i+1}^{hi-1} j
            if (i > 0) run = mulred(run, (double)i, c.m);
        }
        // scanned from the right: slot T-1-t holds this chunk's product
        sh.buf[T - 1 - t] = reduce(run, c.m);
    }
    static CF_HD void ifact_fix(Sh& sh, int t, int src, const Ctx& c) {
        const int ch = chunk(c), lo = t * ch, hi = lo + ch < c.d->F1 ? lo + ch : c.d->F1;
        // product of the chunks to the right of this one
        const double suf = t < T - 1 ? sh.buf[src * T + T - 2 - t] : 1.0;
        const double f = mulred(suf, sh.scal[0], c.m);
        for (int i = lo; i < hi; ++i) c.a.ifact[i] = reduce(mulred(c.a.ifact[i], f, c.m), c.m);
    }
    static CF_HD void inv_table(int t, const Ctx& c) {
        for (int i = c.eb + t; i < c.d->F1; i += c.es) {
            c.a.inv[i] = i ? reduce(mulred(c.a.ifact[i], c.a.fact[i - 1], c.m), c.m) : 0.0;
            c.a.ifact2[i] = reduce(mulred(c.a.ifact[i], c.a.ifact[i], c.m), c.m);
        }
    }
    static CF_HD void zero_staging(int t, const Ctx& c) {
        const int64_t n = 32 * (int64_t)c.d->NS;
        for (int64_t i = c.eb + t; i < n; i += c.es) c.a.Vst[i] = 0.0, c.a.Sst[i] = 0.0;
    }
    static CF_HD double g_of(int v, const Ctx& c) {
        if (c.model) return powmod((double)v, (uint64_t)c.dd, c.m);
        if (v < c.dd) return 0.0;
        return reduce(mulred(mulred(c.a.fact[v], c.a.ifact[c.dd], c.m), c.a.ifact[v - c.dd], c.m), c.m);
    }
    static CF_HD void query_init(int t, const Ctx& c) {
        for (int j = t; j < c.d->q_cnt; j += T) {
            c.a.qtot[2 * j] = 0.0;
            c.a.qtot[2 * j + 1] = g_of(c.q[c.d->q_off + j].M, c);
        }
    }
    // the convolution kernel 1/t (1 <= t < nm) in the round-0 layout
    static CF_HD void kernel_input(Thread& th, int t, const Ctx& c) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * c.es + c.eb + t;
            th.x[k] = (e >= 1 && e < c.d->nm) ? c.a.inv[e] : 0.0;
        }
    }

    // ---------------- per step ----------------
    // Every stage function below is branch-free over its 16 entries: indices are clamped to a
    // valid entry and the result is selected afterwards, so all loads of a phase are in flight
    // together instead of one round trip per entry.

    // Tile number of (Y = r/16, X) from the current row block's segment row (X < Y).
    static CF_HD int tile_of(const Sh& sh, int X) { return sh.segrow[X >> 4] * 16 + (X & 15); }
    // Segment row of row block Y into LDS.
    static CF_HD void load_segrow(Sh& sh, int t, int Y, const Ctx& c) {
        for (int s = t; s < c.d->NSEG; s += T) sh.segrow[s] = c.slots[Y * c.d->NSEG + s];
    }
    // Entries (r, e) of a table for this thread's e in [lo, hi), as raw loads into dst (other
    // entries get the value at a clamped, valid address and must be masked by the consumer).
    // Row block r/16 must be the one whose segment row is in LDS.
    // byte offsets within a table: 32-bit while it spans less than 4 GiB (up to 2^14 rows)
    using Off = typename std::conditional<(CAP > 14), uint64_t, uint32_t>::type;
    static CF_HD void fetch_row(double (&dst)[16], const double* pool, const double* st, const Sh& sh, int t, int r, int lo,
                                int hi, const Ctx& c) {
        const int Y = r >> 4;
        const Off st_off = (Off)((st - pool) << 3);
        const int NS = c.d->NS;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * c.es + c.eb + t, ce = (e >= lo && e < hi) ? e : 0, X = ce >> 4;
            const int tile = X < Y ? tile_of(sh, X) : 0;
            dst[k] = ldo(pool, tab_off<Off>(st_off, tile, NS, r, ce));
        }
    }
    static CF_HD void fetch_hat(Thread& th, int t, const Ctx& c) {
#pragma unroll
        for (int i = 0; i < 16; ++i) th.pa[i] = c.a.hat[16 * t + i];
    }
    // 1/(e - v - off)!^2 for the stage outputs: off = 0 for stage 1 (rows v <= e < nm), off = 1 for
    // stage 2 (rows v < p <= nm, entry 0 standing for p = L).
    static CF_HD void fetch_ifact(Thread& th, int t, int v, int off, const Ctx& c) {
        const int nm = c.d->nm;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * c.es + c.eb + t, p = (off && e == 0) ? c.L : e;
            const bool ok = off ? (p > v && p <= nm) : (e >= v && e < nm);
            th.pa[k] = ldo(c.a.ifact2, ix8(ok ? p - v - off : 0));
        }
    }
    // Stage-1 input a[w] = P2[v][w] / v = V2[v][w] for w < v: columns w < v-1 from th.pa (fetched
    // ahead, or now when fetch_now), column v-1 from the LDS hand-over of the previous step.
    static CF_HD void stage1_pre(Thread& th, const Sh& sh, int t, int v, bool fetch_now, const Ctx& c) {
        if (fetch_now) fetch_row(th.pa, c.a.Vpool, c.a.Vst, sh, t, v, 0, v - 1, c);
        const double iv = c.a.inv[v], hand = sh.scal[1];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * c.es + c.eb + t;
            const double x = e < v - 1 ? th.pa[k] : hand;
            const double a = mulred(x, iv, c.m);
            th.x[k] = e < v ? a : 0.0;
        }
    }
    // Partial sums of the queries j0 .. j0+cnt-1 over this thread's entries of row v.
    static CF_HD void query_partial(const Thread& th, int t, int v, int j0, int cnt, double (&acc)[8], const Ctx& c) {
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            double s = 0.0;
            if (jj < cnt) {
                const int M = c.q[c.d->q_off + j0 + jj].M;
                double iv[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const int e = k * c.es + c.eb + t;
                    iv[k] = ldo(c.a.inv, ix8(M - (e < v ? e : 0)));
                }
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    s += mulred(th.x[k], iv[k], c.m);  // th.x[k] = 0 for e >= v
                    if (k == 7) s = reduce(s, c.m);
                }
            }
            acc[jj] = reduce(s, c.m);
        }
    }
    // Horner update of query j0 + jj with the block-reduced row sum s.
    static CF_HD void query_update(int jj, int v, int j0, double s, const Ctx& c) {
        const int j = j0 + jj;
        const QueryDesc& q = c.q[c.d->q_off + j];
        const double coef = mulred(mulred(c.a.fact[q.n], c.a.ifact[q.n - v], c.m), c.a.ifact[q.M - v], c.m);
        const double tot = c.a.qtot[2 * j], Q = c.a.qtot[2 * j + 1];
        c.a.qtot[2 * j] = reduce(mulred(tot, Q, c.m) + mulred(coef, s, c.m), c.m);
    }
    // Thread t < cnt finishes query j0 + t from the per-wave partial sums.
    static CF_HD void query_finish(const Sh& sh, int t, int v, int j0, int cnt, const Ctx& c) {
        if (t < cnt) {
            double s = 0.0;
            for (int w = 0; w < NW; ++w) s += sh.red[w * 8 + t];  // NW <= 16 values below p/2 + 4
            query_update(t, v, j0, reduce(s, c.m), c);
        }
    }
    // Stage-1 output S2[e][v] = c[e] / (e-v)!^2 for v <= e < nm into the staging column v % 16,
    // with 1/(e-v)!^2 in th.pa (fetched ahead). Then row v of S2 is fetched into th.pa for the
    // stage-2 input of the next phase; its entry (v, v) is the value this thread just computed.
    static CF_HD void stage1_post(Thread& th, const Sh& sh, int t, int v, const Ctx& c) {
        double* col = c.a.Sst + st_col(v, c.d->NS);
        const int nm = c.d->nm;
        double y[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * c.es + c.eb + t;
            y[k] = mulred(th.x[k], th.pa[k], c.m);
            if (e >= v && e < nm) col[e] = y[k];
        }
        fetch_row(th.pa, c.a.Spool, c.a.Sst, sh, t, v, 1, v, c);
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (k * c.es + c.eb + t == v) th.pa[k] = y[k];
    }
    // Column v of a staging area set to zero on rows lo..hi-1.
    static CF_HD void zero_col(double* st, int t, int v, int lo, int hi, const Ctx& c) {
        double* col = st + st_col(v, c.d->NS);
        for (int r = lo + c.eb + t; r < hi; r += c.es) col[r] = 0.0;
    }
    // Power tables of g for step v: lo_pre[i] = v! g^i, lo_post[i] = g^i / v!, hi[j] = g^(64 j).
    static CF_HD void powtab(Sh& sh, int t, int v, double g, const Ctx& c) {
        const int nhi = (c.d->nm >> 6) + 1;
        for (int idx = t; idx < 128 + nhi; idx += T) {
            if (idx < 128) {
                const double x = powmod(g, (uint64_t)(idx & 63), c.m);
                if (idx < 64) sh.lo_pre[idx] = reduce(mulred(x, c.a.fact[v], c.m), c.m);
                else sh.lo_post[idx - 64] = reduce(mulred(x, c.a.ifact[v], c.m), c.m);
            } else {
                const double g64 = powmod(g, 64, c.m);
                sh.hi[idx - 128] = powmod(g64, (uint64_t)(idx - 128), c.m);
            }
        }
    }
    static CF_HD double gpow(const double* lo, const Sh& sh, int e, const Mod& m) { return mulred(sh.hi[e >> 6], lo[e & 63], m); }

    // Stage-2 input: b[q] = S2[v][q] * v! g^(v-q) for 1 <= q <= v (row v of S2 in th.pa).
    static CF_HD void stage2_pre(Thread& th, const Sh& sh, int t, int v, const Ctx& c) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * c.es + c.eb + t, ok = e >= 1 && e <= v, ce = ok ? e : v;
            th.x[k] = ok ? mulred(th.pa[k], gpow(sh.lo_pre, sh, v - ce, c.m), c.m) : 0.0;
        }
    }
    // Stage-2 output P2[p][v] = (g^v/v! - p c[p]) g^(p-v)/v! / (p-v-1)!^2 for v < p <= nm into
    // the staging column v % 16 (c = b (*) k in th.x, 1/(p-v-1)!^2 in th.pa). The entry of row
    // v+1 also goes to the LDS hand-over for the next step's stage-1 input.
    static CF_HD void stage2_post(const Thread& th, Sh& sh, int t, int v, bool conv, const Ctx& c) {
        double* col = c.a.Vst + st_col(v, c.d->NS);
        const int nm = c.d->nm;
        const double s1 = gpow(sh.lo_post, sh, v, c.m);  // g^v / v!
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * c.es + c.eb + t, p = e ? e : c.L;  // c[p mod L]: entry 0 stands for p = L
            const bool ok = p > v && p <= nm;
            const int j = ok ? p - v : 1;
            const double fj = mulred(gpow(sh.lo_post, sh, j, c.m), th.pa[k], c.m);
            // p c[p]: c below 8p + 64 (the inverse transform's last DFT16), p <= 2^16
            const double u = s1 - (conv ? mulred(th.x[k], (double)p, c.m) : 0.0);
            const double y = mulred(u, fj, c.m);
            if (ok) col[p] = y;
            if (p == v + 1) sh.scal[1] = y;
        }
    }
    // Staging column block X -> pool tiles (Y, X) for every row block Y > X, in two phases: the
    // tile numbers go to LDS (the transform buffer is free here), then every thread moves 16
    // entries per trip with all 16 loads in flight.
    static CF_HD void flush_tiles(Sh& sh, int t, int X, const Ctx& c) {
        int* tl = (int*)sh.buf;
        const int NSEG = c.d->NSEG;
        for (int j = t; j < c.d->NB - 1 - X; j += T) tl[j] = c.slots[(X + 1 + j) * NSEG + (X >> 4)] * 16 + (X & 15);
    }
    // Both tables in each trip (their staging and pools share one layout and tile numbering), so
    // every thread has 32 loads in flight before it stores.
    static CF_HD void flush(const Sh& sh, int t, int X, const Ctx& c) {
        const int* tl = (const int*)sh.buf;
        const int n = (c.d->NB - 1 - X) * 256;
        const uint32_t NS = (uint32_t)c.d->NS, row0 = 16 * (uint32_t)(X + 1), cb = (uint32_t)(X & 1) << 4;
        const int gt = c.eb + t, GT = c.es;  // this thread among all threads of the task
        for (int base = 0; base < n; base += 16 * GT) {
            double yv[16], ys[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int idx = base + u * GT + gt, ci = idx < n ? idx : 0;
                const uint32_t j = (uint32_t)ci >> 8, col = ((uint32_t)ci >> 4) & 15, r = (uint32_t)ci & 15;
                const uint32_t off = ((cb + col) * NS + row0 + 16 * j + r) << 3;
                yv[u] = ldo(c.a.Vst, off);
                ys[u] = ldo(c.a.Sst, off);
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int idx = base + u * GT + gt;
                if (idx < n) {
                    const int col = (idx >> 4) & 15, r = idx & 15;
                    const int64_t at = (int64_t)tl[idx >> 8] * 256 + (r << 4) + col;
                    c.a.Vpool[at] = yv[u];
                    c.a.Spool[at] = ys[u];
                }
            }
        }
    }

    // ---- deferred flush (transforms with at least three radix-16 rounds, one workgroup) ----
    // At the end of column block X only tile (X+1, X), which the next step reads, is written
    // (flush_head). The tiles (Y, X), Y >= X+2, are first read in block X+2: they are moved during
    // the 16 steps of block X+1, in 64 slices, by the first forward and the last inverse round of
    // each convolution, whose th.pa is free (loads issued one phase earlier, after the stage
    // input or the pointwise product has used th.pa; stores at the end of the round);
    // the staging is double-buffered by block parity, so block X+1 writes the other half.
    static constexpr bool kDefer = DEFER != 0 && P::NTW >= 3 && CAP == LOG2L;
    static CF_HD void flush_head(int t, int X, const Ctx& c) {
        if (X + 1 >= c.d->NB) return;
        const int NS = c.d->NS, cb = (X & 1) << 4;
        const int64_t tile = (int64_t)c.slots[(X + 1) * c.d->NSEG + (X >> 4)] * 16 + (X & 15);
        for (int gi = t; gi < 512; gi += T) {
            const int tbl = gi >> 8, col = (gi >> 4) & 15, r = gi & 15;
            const double* st = tbl ? c.a.Sst : c.a.Vst;
            double* pool = tbl ? c.a.Spool : c.a.Vpool;
            pool[tile * 256 + r * 16 + col] = st[(int64_t)(cb + col) * NS + 16 * (X + 1) + r];
        }
    }
    // segment numbers of the deferred tiles (X+2+j, X) into LDS (at the start of block X+1)
    static CF_HD void defer_list(Sh& sh, int t, int X, const Ctx& c) {
        const int n = c.d->NB - 2 - X, NSEG = c.d->NSEG;
        for (int j = t; j < n; j += T) sh.dseg[j] = (unsigned short)c.slots[(X + 2 + j) * NSEG + (X >> 4)];
    }
    // Slice k (of 64) of block X's deferred tiles. The tiles are numbered q = 0 .. 2 (NB-2-X) - 1:
    // table q & 1 (0: P2, 1: S2), row block X+2 + q/2. Slice k holds the TPS = ceil(tiles / 64)
    // tiles from k TPS on; virtual thread t moves entry (row t & 15, column (t >> 4) & 15) of its
    // tiles q = k TPS + t/256 + u T/256 (u < 8, while inside the slice). T/256 is even, so a thread
    // stays in one table and its tiles lie T/512 row blocks apart: its staging loads differ by
    // constant offsets, and each store needs one segment lookup (dseg) for its tile.
    static constexpr int kDeferStep = T / 256;  // tiles per pass of the T virtual threads
    static_assert(!kDefer || (kDeferStep % 2 == 0 && 8 * kDeferStep * 64 >= 2 * (P::NBMAX - 2)),
                  "eight passes of T/256 tiles cover a slice, one table per thread");
    // tiles of slice k (0 if none), and this thread's first tile
    static CF_HD int defer_count(int t, int X, int k, const Ctx& c, int& q0) {
        const int ntile = 2 * (c.d->NB - 2 - X);
        if (ntile <= 0) return 0;
        const int tps = (ntile + 63) >> 6, lo = k * tps;
        q0 = lo + (t >> 8);
        return ntile - lo < tps ? ntile - lo : tps;
    }
    // Loads of every pass, from the P2 staging base (the S2 staging follows it at 32 NS doubles).
    // A pass beyond the slice loads the first pass's entry again (a cache hit, no HBM traffic); that
    // entry itself, when this thread has no tile in the slice, is a staging entry of a later row or
    // lies past the S2 staging's end at most 16 (8 T/512) doubles into the factorial tables that
    // follow it (4 F1 + L >= 8192 more doubles of the same arena); such values are never stored.
    static CF_HD void defer_load(Thread& th, int t, int X, int k, const Ctx& c) {
        int q0 = 0;
        const int cnt = defer_count(t, X, k, c, q0), tq = t >> 8;
        if (cnt <= 0) return;
        const uint32_t NS = (uint32_t)c.d->NS, cb = (uint32_t)(X & 1) << 4, col = (t >> 4) & 15, r = t & 15;
        // staging entry (row 16 (X+2+j) + r, column cb + col) of table q0 & 1, j = q0/2 + u T/512
        const uint32_t off0 = (((q0 & 1) ? 32 * NS : 0u) + (cb + col) * NS +
                               16 * ((uint32_t)X + 2 + (uint32_t)(q0 >> 1)) + r) << 3;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t du = tq + kDeferStep * u < cnt ? (uint32_t)(u * (kDeferStep / 2) * 16 * 8) : 0u;
            th.pa[u] = ldo(c.a.Vst, off0 + du);
        }
    }
    static CF_HD void defer_store(const Thread& th, const Sh& sh, int t, int X, int k, const Ctx& c) {
        int q0 = 0;
        const int cnt = defer_count(t, X, k, c, q0), tq = t >> 8;
        if (cnt <= 0) return;
        // from the P2 pool base (the S2 pool follows it at pool tiles of 2048 bytes): tile
        // dseg[j] 16 + X % 16, entry r 16 + col inside it
        const Off base = ((q0 & 1) ? (Off)c.d->pool << 11 : (Off)0) +
                         ((Off)(((X & 15) << 8) + ((t & 15) << 4) + ((t >> 4) & 15)) << 3);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (tq + kDeferStep * u < cnt) {
                const int j = (q0 >> 1) + u * (kDeferStep / 2);
                *(double*)((char*)c.a.Vpool + (base + ((Off)sh.dseg[j] << 15))) = th.pa[u];
            }
        }
    }
    // slices k0 .. k1-1 in a phase of their own (a step of block X+1 without a convolution)
    template <class R>
    static CF_HD void defer_phase(R& run, Sh& sh, int X, int k0, int k1, const Ctx& c) {
        run.phase([&](Thread& th, int t) {
            for (int k = k0; k < k1; ++k) {
                defer_load(th, t, X, k, c);
                defer_store(th, sh, t, X, k, c);
            }
        });
    }

    // ---------------- the convolution, split at its first and last phase ----------------
    // The caller fuses its own work into the first phase (before fwd0) and the last (after inv0).
    // The middle phases also fetch ahead: the kernel spectrum for the final round, and the
    // factorials 1/(e - v - off)! that the caller's last phase needs.
    // With kDefer and X >= 0, the first forward and the last inverse round also carry the
    // deferred-flush slices k and k + 1 of block X.
    template <class R>
    static CF_HD void conv_middle(R& run, Sh& sh, int v, int off, const Ctx& c, int X = -1, int k = 0) {
        const bool carry = kDefer && X >= 0;
        // slice k was loaded at the end of the caller's first phase (after its stage input)
        static_for<1, P::NTW>([&](auto RI) {
            constexpr int r = decltype(RI)::value;
            run.phase([&](Thread& th, int t) {
                if (r == P::NTW - 1) fetch_hat(th, t, c);
                E::template fwdr<r>(th, sh, t, *c.d, c.m);
                if (carry && r == 1) defer_store(th, sh, t, X, k, c);
            });
        });
        run.phase([&](Thread& th, int t) {
            E::final_round(false, th, sh, t, *c.d, c.m, c.a.hat);
            if (carry) defer_load(th, t, X, k + 1, c);
        });
        static_rfor<1, P::NTW>([&](auto RI) {
            constexpr int r = decltype(RI)::value;
            run.phase([&](Thread& th, int t) {
                if (r == 1) fetch_ifact(th, t, v, off, c);
                E::template invr<r>(th, sh, t, *c.d, c.m);
                if (carry && r == P::NTW - 1) defer_store(th, sh, t, X, k + 1, c);
            });
        });
    }

    static CF_HD int active_queries(int v, const Ctx& c) {
        // queries are sorted by n, largest first: the active ones (n >= v) are a prefix
        int nact = 0;
        while (nact < c.d->q_cnt && c.q[c.d->q_off + nact].n >= v) ++nact;
        return nact;
    }

    // Row v of Vh is in th.x (stage-1 input form): add its query terms, in phases of their own.
    template <class R>
    static CF_HD void queries(R& run, Sh& sh, int v, int j_from, const Ctx& c) {
        const int nact = active_queries(v, c);
        for (int j0 = j_from; j0 < nact; j0 += 8) {
            const int cnt = nact - j0 < 8 ? nact - j0 : 8;
            run.wave_reduce8(sh.red, [&](Thread& th, int t, double (&acc)[8]) {
                query_partial(th, t, v, j0, cnt, acc, c);
                return cnt;
            }, c.m);
            run.phase([&](Thread&, int t) { query_finish(sh, t, v, j0, cnt, c); });
        }
    }

    // Retain mode: row v of the stage-1 input (a[w] for w < v, in th.x) into the resident store.
    static CF_HD void keep_row_out(const Thread& th, int t, int v, const Ctx& c) {
        double* keep = c.d->keep;
        if (!keep) return;
        double* row = keep + keep_row(c.d->F1, v);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * c.es + c.eb + t;
            if (e < v) row[e] = th.x[k];
        }
    }
    static CF_HD void keep_factorials(int t, const Ctx& c) {
        double* keep = c.d->keep;
        const int F1 = c.d->F1;
        for (int i = c.eb + t; i < F1; i += c.es) {
            keep[i] = c.a.fact[i];
            keep[F1 + i] = c.a.ifact[i];
            keep[2 * F1 + i] = c.a.inv[i];
        }
    }

    // Fetch ahead for step v + 1: its row of P2 except column v (which the hand-over carries).
    static CF_HD void fetch_next(Thread& th, const Sh& sh, int t, int v, const Ctx& c) {
        fetch_row(th.pa, c.a.Vpool, c.a.Vst, sh, t, v + 1, 0, v, c);
    }

    template <class R>
    static CF_HD void run_task(R& run, Sh& sh, const Ctx& c) {
        const int nm = c.d->nm;
        // ---- prologue ----
        run.phase([&](Thread&, int t) { zero_staging(t, c); fact_local(sh, t, c); });
        int src = 0;
        for (int s = 1; s < T; s <<= 1, src ^= 1) run.phase([&](Thread&, int t) { scan_step(sh, t, s, src, c); });
        run.phase([&](Thread&, int t) { fact_fix(sh, t, src, c); });
        run.phase([&](Thread&, int t) {
            if (t == 0) sh.scal[0] = powmod(c.a.fact[c.d->F1 - 1], (uint64_t)c.d->p - 2, c.m);
        });
        run.phase([&](Thread&, int t) { ifact_local(sh, t, c); });
        src = 0;
        for (int s = 1; s < T; s <<= 1, src ^= 1) run.phase([&](Thread&, int t) { scan_step(sh, t, s, src, c); });
        run.phase([&](Thread&, int t) { ifact_fix(sh, t, src, c); });
        run.phase([&](Thread&, int t) {
            inv_table(t, c);
            E::twiddles(sh, t, *c.d, c.m);
        });
        // kernel spectrum (times 1/L) into the arena
        run.phase([&](Thread& th, int t) {
            query_init(t, c);
            kernel_input(th, t, c);
            E::fwd0(th, sh, t, *c.d, c.m);
        });
        static_for<1, P::NTW>([&](auto RI) {
            run.phase([&](Thread& th, int t) { E::template fwdr<decltype(RI)::value>(th, sh, t, *c.d, c.m); });
        });
        run.phase([&](Thread& th, int t) { E::final_round(true, th, sh, t, *c.d, c.m, c.a.hat); });

        // ---- the v-loop ----
        int segY = -1;       // row block whose segment row is in LDS
        bool pref = false;  // th.pa holds row v of P2 (columns < v-1), fetched by the previous step
        for (int v = 1; v < nm; ++v) {
            const bool s1 = v > c.v0;   // rows v <= v0 of V are zero, so column v of S is zero
            const bool s2 = v >= c.v0;  // g(v) = 0 below v0, so column v of V is zero
            const double g = s2 ? g_of(v, c) : 0.0;
            const int Xd = kDefer ? (v >> 4) - 1 : -1;  // block whose deferred tiles move this step
            const int kd = 4 * (v & 15);                 // its slices kd .. kd+3 belong to this step
            if ((v >> 4) != segY && (s1 || s2 || Xd >= 0)) {
                segY = v >> 4;
                run.phase([&](Thread&, int t) {
                    load_segrow(sh, t, segY, c);
                    if (Xd >= 0) defer_list(sh, t, Xd, c);
                });
            }
            if (s1) {
                const int nact = active_queries(v, c);
                const int cnt0 = nact < 8 ? nact : 8;
                const bool now = !pref;
                if (nact <= 8) {
                    // stage-1 input, its query terms and the first forward round in one phase
                    run.phase([&](Thread& th, int t) {
                        stage1_pre(th, sh, t, v, now, c);
                        keep_row_out(th, t, v, c);
                        if (cnt0) run.wave_partial8(sh.red, th, t, [&](Thread& th2, int t2, double (&acc)[8]) {
                            query_partial(th2, t2, v, 0, cnt0, acc, c);
                            return cnt0;
                        }, c.m);
                        E::fwd0(th, sh, t, *c.d, c.m);
                        if (kDefer && Xd >= 0) defer_load(th, t, Xd, kd, c);
                    });
                } else {
                    run.phase([&](Thread& th, int t) {
                        stage1_pre(th, sh, t, v, now, c);
                        keep_row_out(th, t, v, c);
                    });
                    queries(run, sh, v, 0, c);
                    run.phase([&](Thread& th, int t) {
                        E::fwd0(th, sh, t, *c.d, c.m);
                        if (kDefer && Xd >= 0) defer_load(th, t, Xd, kd, c);
                    });
                }
                conv_middle(run, sh, v, 0, c, Xd, kd);
                // last inverse round, stage-1 output (then row v of S2 is fetched), query finish,
                // power tables of stage 2
                run.phase([&](Thread& th, int t) {
                    if (nact <= 8 && cnt0) query_finish(sh, t, v, 0, cnt0, c);
                    E::inv0(th, sh, t, *c.d, c.m);
                    stage1_post(th, sh, t, v, c);
                    if (s2) powtab(sh, t, v, g, c);
                });
            } else {
                run.phase([&](Thread&, int t) {
                    zero_col(c.a.Sst, t, v, v, nm, c);
                    if (s2) powtab(sh, t, v, g, c);
                });
                if (Xd >= 0) defer_phase(run, sh, Xd, kd, kd + 2, c);
            }
            // fetch ahead for step v + 1 unless a flush (and a new row block) comes in between
            const bool next = s2 && (v & 15) != 15;
            if (s2) {
                const bool conv = v > c.v0;  // row v of S is zero unless v > v0
                if (conv) {
                    run.phase([&](Thread& th, int t) {
                        stage2_pre(th, sh, t, v, c);
                        E::fwd0(th, sh, t, *c.d, c.m);
                        if (kDefer && Xd >= 0) defer_load(th, t, Xd, kd + 2, c);
                    });
                    conv_middle(run, sh, v, 1, c, Xd, kd + 2);
                    run.phase([&](Thread& th, int t) {
                        E::inv0(th, sh, t, *c.d, c.m);
                        stage2_post(th, sh, t, v, true, c);
                        if (next) fetch_next(th, sh, t, v, c);
                    });
                } else {
                    run.phase([&](Thread& th, int t) {
                        fetch_ifact(th, t, v, 1, c);
                        stage2_post(th, sh, t, v, false, c);
                        if (next) fetch_next(th, sh, t, v, c);
                    });
                    if (Xd >= 0) defer_phase(run, sh, Xd, kd + 2, kd + 4, c);
                }
            } else {
                run.phase([&](Thread&, int t) {
                    zero_col(c.a.Vst, t, v, v + 1, nm + 1, c);
                    if (t == 0) sh.scal[1] = 0.0;  // the hand-over entry (v+1, v) is zero too
                });
                if (Xd >= 0) defer_phase(run, sh, Xd, kd + 2, kd + 4, c);
            }
            pref = next;
            if ((v & 15) == 15) {
                if (kDefer) {
                    run.phase([&](Thread&, int t) { flush_head(t, v >> 4, c); });
                } else {
                    run.phase([&](Thread&, int t) { flush_tiles(sh, t, v >> 4, c); });
                    run.phase([&](Thread&, int t) { flush(sh, t, v >> 4, c); });
                }
            }
        }
        // row nm of V only feeds the queries with n = nm
        if (nm > c.v0) {
            if ((nm >> 4) != segY) run.phase([&](Thread&, int t) { load_segrow(sh, t, nm >> 4, c); });
            const bool now = !pref;
            run.phase([&](Thread& th, int t) {
                stage1_pre(th, sh, t, nm, now, c);
                keep_row_out(th, t, nm, c);
            });
            queries(run, sh, nm, 0, c);
        }
        if (c.d->keep) run.phase([&](Thread&, int t) { keep_factorials(t, c); });
    }

    // Canonical N_bad mod p of each query (fast/cuckoo_fast.cu:884-893).
    static CF_HD double result(int j, const Ctx& c) {
        const QueryDesc& q = c.q[c.d->q_off + j];
        return canon(mulred(c.a.qtot[2 * j], c.a.fact[q.M], c.m), c.m);
    }
};

}  // namespace cf
