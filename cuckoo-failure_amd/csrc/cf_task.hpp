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
};

template <int LOG2L>
struct Task {
    using P = Plan<LOG2L>;
    using E = Engine<LOG2L>;
    using Sh = Shared<LOG2L>;
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
        for (int i = t; i < c.d->F1; i += T) c.a.inv[i] = i ? reduce(mulred(c.a.ifact[i], c.a.fact[i - 1], c.m), c.m) : 0.0;
    }
    static CF_HD void zero_staging(int t, const Ctx& c) {
        const int64_t n = 16 * (int64_t)c.d->NS;
        for (int64_t i = t; i < n; i += T) c.a.Vst[i] = 0.0, c.a.Sst[i] = 0.0;
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
            const int e = k * T + t;
            th.x[k] = (e >= 1 && e < c.d->nm) ? c.a.inv[e] : 0.0;
        }
    }

    // ---------------- per step ----------------
    static CF_HD void stage1_pre(Thread& th, int t, int v, const Ctx& c) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * T + t;
            double a = 0.0;
            if (e < v) a = mulred(load_tab(c.a.Vpool, c.a.Vst, c.slots, c.d->NS, v, e), c.a.ifact[v - 1 - e], c.m);
            th.x[k] = a;
        }
    }
    // Partial sums of the queries j0 .. j0+cnt-1 over this thread's entries of row v.
    static CF_HD void query_partial(const Thread& th, int t, int v, int j0, int cnt, double (&acc)[8], const Ctx& c) {
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            double s = 0.0;
            if (jj < cnt) {
                const int M = c.q[c.d->q_off + j0 + jj].M;
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const int e = k * T + t;
                    if (e < v) s += mulred(th.x[k], c.a.inv[M - e], c.m);
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
    static CF_HD void stage1_post(const Thread& th, int t, int v, const Ctx& c) {
        double* col = c.a.Sst + (int64_t)(v & 15) * c.d->NS;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * T + t;
            if (e >= v && e < c.d->nm) col[e] = mulred(th.x[k], c.a.ifact[e - v], c.m);
        }
    }
    // Column v of a staging area set to zero on rows lo..hi-1.
    static CF_HD void zero_col(double* st, int t, int v, int lo, int hi, const Ctx& c) {
        double* col = st + (int64_t)(v & 15) * c.d->NS;
        for (int r = lo + t; r < hi; r += T) col[r] = 0.0;
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

    static CF_HD void stage2_pre(Thread& th, const Sh& sh, int t, int v, const Ctx& c) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * T + t;
            double b = 0.0;
            if (e >= 1 && e <= v) {
                const double x = load_tab(c.a.Spool, c.a.Sst, c.slots, c.d->NS, v, e);
                const double f = mulred(gpow(sh.lo_pre, sh, v - e, c.m), c.a.ifact[v - e], c.m);
                b = mulred(x, f, c.m);
            }
            th.x[k] = b;
        }
    }
    static CF_HD void stage2_post(const Thread& th, const Sh& sh, int t, int v, bool conv, const Ctx& c) {
        double* col = c.a.Vst + (int64_t)(v & 15) * c.d->NS;
        const double s1 = gpow(sh.lo_post, sh, v, c.m);  // g^v / v!
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int e = k * T + t;
            const int p = e ? e : L;  // c[p mod L]: entry 0 stands for p = L
            if (p > v && p <= c.d->nm) {
                const int j = p - v;
                const double f = mulred(gpow(sh.lo_post, sh, j, c.m), c.a.ifact[j - 1], c.m);
                const double u = mulred(s1, c.a.inv[p], c.m) - (conv ? reduce(th.x[k], c.m) : 0.0);
                col[p] = mulred(u, f, c.m);
            }
        }
    }
    // Staging column block X -> pool tiles (Y, X) for every row block Y > X.
    static CF_HD void flush(double* pool, const double* st, int t, int X, const Ctx& c) {
        const int n = (c.d->NB - 1 - X) * 256;
        for (int idx = t; idx < n; idx += T) {
            const int Y = X + 1 + (idx >> 8), col = (idx >> 4) & 15, r = idx & 15;
            pool[(int64_t)c.slots[Y * (Y - 1) / 2 + X] * 256 + (r << 4) + col] = st[(int64_t)col * c.d->NS + 16 * Y + r];
        }
    }

    // ---------------- the convolution: x <- x (*) k ----------------
    template <class R>
    static CF_HD void convolve(R& run, Sh& sh, const Ctx& c) {
        run.phase([&](Thread& th, int t) { E::fwd0(th, sh, t, *c.d, c.m, c.a.tw0); });
        for (int r = 1; r < P::NTW; ++r) run.phase([&](Thread& th, int t) { E::fwdr(r, th, sh, t, *c.d, c.m); });
        run.phase([&](Thread& th, int t) { E::final_round(false, th, sh, t, *c.d, c.m, c.a.hat); });
        for (int r = P::NTW - 1; r >= 1; --r) run.phase([&](Thread& th, int t) { E::invr(r, th, sh, t, *c.d, c.m); });
        run.phase([&](Thread& th, int t) { E::inv0(th, sh, t, *c.d, c.m, c.a.tw0); });
    }

    // Row v of Vh is in th.x (stage-1 input form); add its query terms.
    template <class R>
    static CF_HD void queries(R& run, Sh& sh, int v, const Ctx& c) {
        // queries are sorted by n, largest first: the active ones (n >= v) are a prefix
        int nact = 0;
        while (nact < c.d->q_cnt && c.q[c.d->q_off + nact].n >= v) ++nact;
        for (int j0 = 0; j0 < nact; j0 += 8) {
            const int cnt = nact - j0 < 8 ? nact - j0 : 8;
            run.wave_reduce8(sh.red, [&](Thread& th, int t, double (&acc)[8]) { query_partial(th, t, v, j0, cnt, acc, c); }, c.m);
            run.phase([&](Thread&, int t) {
                if (t < cnt) {
                    double s = 0.0;
                    for (int w = 0; w < NW; ++w) s += sh.red[w * 8 + t];  // NW <= 16 values below p/2 + 4
                    query_update(t, v, j0, reduce(s, c.m), c);
                }
            });
        }
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
            E::twiddles(sh, t, *c.d, c.m, c.a.tw0);
        });
        run.phase([&](Thread& th, int t) {
            query_init(t, c);
            kernel_input(th, t, c);
        });
        // kernel spectrum (times 1/L) into the arena
        run.phase([&](Thread& th, int t) { E::fwd0(th, sh, t, *c.d, c.m, c.a.tw0); });
        for (int r = 1; r < P::NTW; ++r) run.phase([&](Thread& th, int t) { E::fwdr(r, th, sh, t, *c.d, c.m); });
        run.phase([&](Thread& th, int t) { E::final_round(true, th, sh, t, *c.d, c.m, c.a.hat); });

        // ---- the v-loop ----
        for (int v = 1; v < nm; ++v) {
            // stage 1: rows v <= v0 of V are zero, so column v of S is zero
            if (v > c.v0) {
                run.phase([&](Thread& th, int t) { stage1_pre(th, t, v, c); });
                queries(run, sh, v, c);
                convolve(run, sh, c);
                run.phase([&](Thread& th, int t) { stage1_post(th, t, v, c); });
            } else {
                run.phase([&](Thread&, int t) { zero_col(c.a.Sst, t, v, v, nm, c); });
            }
            // stage 2: g(v) = 0 below v0, so column v of V is zero
            if (v >= c.v0) {
                const double g = g_of(v, c);
                run.phase([&](Thread&, int t) { powtab(sh, t, v, g, c); });
                const bool conv = v > c.v0;  // row v of S is zero unless v > v0
                if (conv) {
                    run.phase([&](Thread& th, int t) { stage2_pre(th, sh, t, v, c); });
                    convolve(run, sh, c);
                }
                run.phase([&](Thread& th, int t) { stage2_post(th, sh, t, v, conv, c); });
            } else {
                run.phase([&](Thread&, int t) { zero_col(c.a.Vst, t, v, v + 1, nm + 1, c); });
            }
            if ((v & 15) == 15) {
                run.phase([&](Thread&, int t) {
                    flush(c.a.Vpool, c.a.Vst, t, v >> 4, c);
                    flush(c.a.Spool, c.a.Sst, t, v >> 4, c);
                });
            }
        }
        // row nm of V only feeds the queries with n = nm
        if (nm > c.v0) {
            run.phase([&](Thread& th, int t) { stage1_pre(th, t, nm, c); });
            queries(run, sh, nm, c);
        }
    }

    // Canonical N_bad mod p of each query (fast/cuckoo_fast.cu:884-893).
    static CF_HD double result(int j, const Ctx& c) {
        const QueryDesc& q = c.q[c.d->q_off + j];
        return canon(mulred(c.a.qtot[2 * j], c.a.fact[q.M], c.m), c.m);
    }
};

}  // namespace cf
