// cf_wide.hpp -- one prime on kWideG = 16 workgroups: the suffix-table recurrence with transforms
// of length L = 16 N2 (N2 = 2^INNER: L = 2^16 with INNER = 12, 2^17 with INNER = 13), which do not
// fit one CU's LDS.
//
// The transform is split four-step: entry e = N2 n1 + n2 (n1 < 16, n2 < N2). Workgroup g owns the
// columns n2 = T g + t (t < T = N2 / 16): its thread t holds the 16 entries N2 k + T g + t in
// registers, which is also the element map of every stage function (Ctx::es = N2, Ctx::eb = T g).
// A convolution is then
//   outer: DFT16 over n1 in registers, twiddle w^(n2 k1) (w of order L), store row k1 of the
//          exchange buffer (16 x N2 doubles in HBM);
//   group barrier;
//   inner: workgroup g loads row k1 = g and runs the one-workgroup engine of length N2
//          (Engine<INNER>: rounds in LDS, pointwise product with its slice of the kernel spectrum,
//          inverse rounds), storing the row back;
//   group barrier;
//   outer inverse: load the thread's column, twiddle w^-(n2 k1), inverse DFT16.
// Query terms are reduced per workgroup, written to per-workgroup slots (one per query) and added
// up by workgroup 0 after the next group barrier; up to 8 queries ride in the stage-1 input phase,
// more take two phases per group of 8 of their own before the outer transform. The factorial tables
// are built by workgroup 0; staging zeroing and the tile flush are split over all workgroups.
//
// Reference correspondence is that of cf_task.hpp (fast/cuckoo_fast.cu:787-898); only the
// transform's distribution differs.
#pragma once
#include "cf_task.hpp"

namespace cf {

constexpr int kWideG = 16;  // workgroups per prime
static_assert(kWideG == 16, "wide_extra_doubles() assumes 16 workgroups per prime");

// Extra arena of a wide task, after the one-workgroup layout of length L = 16 N2 (doubles).
struct WideArena {
    double* exch;   // 16 x N2 exchange rows
    double* qpart;  // [g][qs] per-workgroup query sums
    double* hand;   // [0]: the hand-over entry when it crosses workgroups
    int qs;         // query slots per workgroup: max(8, the task's query count)
};
template <int INNER>
CF_HD WideArena wide_arena_of(double* base, const TaskDesc& t) {
    constexpr int64_t N2 = int64_t(1) << INNER;
    WideArena w;
    w.exch = base + t.arena + arena_doubles(t.pool, t.NS, t.F1, t.q_cnt, (int)(16 * N2));
    w.qpart = w.exch + 16 * N2;
    w.qs = wide_qslots(t.q_cnt);
    w.hand = w.qpart + (int64_t)kWideG * w.qs;
    return w;
}

template <int INNER>
struct Wide {
    static constexpr int LOG2L = INNER + 4;
    using TK = Task<INNER, LOG2L>;  // stage functions (wide element map), LDS layout
    using E = Engine<INNER>;        // the inner transform
    using P = Plan<INNER>;
    static constexpr int T = P::T, G = kWideG, NW = T / 64, N2 = 1 << INNER;
    // LDS: the one-workgroup layout, and the outer twiddles of this workgroup's columns:
    // otw[k T + t] = w^(n2 2^k) (k < 4), otw[(4 + k) T + t] = w^-(n2 2^k), n2 = T g + t
    struct Sh : Shared<INNER, LOG2L> {
        double otw[8 * T];
    };

    // The context of workgroup g: the wide element map and its slice of the spectrum.
    static CF_HD Ctx ctx(const TaskDesc* d, const QueryDesc* qs, const int* slots, double* mem, int g, int model, int dd) {
        Ctx c;
        c.d = d;
        c.m.p = d->p;
        c.m.pinv = d->pinv;
        c.a = arena_of(mem, *d);
        c.a.hat += (int64_t)N2 * g;
        c.slots = slots + d->slot_off;
        c.q = qs;
        c.model = model;
        c.dd = dd;
        c.v0 = model ? 1 : dd;
        c.es = N2;
        c.eb = T * g;
        c.L = 1 << LOG2L;
        return c;
    }

    // ---- the outer step ----
    static CF_HD void outer_twiddles(Sh& sh, int g, int t, const TaskDesc& d, const Mod& m) {
        const uint64_t n2 = (uint64_t)(T * g + t);
        double f = powmod(d.wroot, n2, m), i = powmod(d.wrooti, n2, m);
        for (int k = 0; k < 4; ++k) {
            sh.otw[k * T + t] = f, sh.otw[4 * T + k * T + t] = i;
            f = reduce(mulred(f, f, m), m), i = reduce(mulred(i, i, m), m);
        }
    }
    // this thread's w^(n2 2^k), k < 4 (dir 0), or w^-(n2 2^k) (dir 1)
    static CF_HD void outer_bases(double (&b4)[4], const Sh& sh, int t, int dir) {
#pragma unroll
        for (int k = 0; k < 4; ++k) b4[k] = sh.otw[(4 * dir + k) * T + t];
    }
    // Power tables of g for step v (lo_pre[i] = v! g^i, lo_post[i] = g^i / v!, hi[j] = g^(64 j)):
    // thread t makes hi[t + T i] as (g^64)^t times (g^(64 T))^i, a chain of about 20 products
    // instead of one power per entry.
    static CF_HD void powtab(Sh& sh, int t, int v, double gv, const Ctx& c) {
        const Mod& m = c.m;
        if (t < 128) {
            const double x = powmod(gv, (uint64_t)(t & 63), m);
            if (t < 64) sh.lo_pre[t] = reduce(mulred(x, c.a.fact[v], m), m);
            else sh.lo_post[t - 64] = reduce(mulred(x, c.a.ifact[v], m), m);
        }
        const int nhi = (c.d->nm >> 6) + 1;
        if (t < nhi) {
            const double g64 = powmod(gv, 64, m), step = powmod(g64, (uint64_t)T, m);
            double h = powmod(g64, (uint64_t)t, m);
            for (int j = t; j < nhi; j += T) {
                sh.hi[j] = h;
                h = reduce(mulred(h, step, m), m);
            }
        }
    }
    static CF_HD void outer_fwd(Thread& th, const Sh& sh, int t, const TaskDesc& d, const Mod& m, const WideArena& w,
                                int g) {
        double b4[4];
        outer_bases(b4, sh, t, 0);
        reduce16(th.x, m);
        dft_groups<16>(th.x, d.W, m);
        twiddle_chain(th.x, b4, m);
#pragma unroll
        for (int k = 0; k < 16; ++k) w.exch[k * N2 + T * g + t] = th.x[k];
    }
    static CF_HD void outer_inv(Thread& th, const Sh& sh, int t, const TaskDesc& d, const Mod& m, const WideArena& w,
                                int g) {
#pragma unroll
        for (int k = 0; k < 16; ++k) th.x[k] = w.exch[k * N2 + T * g + t];
        double b4[4];
        outer_bases(b4, sh, t, 1);
        twiddle_chain(th.x, b4, m);
        reduce16(th.x, m);
        dft_groups<16>(th.x, d.Wi, m);
    }
    // inner row g: exchange -> registers (entries T k + t) and back
    static CF_HD void inner_load(Thread& th, const WideArena& w, int g, int t) {
#pragma unroll
        for (int k = 0; k < 16; ++k) th.x[k] = w.exch[(int64_t)g * N2 + T * k + t];
    }
    static CF_HD void inner_store(const Thread& th, const WideArena& w, int g, int t) {
#pragma unroll
        for (int k = 0; k < 16; ++k) w.exch[(int64_t)g * N2 + T * k + t] = th.x[k];
    }

    // The inner part of a convolution (between the two group barriers). The factorials of the
    // stage output (off: 0 stage 1, 1 stage 2) are fetched into th.pa on the way; with qn > 0 the
    // first phase also moves this workgroup's sums of queries 0 .. qn-1 of the phase before the
    // barrier (in LDS) to its slots, which workgroup 0 reads after the next group barrier; with
    // tab the second phase makes the power tables of g(v) for the stage-2 phases that follow.
    template <class R>
    static CF_HD void inner(R& run, Sh& sh, const WideArena& w, int g, int v, int off, bool make_hat,
                            const Ctx& c, int qn = 0, bool tab = false, double gv = 0.0) {
        run.phase([&](Thread& th, int t) {
            if (qn) query_slots(sh, w, g, t, 0, qn, c.m);
            inner_load(th, w, g, t);
            E::fwd0(th, sh, t, *c.d, c.m);
        });
        static_for<1, P::NTW>([&](auto RI) {
            constexpr int r = decltype(RI)::value;
            run.phase([&](Thread& th, int t) {
                if (!make_hat && r == P::NTW - 1) TK::fetch_hat(th, t, c);
                E::template fwdr<r>(th, sh, t, *c.d, c.m);
                if (r == 1 && tab) powtab(sh, t, v, gv, c);
            });
        });
        run.phase([&](Thread& th, int t) { E::final_round(make_hat, th, sh, t, *c.d, c.m, c.a.hat); });
        if (make_hat) return;
        static_rfor<1, P::NTW>([&](auto RI) {
            constexpr int r = decltype(RI)::value;
            run.phase([&](Thread& th, int t) {
                if (r == 1) TK::fetch_ifact(th, t, v, off, c);
                E::template invr<r>(th, sh, t, *c.d, c.m);
            });
        });
        run.phase([&](Thread& th, int t) {
            E::inv0(th, sh, t, *c.d, c.m);
            inner_store(th, w, g, t);
        });
    }

    // Query terms of row v (th.x, stage-1 input form) for queries j0 .. j0+cnt-1 (cnt <= 8): this
    // workgroup's wave sums to LDS (inside the caller's phase), then to the workgroup's slots.
    template <class R>
    static CF_HD void query_wave(R& run, Thread& th, Sh& sh, int t, int v, int j0, int cnt, const Ctx& c) {
        if (cnt)
            run.wave_partial8(sh.red, th, t, [&](Thread& th2, int t2, double (&acc)[8]) {
                TK::query_partial(th2, t2, v, j0, cnt, acc, c);
                return cnt;
            }, c.m);
    }
    static CF_HD void query_slots(const Sh& sh, const WideArena& w, int g, int t, int j0, int cnt, const Mod& m) {
        if (t < cnt) {
            double s = 0.0;
            for (int wv = 0; wv < NW; ++wv) s += sh.red[wv * 8 + t];  // up to 8 values below p/2 + 4
            w.qpart[(int64_t)g * w.qs + j0 + t] = reduce(s, m);
        }
    }
    // Queries beyond the first 8 of row v: per group of 8, the wave sums in one phase and the
    // workgroup's slots in the next (th.x keeps the stage-1 input throughout).
    template <class R>
    static CF_HD void query_groups(R& run, Sh& sh, const WideArena& w, int g, int v, int nact, const Ctx& c) {
        for (int j0 = 8; j0 < nact; j0 += 8) {
            const int cnt = nact - j0 < 8 ? nact - j0 : 8;
            run.phase([&](Thread& th, int t) { query_wave(run, th, sh, t, v, j0, cnt, c); });
            run.phase([&](Thread&, int t) { query_slots(sh, w, g, t, j0, cnt, c.m); });
        }
    }
    // Workgroup 0 finishes the queries from all workgroups' slots (after a group barrier).
    static CF_HD void query_total(const WideArena& w, int g, int t, int v, int cnt, const Ctx& c) {
        if (g != 0) return;
        for (int j = t; j < cnt; j += T) {
            double s = 0.0;
            for (int h = 0; h < G; ++h) {
                s += w.qpart[(int64_t)h * w.qs + j];
                if ((h & 3) == 3) s = reduce(s, c.m);
            }
            TK::query_update(j, v, 0, reduce(s, c.m), c);
        }
    }

    template <class R>
    static CF_HD bool run_task(R& run, Sh& sh, const Ctx& c, const WideArena& w, int g) {
        const int nm = c.d->nm;
        const TaskDesc& d = *c.d;
        // ---- prologue: workgroup 0 builds the factorial tables ----
        run.phase([&](Thread&, int t) {
            TK::zero_staging(t, c);
            if (g == 0) TK::fact_local(sh, t, c);
        });
        if (g == 0) {
            int src = 0;
            for (int s = 1; s < T; s <<= 1, src ^= 1) run.phase([&](Thread&, int t) { TK::scan_step(sh, t, s, src, c); });
            run.phase([&](Thread&, int t) { TK::fact_fix(sh, t, src, c); });
            run.phase([&](Thread&, int t) {
                if (t == 0) sh.scal[0] = powmod(c.a.fact[d.F1 - 1], (uint64_t)d.p - 2, c.m);
            });
            run.phase([&](Thread&, int t) { TK::ifact_local(sh, t, c); });
            src = 0;
            for (int s = 1; s < T; s <<= 1, src ^= 1) run.phase([&](Thread&, int t) { TK::scan_step(sh, t, s, src, c); });
            run.phase([&](Thread&, int t) { TK::ifact_fix(sh, t, src, c); });
        }
        if (run.gsync()) return false;
        run.phase([&](Thread&, int t) {
            TK::inv_table(t, c);
            E::twiddles(sh, t, d, c.m);
            outer_twiddles(sh, g, t, d, c.m);
            if (g == 0) TK::query_init(t, c);
        });
        if (run.gsync()) return false;
        // kernel spectrum (times 1/L), this workgroup's slice
        run.phase([&](Thread& th, int t) {
            TK::kernel_input(th, t, c);
            outer_fwd(th, sh, t, d, c.m, w, g);
        });
        if (run.gsync()) return false;
        inner(run, sh, w, g, 0, 0, true, c);

        // ---- the v-loop ----
        int segY = -1;
        bool pref = false;
        for (int v = 1; v < nm; ++v) {
            const bool s1 = v > c.v0, s2 = v >= c.v0;
            const double gv = s2 ? TK::g_of(v, c) : 0.0;
            if ((v >> 4) != segY && (s1 || s2)) {
                segY = v >> 4;
                run.phase([&](Thread&, int t) { TK::load_segrow(sh, t, segY, c); });
            }
            if (s1) {
                const int nact = TK::active_queries(v, c);
                const bool now = !pref;
                if (nact <= 8) {
                    // the stage-1 input, its query terms and the outer transform in one phase
                    run.phase([&](Thread& th, int t) {
                        TK::stage1_pre(th, sh, t, v, now, c);
                        query_wave(run, th, sh, t, v, 0, nact, c);
                        outer_fwd(th, sh, t, d, c.m, w, g);
                    });
                } else {
                    run.phase([&](Thread& th, int t) {
                        TK::stage1_pre(th, sh, t, v, now, c);
                        query_wave(run, th, sh, t, v, 0, 8, c);
                    });
                    run.phase([&](Thread&, int t) { query_slots(sh, w, g, t, 0, 8, c.m); });
                    query_groups(run, sh, w, g, v, nact, c);
                    run.phase([&](Thread& th, int t) { outer_fwd(th, sh, t, d, c.m, w, g); });
                }
                if (run.gsync()) return false;
                inner(run, sh, w, g, v, 0, false, c, nact <= 8 ? nact : 0, s2, gv);
                if (run.gsync()) return false;
                run.phase([&](Thread& th, int t) {
                    query_total(w, g, t, v, nact, c);
                    outer_inv(th, sh, t, d, c.m, w, g);
                    TK::stage1_post(th, sh, t, v, c);
                });
            } else {
                run.phase([&](Thread&, int t) {
                    TK::zero_col(c.a.Sst, t, v, v, nm, c);
                    if (s2) powtab(sh, t, v, gv, c);
                });
            }
            const bool next = s2 && (v & 15) != 15;
            if (s2) {
                if (v > c.v0) {
                    run.phase([&](Thread& th, int t) {
                        TK::stage2_pre(th, sh, t, v, c);
                        outer_fwd(th, sh, t, d, c.m, w, g);
                    });
                    if (run.gsync()) return false;
                    inner(run, sh, w, g, v, 1, false, c);
                    if (run.gsync()) return false;
                    run.phase([&](Thread& th, int t) {
                        outer_inv(th, sh, t, d, c.m, w, g);
                        TK::stage2_post(th, sh, t, v, true, c);
                        if (next) TK::fetch_next(th, sh, t, v, c);
                    });
                } else {
                    run.phase([&](Thread& th, int t) {
                        TK::fetch_ifact(th, t, v, 1, c);
                        TK::stage2_post(th, sh, t, v, false, c);
                        if (next) TK::fetch_next(th, sh, t, v, c);
                    });
                }
            } else {
                run.phase([&](Thread&, int t) {
                    TK::zero_col(c.a.Vst, t, v, v + 1, nm + 1, c);
                    if (t == 0) sh.scal[1] = 0.0;
                });
            }
            pref = next;
            // a step without convolutions (v <= v0) had no group barrier, but its staging writes
            // (zeroed columns, the stage-2 output at v = v0) are spread over all workgroups
            if (!s1 && !(s2 && v > c.v0) && run.gsync()) return false;
            // the hand-over entry (v+1, v) was made by the owner of entry v+1; when the owner of
            // entry v (which uses it next step) is another workgroup, it goes through memory
            const bool cross = ((v + 1) & (T - 1)) == 0;
            if (cross) {
                run.phase([&](Thread&, int t) {
                    if (((v + 1) & (N2 - 1)) / T == g && t == 0) w.hand[0] = sh.scal[1];
                });
                if (run.gsync()) return false;
                run.phase([&](Thread&, int t) {
                    if (t == 0) sh.scal[1] = w.hand[0];
                });
            }
            if ((v & 15) == 15) {
                if (!cross && run.gsync()) return false;  // every workgroup's staging columns
                run.phase([&](Thread&, int t) { TK::flush_tiles(sh, t, v >> 4, c); });
                run.phase([&](Thread&, int t) { TK::flush(sh, t, v >> 4, c); });
                if (run.gsync()) return false;  // the tiles, before any row fetch
            }
        }
        // row nm of V only feeds the queries with n = nm
        if (nm > c.v0) {
            if ((nm >> 4) != segY) run.phase([&](Thread&, int t) { TK::load_segrow(sh, t, nm >> 4, c); });
            const bool now = !pref;
            const int nact = TK::active_queries(nm, c);
            const int cnt0 = nact < 8 ? nact : 8;
            run.phase([&](Thread& th, int t) {
                TK::stage1_pre(th, sh, t, nm, now, c);
                query_wave(run, th, sh, t, nm, 0, cnt0, c);
            });
            if (nact) run.phase([&](Thread&, int t) { query_slots(sh, w, g, t, 0, cnt0, c.m); });
            query_groups(run, sh, w, g, nm, nact, c);
            if (run.gsync()) return false;
            run.phase([&](Thread&, int t) { query_total(w, g, t, nm, nact, c); });
        }
        return !run.gsync();  // workgroup 0's query totals are final
    }
};

}  // namespace cf
