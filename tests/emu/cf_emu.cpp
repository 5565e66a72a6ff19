// cf_emu.cpp -- CPU emulation of the engine's device kernels, for tests only (never a product
// path). The kernel headers (cf_core.hpp, cf_task.hpp, cf_wide.hpp) are compiled for the host and
// driven by an emulated runner: a phase runs every thread of the workgroup in turn (the device
// runs them at once and separates phases by a workgroup barrier), wave sums add the lanes of a
// wave, and a task spread over 16 workgroups runs them on 16 host threads whose group barrier is
// a host barrier. Built with CF_EMU_CHECK, every modular product of the kernels is checked
// exactly with 128-bit integers: its inputs and result are integers, the result is congruent to
// the product modulo p, and |result| stays below 2^52 (one spare bit under the mantissa).
//
//   cf_emu [--lg L] [--defer D] [--reverse] [--serial-wg 0|1|2] [--poison] d model p n1 M1 [n2 M2 ...]
//
// prints one line per query, "n M residue", in the order given, then a line
// "products N max_ratio R" (the largest |product result| / p seen) and exits 0; a failed check
// prints "EXACTNESS ..." and exits 3. --lg picks the transform length (default: the engine's
// choice for the largest n; any L >= that n is valid, 16 and 17 select the 16-workgroup kernel),
// --defer -1/0 the tile-flush variant of the one-workgroup kernel, --reverse runs the threads of
// every phase in descending order (results must not change: a phase may not read what another
// thread of the same phase writes).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "cf_host.hpp"
#include "cf_task.hpp"
#include "cf_wide.hpp"

using namespace cf;

namespace {

std::atomic<long long> g_products{0};
std::atomic<bool> g_bad{false};
std::mutex g_ratio_mu;
double g_max_ratio = 0;
bool g_reverse = false;
bool g_poison = false;

}  // namespace

void cf::emu_check_mulred(double a, double b, double r, const Mod& m) {
    thread_local long long local = 0;
    thread_local double local_max = 0;
    const double lim = 4503599627370496.0;  // 2^52
    bool ok = a == std::rint(a) && b == std::rint(b) && r == std::rint(r) && std::fabs(a) < 2 * lim &&
              std::fabs(b) < 2 * lim && std::fabs(r) < lim;
    if (ok) {
        const __int128 p = (__int128)(int64_t)m.p;
        const __int128 diff = (__int128)(int64_t)a * (__int128)(int64_t)b - (__int128)(int64_t)r;
        ok = diff % p == 0;
    }
    if (!ok && !g_bad.exchange(true))
        std::printf("EXACTNESS a=%.0f b=%.0f r=%.0f p=%.0f\n", a, b, r, m.p);
    const double ratio = std::fabs(r) / m.p;
    if (ratio > local_max) {
        local_max = ratio;
        std::lock_guard<std::mutex> lk(g_ratio_mu);
        g_max_ratio = std::max(g_max_ratio, ratio);
    }
    if (++local == 4096) {  // counted in blocks (the total is short by under 4096 per host thread)
        g_products += local;
        local = 0;
    }
}

namespace {

// The emulated workgroup runner: T threads' states, phases run thread by thread.
template <int T>
struct EmuRunner {
    std::vector<Thread> th = std::vector<Thread>(T);
    // per destination array: the running sums of each wave (a phase may reduce into several)
    struct Acc {
        double* red;
        double w[T / 64][8];
    };
    std::vector<Acc> accs;
    double (&acc_of(double* red))[T / 64][8] {
        for (auto& a : accs)
            if (a.red == red) return a.w;
        accs.push_back(Acc{red, {}});
        return accs.back().w;
    }
    template <class F>
    void phase(F f) {
        if (g_reverse)
            for (int t = T - 1; t >= 0; --t) f(th[t], t);
        else
            for (int t = 0; t < T; ++t) f(th[t], t);
    }
    // Inside a phase: this wave's sums of the cnt partials (the device adds them across lanes with
    // DPP and shuffles; here they accumulate lane by lane), written to red[wave][8] once the whole
    // wave has contributed.
    template <class F>
    void wave_partial8(double* red, Thread& th_, int t_, F f, const Mod& m) {
        double acc[8];
        const int cnt = f(th_, t_, acc);
        const int w = t_ >> 6;
        double(&wacc)[T / 64][8] = acc_of(red);
        const int lane = t_ & 63, first = g_reverse ? 63 : 0, last = g_reverse ? 0 : 63;
        for (int j = 0; j < 8; ++j) wacc[w][j] = lane == first ? acc[j] : (j < cnt ? reduce(wacc[w][j] + acc[j], m) : wacc[w][j]);
        if (lane == last)
            for (int j = 0; j < 8; ++j) red[w * 8 + j] = wacc[w][j];
    }
    template <class F>
    void wave_reduce8(double* red, F f, const Mod& m) {
        phase([&](Thread& th_, int t_) { wave_partial8(red, th_, t_, f, m); });
    }
};

// A host barrier for the workgroups of one wide task.
struct HostBarrier {
    std::mutex mu;
    std::condition_variable cv;
    int n, waiting = 0;
    long gen = 0;
    explicit HostBarrier(int k) : n(k) {}
    void wait() {
        std::unique_lock<std::mutex> lk(mu);
        const long g = gen;
        if (++waiting == n) {
            waiting = 0;
            ++gen;
            cv.notify_all();
        } else {
            cv.wait(lk, [&] { return gen != g; });
        }
    }
};

// Workgroups run one at a time in a fixed order between group barriers (--serial-wg 1: 0, 1, ..,
// 15; 2: 15, .., 0): a phase that reads another workgroup's writes without a group barrier between
// them then sees the same (old or new) value on every run, and the two orders see different ones.
struct TurnBarrier {
    std::mutex mu;
    std::condition_variable cv;
    int n, arrived = 0, turn;
    bool rev;
    TurnBarrier(int k, bool reverse) : n(k), turn(reverse ? k - 1 : 0), rev(reverse) {}
    int first() const { return rev ? n - 1 : 0; }
    int after(int g) const { return rev ? g - 1 : g + 1; }
    void begin(int g) {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return turn == g; });
    }
    void wait(int g) {
        std::unique_lock<std::mutex> lk(mu);
        turn = ++arrived == n ? (arrived = 0, first()) : after(g);
        cv.notify_all();
        cv.wait(lk, [&] { return turn == g; });
    }
    void end(int g) {
        std::lock_guard<std::mutex> lk(mu);
        turn = after(g);
        cv.notify_all();
    }
};

template <int T>
struct WideEmuRunner : EmuRunner<T> {
    HostBarrier* bar = nullptr;
    TurnBarrier* turns = nullptr;
    int g = 0;
    bool gsync() {
        if (turns) turns->wait(g);
        else bar->wait();
        return false;
    }
};

int g_serial = 0;

struct Job {
    int d = 3, model = 0, lg = 0, defer = -1;
    uint64_t p = 0;
    std::vector<QueryDesc> q;  // sorted by n, largest first; out = input position
    int nm = 0, F1 = 0;
    SlotPlan sp;
    TaskDesc td{};
    std::vector<double> mem;
    std::vector<uint64_t> res;
};

void describe(Job& j) {
    j.sp = plan_slots(j.nm);
    TaskDesc& t = j.td;
    std::memset(&t, 0, sizeof t);
    if (j.lg >= 16) wide_constants(j.p, j.lg, t);
    else task_constants(j.p, j.lg, t);
    t.arena = 0;
    t.nm = j.nm;
    t.F1 = j.F1;
    t.q_off = 0;
    t.q_cnt = (int)j.q.size();
    t.slot_off = 0;
    t.pool = j.sp.pool;
    t.NB = j.sp.NB;
    t.NSEG = j.sp.NSEG;
    t.NS = j.sp.NS;
    t.NT = (1 << j.lg) / 16;
    t.lg = j.lg;
    t.dd = j.d;
    t.model = j.model;
    t.keep = nullptr;
    // --poison: the arena starts as NaN, LDS and the threads' register arrays as all-ones bytes
    // (none is cleared between launches), so a value read before the kernel wrote it fails the
    // product checks
    j.mem.assign(task_bytes(j.sp, j.F1, (int)j.q.size(), j.lg) / 8 + 64, g_poison ? std::nan("") : 0.0);
    j.res.assign(j.q.size(), 0);
}

template <int LG, int DF>
void run_one_workgroup(Job& j) {
    using TK = Task<LG, LG, DF>;
    auto run = std::make_unique<EmuRunner<Plan<LG>::T>>();
    if (g_poison) std::memset((void*)run->th.data(), 0xff, sizeof(Thread) * run->th.size());
    auto sh = std::make_unique<typename TK::Sh>();
    if (g_poison) std::memset((void*)sh.get(), 0xff, sizeof(typename TK::Sh));
    Ctx c;
    c.d = &j.td;
    c.m.p = j.td.p;
    c.m.pinv = j.td.pinv;
    c.a = arena_of(j.mem.data(), j.td);
    c.slots = j.sp.segs.data();
    c.q = j.q.data();
    c.model = j.model;
    c.dd = j.d;
    c.v0 = j.model ? 1 : j.d;
    c.es = Plan<LG>::T;
    c.eb = 0;
    c.L = Plan<LG>::L;
    TK::run_task(*run, *sh, c);
    for (size_t k = 0; k < j.q.size(); ++k) j.res[(size_t)j.q[k].out] = (uint64_t)TK::result((int)k, c);
}

template <int INNER>
void run_wide(Job& j) {
    using W = Wide<INNER>;
    HostBarrier bar(kWideG);
    TurnBarrier turns(kWideG, g_serial == 2);
    std::vector<std::thread> ws;
    for (int g = 0; g < kWideG; ++g)
        ws.emplace_back([&, g] {
            auto run = std::make_unique<WideEmuRunner<W::T>>();
            if (g_poison) std::memset((void*)run->th.data(), 0xff, sizeof(Thread) * run->th.size());  // nor are registers
            run->bar = &bar;
            run->g = g;
            if (g_serial) {
                run->turns = &turns;
                turns.begin(g);
            }
            auto sh = std::make_unique<typename W::Sh>();
            if (g_poison) std::memset((void*)sh.get(), 0xff, sizeof(typename W::Sh));  // LDS is not cleared either
            const Ctx c = W::ctx(&j.td, j.q.data(), j.sp.segs.data(), j.mem.data(), g, j.model, j.d);
            const WideArena w = wide_arena_of<INNER>(j.mem.data(), j.td);
            W::run_task(*run, *sh, c, w, g);
            if (g == 0)
                for (size_t k = 0; k < j.q.size(); ++k) j.res[(size_t)j.q[k].out] = (uint64_t)W::TK::result((int)k, c);
            if (g_serial) turns.end(g);
        });
    for (auto& t : ws) t.join();
}

template <int DF>
bool dispatch(Job& j) {
    switch (j.lg) {
        case 10: run_one_workgroup<10, DF>(j); return true;
        case 11: run_one_workgroup<11, DF>(j); return true;
        case 12: run_one_workgroup<12, DF>(j); return true;
        case 13: run_one_workgroup<13, DF>(j); return true;
        case 14: run_one_workgroup<14, DF>(j); return true;
        default: return false;
    }
}

}  // namespace

int main(int argc, char** argv) {
    Job j;
    int i = 1;
    for (; i < argc && argv[i][0] == '-' && argv[i][1] == '-'; ++i) {
        if (!std::strcmp(argv[i], "--lg") && i + 1 < argc) j.lg = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--defer") && i + 1 < argc) j.defer = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--reverse")) g_reverse = true;
        else if (!std::strcmp(argv[i], "--serial-wg") && i + 1 < argc) g_serial = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--poison")) g_poison = true;
        else {
            std::fprintf(stderr, "unknown option %s\n", argv[i]);
            return 2;
        }
    }
    if (argc - i < 5 || (argc - i - 3) % 2) {
        std::fprintf(stderr, "usage: cf_emu [--lg L] [--defer D] [--reverse] [--serial-wg 0|1|2] [--poison] d model p n1 M1 [n2 M2 ...]\n");
        return 2;
    }
    j.d = std::atoi(argv[i]);
    j.model = std::atoi(argv[i + 1]);
    j.p = std::strtoull(argv[i + 2], nullptr, 10);
    int Mmax = 0;
    for (int k = i + 3, out = 0; k + 1 < argc; k += 2, ++out) {
        QueryDesc q;
        q.n = std::atoi(argv[k]);
        q.M = std::atoi(argv[k + 1]);
        q.out = out;
        if (q.n < 1 || q.M <= q.n || (uint64_t)q.M >= j.p) {
            std::fprintf(stderr, "bad query %d %d\n", q.n, q.M);
            return 2;
        }
        j.q.push_back(q);
        j.nm = std::max(j.nm, q.n);
        Mmax = std::max(Mmax, q.M);
    }
    std::vector<QueryDesc> given = j.q;
    std::stable_sort(j.q.begin(), j.q.end(), [](const QueryDesc& a, const QueryDesc& b) { return a.n > b.n; });
    j.F1 = std::max(j.nm, Mmax) + 1;
    const int need_lg = engine_log2_length(j.nm, j.p);
    if (need_lg < 0) {
        std::fprintf(stderr, "p - 1 lacks the transform's power of two\n");
        return 2;
    }
    if (!j.lg) j.lg = need_lg;
    if (j.lg < need_lg || (j.lg > 14 && j.lg != 16 && j.lg != 17) || two_adic_valuation(j.p - 1) < j.lg) {
        std::fprintf(stderr, "transform 2^%d cannot run n = %d modulo %llu\n", j.lg, j.nm, (unsigned long long)j.p);
        return 2;
    }
    describe(j);
    const bool ok = j.lg == 16   ? (run_wide<12>(j), true)
                    : j.lg == 17 ? (run_wide<13>(j), true)
                                 : (j.defer == 0 ? dispatch<0>(j) : dispatch<-1>(j));
    if (!ok) return 2;
    for (size_t k = 0; k < given.size(); ++k) std::printf("%d %d %llu\n", given[k].n, given[k].M, (unsigned long long)j.res[k]);
    std::printf("products %lld max_ratio %.6f\n", (long long)g_products.load(), g_max_ratio);
    return g_bad ? 3 : 0;
}
