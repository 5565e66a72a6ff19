// cf_p62_host.cpp -- host side of the tables modulo primes of 50 to 63 bits (cf_p62.hip): the
// per-prime tables in Montgomery form (R = 2^64), built on host threads as the reference builds
// its per-prime tables on the host (fast/cuckoo_fast.cu:787-837), the launches grouped by
// transform length and sized by free device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "cf_engine.hpp"
#include "cf_host.hpp"
#include "cf_p62.hpp"

using namespace cf;
using namespace cf_int;

namespace {

using u64 = uint64_t;
using u128 = unsigned __int128;

// Montgomery arithmetic modulo p (host), the same reduction as the device's.
struct Mont62 {
    Mod62 m;
    explicit Mont62(u64 p) {
        m.p = p;
        u64 x = p;  // Newton for p^-1 mod 2^64: correct to 3 bits, doubling each step
        for (int i = 0; i < 6; ++i) x *= 2 - p * x;
        m.pinv = x;
        m.one = (u64)(((u128)1 << 64) % p);
        m.r2 = (u64)((u128)m.one * m.one % p);
    }
    u64 mul(u64 a, u64 b) const {
        const u128 t = (u128)a * b;
        const u64 q = (u64)t * m.pinv;
        const u64 hi = (u64)(t >> 64), h2 = (u64)(((u128)q * m.p) >> 64);
        return hi >= h2 ? hi - h2 : hi - h2 + m.p;
    }
    u64 to(u64 x) const { return mul(x % m.p, m.r2); }
    u64 pow(u64 b, u64 e) const {
        u64 r = m.one;
        for (; e; e >>= 1) {
            if (e & 1) r = mul(r, b);
            b = mul(b, b);
        }
        return r;
    }
};

struct Plan62 {
    size_t prime;  // index into the call's primes
    u64 p;
    int nm, F1, lg;
    std::vector<QueryDesc> q;  // sorted by n, largest first
    const SlotPlan* sp;        // tile slots of the tables (shared by the tasks of one nm)
    int64_t words;
};

int log2_length62(int nm) {
    int lg = 4;
    while ((1 << lg) < nm) ++lg;
    return lg;
}

// Words of one prime's host-built tables: fact, ifact, inv, ifact2, mont (F1 each), the forward and
// inverse twiddles (L each), then the roots w_16^r and w_16^-r (r < 8) as plain residues with their
// Shoup quotients floor(v 2^64 / p) (32 words).
size_t table_words62(int F1, int lg) { return 5 * (size_t)F1 + 2 * ((size_t)1 << lg) + 32; }

void prime_tables(const Plan62& pl, Task62& td, u64* out) {
    const Mont62 mt(pl.p);
    const int F1 = pl.F1, L = 1 << pl.lg;
    u64 *fact = out, *ifact = fact + F1, *inv = ifact + F1, *ifact2 = inv + F1, *mont = ifact2 + F1;
    u64 *tw = mont + F1, *twi = tw + L;
    for (int i = 0; i < F1; ++i) mont[i] = mt.to((u64)i);
    fact[0] = mt.m.one;
    for (int i = 1; i < F1; ++i) fact[i] = mt.mul(fact[i - 1], mont[i]);
    ifact[F1 - 1] = mt.pow(fact[F1 - 1], pl.p - 2);
    for (int i = F1 - 1; i > 0; --i) ifact[i - 1] = mt.mul(ifact[i], mont[i]);
    inv[0] = 0;
    for (int i = 1; i < F1; ++i) inv[i] = mt.mul(ifact[i], fact[i - 1]);
    for (int i = 0; i < F1; ++i) ifact2[i] = mt.mul(ifact[i], ifact[i]);
    // a quadratic non-residue x gives a root x^((p-1)/L) of order exactly L
    u64 x = 2;
    while (mt.pow(mt.to(x), (pl.p - 1) / 2) == mt.m.one) ++x;
    const u64 w = mt.pow(mt.to(x), (pl.p - 1) / (u64)L), wi = mt.pow(w, pl.p - 2);
    tw[0] = twi[0] = 0;
    for (int h = 1; h < L; h <<= 1) {  // tw[h + j] = w_{2h}^j, w_{2h} = w^(L / 2h)
        const u64 s = mt.pow(w, (u64)(L / (2 * h))), si = mt.pow(wi, (u64)(L / (2 * h)));
        u64 f = mt.m.one, fi = mt.m.one;
        for (int j = 0; j < h; ++j) {
            tw[h + j] = f, twi[h + j] = fi;
            f = mt.mul(f, s), fi = mt.mul(fi, si);
        }
    }
    u64* sh = twi + L;
    for (int r = 0; r < 8; ++r) {
        const u64 v = mt.mul(tw[8 + r], 1), vi = mt.mul(twi[8 + r], 1);  // out of Montgomery form
        sh[r] = v, sh[8 + r] = (u64)(((u128)v << 64) / pl.p);
        sh[16 + r] = vi, sh[24 + r] = (u64)(((u128)vi << 64) / pl.p);
    }
    td.m = mt.m;
    td.linv = mt.pow(mt.to((u64)L), pl.p - 2);  // 1/L, Montgomery form
    td.nm = pl.nm;
    td.F1 = F1;
    td.L = L;
}

}  // namespace

namespace cf_int {

// Device bytes of one such task (for the drivers' batch sizing).
size_t p62_task_bytes(int nmax, int Mmax, size_t J) {
    const int F1 = std::max(nmax, Mmax) + 1;
    const SlotPlan sp = plan_slots(nmax);
    return 8 * (size_t)arena62_words(sp.NS, sp.pool, F1, 1 << log2_length62(nmax), (int)J);
}

int run_p62(cf_engine* e, int d, int model, const int32_t* n, const int32_t* M, size_t J, const uint64_t* primes,
            const std::vector<size_t>& idx, const uint8_t* mask, uint64_t* res) {
    std::vector<Plan62> plans;
    std::map<int, SlotPlan> slot_plans;  // by nm
    for (size_t b : idx) {
        Plan62 pl;
        pl.prime = b;
        pl.p = primes[b];
        pl.nm = 0;
        int Mmax = 0;
        for (size_t j = 0; j < J; ++j)
            if (!mask || mask[b * J + j]) {
                pl.q.push_back(QueryDesc{n[j], M[j], (int64_t)(b * J + j)});
                pl.nm = std::max(pl.nm, (int)n[j]);
                Mmax = std::max(Mmax, (int)M[j]);
            }
        if (pl.q.empty()) continue;
        if (!(pl.p & 1) || pl.p >= (u64(1) << 63) || (u64)Mmax >= pl.p)
            return fail(e, CF_E_ARGS, "prime " + std::to_string(pl.p) + " is not an odd modulus in (M, 2^63)");
        pl.lg = log2_length62(pl.nm);
        if (pl.lg > kP62MaxLog2L)
            return fail(e, CF_E_UNSUPPORTED, "n = " + std::to_string(pl.nm) + " modulo a prime above 2^50 needs a transform "
                                             "longer than 2^" + std::to_string(kP62MaxLog2L) + ", which this build does not run");
        if (two_adic_valuation(pl.p - 1) < pl.lg)
            return fail(e, CF_E_UNSUPPORTED, "p - 1 of " + std::to_string(pl.p) + " has too few factors of two for n = " +
                                                 std::to_string(pl.nm));
        std::stable_sort(pl.q.begin(), pl.q.end(), [](const QueryDesc& a, const QueryDesc& c) { return a.n > c.n; });
        pl.F1 = std::max(pl.nm, Mmax) + 1;
        auto it = slot_plans.find(pl.nm);
        if (it == slot_plans.end()) it = slot_plans.emplace(pl.nm, plan_slots(pl.nm)).first;
        pl.sp = &it->second;
        pl.words = arena62_words(pl.sp->NS, pl.sp->pool, pl.F1, 1 << pl.lg, (int)pl.q.size());
        plans.push_back(std::move(pl));
    }
    // launches: tasks of one transform length, as many as fit in the device memory
    std::stable_sort(plans.begin(), plans.end(), [](const Plan62& a, const Plan62& b) { return a.lg > b.lg; });
    const size_t usable = usable_bytes(e);
    const cf_stats keep = e->stats;
    float kernel_ms = 0;
    int launches = 0;
    std::vector<u64> out;
    for (size_t at = 0; at < plans.size();) {
        size_t end = at, bytes = 0;
        while (end < plans.size() && plans[end].lg == plans[at].lg && end - at < 65535 &&
               bytes + 8 * (size_t)plans[end].words <= usable)
            bytes += 8 * (size_t)plans[end].words, ++end;
        if (end == at) return fail(e, CF_E_NOMEM, "one prime's tables need " + std::to_string(8 * plans[at].words) + " bytes");
        const size_t k = end - at;
        std::vector<Task62> td(k);
        std::vector<QueryDesc> qd;
        std::vector<int> slots;            // the slot plans of this launch's table sizes
        std::map<int, int> slot_at;        // nm -> offset in slots
        std::vector<size_t> table_off(k);
        size_t table_words = 0;
        int64_t arena = 0;
        for (size_t i = 0; i < k; ++i) {
            const Plan62& pl = plans[at + i];
            td[i].arena = arena;
            td[i].q_off = (int)qd.size();
            td[i].q_cnt = (int)pl.q.size();
            td[i].dd = d;
            td[i].model = model;
            auto so = slot_at.find(pl.nm);
            if (so == slot_at.end()) {
                so = slot_at.emplace(pl.nm, (int)slots.size()).first;
                slots.insert(slots.end(), pl.sp->segs.begin(), pl.sp->segs.end());
            }
            td[i].slot_off = so->second;
            td[i].NSEG = pl.sp->NSEG;
            td[i].NS = pl.sp->NS;
            td[i].pool = pl.sp->pool;
            for (size_t j = 0; j < pl.q.size(); ++j) {
                QueryDesc qq = pl.q[j];
                qq.out = (int64_t)qd.size();  // this launch's result slot
                qd.push_back(qq);
            }
            table_off[i] = table_words;
            table_words += table_words62(pl.F1, pl.lg);
            arena += pl.words;
        }
        std::vector<u64> tables(table_words);
        // the per-prime tables on up to 16 host threads
        const unsigned nth = std::max(1u, std::min<unsigned>(16, std::thread::hardware_concurrency()));
        std::vector<std::thread> ths;
        for (unsigned h = 0; h < nth; ++h)
            ths.emplace_back([&, h] {
                for (size_t i = h; i < k; i += nth) prime_tables(plans[at + i], td[i], tables.data() + table_off[i]);
            });
        for (auto& th : ths) th.join();
        int rc = ensure(e, &e->arena, &e->arena_cap, bytes, "the 62-bit tables");
        if (rc) return rc;
        const size_t small = sizeof(Task62) * k + sizeof(QueryDesc) * qd.size() + sizeof(u64) * qd.size() +
                             sizeof(int) * slots.size() + 1024;
        rc = ensure(e, &e->small, &e->small_cap, small, "the 62-bit descriptors");
        if (rc) return rc;
        char* sp = (char*)e->small;
        Task62* dtd = (Task62*)sp;
        QueryDesc* dqd = (QueryDesc*)(sp + ((sizeof(Task62) * k + 255) & ~(size_t)255));
        u64* dres = (u64*)((char*)dqd + ((sizeof(QueryDesc) * qd.size() + 255) & ~(size_t)255));
        int* dslots = (int*)((char*)dres + ((sizeof(u64) * qd.size() + 255) & ~(size_t)255));
        u64* mem = (u64*)e->arena;
        for (size_t i = 0; i < k; ++i) {
            const size_t w = table_words62(plans[at + i].F1, plans[at + i].lg);
            HIP_TRY(e, hipMemcpyAsync(mem + td[i].arena, tables.data() + table_off[i], 8 * w, hipMemcpyHostToDevice, e->stream),
                    "upload");
        }
        HIP_TRY(e, hipMemcpyAsync(dtd, td.data(), sizeof(Task62) * k, hipMemcpyHostToDevice, e->stream), "upload");
        HIP_TRY(e, hipMemcpyAsync(dqd, qd.data(), sizeof(QueryDesc) * qd.size(), hipMemcpyHostToDevice, e->stream), "upload");
        HIP_TRY(e, hipMemcpyAsync(dslots, slots.data(), sizeof(int) * slots.size(), hipMemcpyHostToDevice, e->stream),
                "upload");
        HIP_TRY(e, hipEventRecord(e->ev0, e->stream), "hipEventRecord");
        HIP_TRY(e, launch_p62(plans[at].lg, (int)k, dtd, dqd, dslots, mem, dres, e->stream), "62-bit launch");
        HIP_TRY(e, hipEventRecord(e->ev1, e->stream), "hipEventRecord");
        out.assign(qd.size(), 0);
        HIP_TRY(e, hipMemcpyAsync(out.data(), dres, sizeof(u64) * qd.size(), hipMemcpyDeviceToHost, e->stream), "download");
        HIP_TRY(e, hipStreamSynchronize(e->stream), "62-bit execution");
        float ms = 0;
        HIP_TRY(e, hipEventElapsedTime(&ms, e->ev0, e->ev1), "hipEventElapsedTime");
        kernel_ms += ms;
        ++launches;
        size_t s = 0;
        for (size_t i = 0; i < k; ++i)
            for (const QueryDesc& qq : plans[at + i].q) res[qq.out] = out[s++];
        at = end;
    }
    e->stats = keep;
    e->stats.kernel_ms += kernel_ms;
    e->stats.launches += launches;
    e->stats.tasks += (int)plans.size();
    return CF_OK;
}

}  // namespace cf_int
