// cf_serve_host.cpp -- resident table sets, batched queries and the serving loop (C ABI).
//
// A table set holds, for every prime of a list, the rows a_v[w] of its one-table suffix table
// (v <= nmax) and its factorial tables, built once by the task kernel in retain mode. Queries
// N_bad(n, M) with n <= nmax, M <= Mmax are then answered by the batched row kernel
// (cf_serve.hip). The server runs a request loop over such a set: requests that have arrived
// are batched into the next free slot (several slots in flight on their own streams, so batches
// overlap on the device), and each finished query is combined by CRT on host worker threads
// into log2 N_bad and log2 p_fail.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "cf_engine.hpp"
#include "cf_kernels.hpp"
#include "cf_serve.hpp"

using namespace cf;
using namespace cf_int;

struct cf_tables {
    cf_engine* e = nullptr;  // must outlive the set (queries report through it)
    int device = 0;
    int d = 0, model = 0, nmax = 0, Mmax = 0, F1 = 0, v0 = 0;
    std::vector<uint64_t> primes;
    double* store = nullptr;  // np x stride doubles
    int64_t stride = 0;
    ServePrime* dprimes = nullptr;
    hipStream_t stream = nullptr;
    // scratch of cf_query, grown on demand
    void* scratch = nullptr;
    size_t scratch_cap = 0;
};

namespace {

// need(n, M) for a table set: the reference's count (fast/cuckoo_fast.cu:452-459) for its 50-bit
// primes; for the replacement model log2 Q = d log2 M in the same formula.
long need_of(const cf_tables* t, int n, int M) {
    if (!t->model) return need(n, M, t->d, 50);
    return (long)std::ceil(((long double)n * log2_q_replacement(M, t->d) + 3) / 49) + 1;
}

// Range groups (workgroups) per prime and pass: enough workgroups for the device, at least 8
// rows per wave.
int groups_for(int np, int nq) {
    const int waves = kServeThreads / 64;
    int y = std::max(1, (512 + np - 1) / std::max(np, 1));
    return std::min(y, std::max(1, nq / (8 * waves)));
}

// The batch sorted by n (largest first) into passes of kServeJ; per pass the rows 1 .. max n in
// ranges of about equal work (row v costs v entries for every query of the pass with n >= v).
struct Batch {
    std::vector<ServeQuery> q;  // sorted
    std::vector<size_t> order;  // q[k] is the caller's query order[k]
    std::vector<int> bounds;    // passes x (ranges + 1)
    int passes = 0, groups = 0, ranges = 0;
};

void plan_batch(int np, int max_groups, Batch& bt) {
    const size_t J = bt.q.size();
    std::vector<size_t> idx(J);
    for (size_t k = 0; k < J; ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return bt.q[x].n > bt.q[y].n; });
    std::vector<ServeQuery> sq(J);
    std::vector<size_t> ord(J);
    for (size_t k = 0; k < J; ++k) sq[k] = bt.q[idx[k]], ord[k] = bt.order.empty() ? idx[k] : bt.order[idx[k]];
    bt.q.swap(sq);
    bt.order.swap(ord);
    bt.passes = (int)((J + kServeJ - 1) / kServeJ);
    bt.groups = std::min(max_groups, groups_for(np, J ? bt.q[0].n : 1));
    bt.ranges = bt.groups * (kServeThreads / 64);
    const int R = bt.ranges;
    bt.bounds.assign((size_t)bt.passes * (R + 1), 0);
    std::vector<double> cum;
    for (int z = 0; z < bt.passes; ++z) {
        const size_t j0 = (size_t)z * kServeJ, j1 = std::min(J, j0 + kServeJ);
        const int nq = bt.q[j0].n;  // the pass's largest n
        // work of row v: v entries for each query of the pass with n >= v
        cum.assign(nq + 1, 0.0);
        size_t act = j1 - j0;
        for (int v = 1; v <= nq; ++v) {
            while (act > 0 && bt.q[j0 + act - 1].n < v) --act;
            cum[v] = cum[v - 1] + (double)v * (double)act;
        }
        int* b = bt.bounds.data() + (size_t)z * (R + 1);
        int v = 0;
        for (int r = 1; r < R; ++r) {
            const double target = cum[nq] * r / R;
            while (v < nq && cum[v] < target) ++v;
            b[r] = v;
        }
        b[R] = nq;
        for (int r = 1; r <= R; ++r) b[r] = std::max(b[r], b[r - 1]);
    }
}

ServeArgs args_of(const cf_tables* t) {
    ServeArgs a{};
    a.store = t->store;
    a.stride = t->stride;
    a.primes = t->dprimes;
    a.np = (int)t->primes.size();
    a.F1 = t->F1;
    a.dd = t->d;
    a.model = t->model;
    a.v0 = t->v0;
    return a;
}

size_t part_bytes(size_t np, int R, size_t J) { return sizeof(double) * np * (size_t)R * J; }
size_t bounds_count(int R, size_t J) { return ((J + kServeJ - 1) / kServeJ) * (size_t)(R + 1); }

}  // namespace

// ---- host CRT (Garner) over a prefix of the table set's primes ----
// x = sum_i c_i P_i with P_i = p_0 ... p_{i-1} and digits c_i < p_i; row i of the table holds
// P_j mod p_i for j < i, so each digit is one 128-bit dot product and one reduction.
struct Crt {
    std::vector<uint64_t> p;
    std::vector<uint64_t> tab;      // row i at i (i - 1) / 2
    std::vector<uint64_t> invP;     // (p_0 ... p_{i-1})^-1 mod p_i
    std::vector<long double> lp;    // sum_{j<i} log2 p_j
    explicit Crt(const std::vector<uint64_t>& primes) : p(primes) {
        const size_t k = p.size();
        tab.resize(k * (k - (k ? 1 : 0)) / 2 + 1);
        invP.resize(k);
        lp.resize(k + 1);
        lp[0] = 0;
        for (size_t i = 0; i < k; ++i) {
            uint64_t acc = 1 % p[i];
            uint64_t* row = tab.data() + i * (i - (i ? 1 : 0)) / 2;
            for (size_t j = 0; j < i; ++j) {
                row[j] = acc;
                acc = mulmod_u64(acc, p[j] % p[i], p[i]);
            }
            invP[i] = powmod_u64(acc, p[i] - 2, p[i]);
            lp[i + 1] = lp[i] + std::log2((long double)p[i]);
        }
    }
    // log2 of the x with x = r_i mod p_i (i < k); -inf for x = 0.
    double log2x(const uint64_t* r, size_t k) const {
        double out;
        log2x_many(&r, &k, 1, &out);
        return out;
    }
    // The same for cnt residue vectors at once, digit by digit in lockstep: each table row is
    // read once for all of them (the row stays in L1 while the cnt dot products use it), so a
    // group costs one pass over the table instead of cnt.
    void log2x_many(const uint64_t* const* r, const size_t* k, size_t cnt, double* out) const {
        size_t kmax = 0;
        std::vector<std::vector<uint64_t>> c(cnt);
        for (size_t q = 0; q < cnt; ++q) c[q].assign(k[q], 0), kmax = std::max(kmax, k[q]);
        for (size_t i = 0; i < kmax; ++i) {
            const uint64_t* row = tab.data() + i * (i - (i ? 1 : 0)) / 2;
            for (size_t q = 0; q < cnt; ++q) {
                if (i >= k[q]) continue;
                const uint64_t* cq = c[q].data();
                unsigned __int128 s = 0;
                for (size_t j = 0; j < i; ++j) s += (unsigned __int128)cq[j] * row[j];
                const uint64_t pre = (uint64_t)(s % p[i]);
                const uint64_t t = (r[q][i] % p[i] + p[i] - pre) % p[i];
                c[q][i] = mulmod_u64(t, invP[i], p[i]);
            }
        }
        for (size_t q = 0; q < cnt; ++q) {
            size_t top = k[q];
            while (top > 0 && c[q][top - 1] == 0) --top;
            if (!top) {
                out[q] = -INFINITY;
                continue;
            }
            const size_t i = top - 1;
            long double lead = (long double)c[q][i];
            if (i >= 1) lead += (long double)c[q][i - 1] / (long double)p[i - 1];
            if (i >= 2) lead += (long double)c[q][i - 2] / ((long double)p[i - 1] * (long double)p[i - 2]);
            out[q] = (double)(lp[i] + std::log2(lead));
        }
    }
};

// ---- the serving loop ----
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    ServeQuery* hq = nullptr;   // pinned: the batch
    int* hb = nullptr;          // pinned: row bounds
    uint64_t* hres = nullptr;   // pinned: results
    ServeQuery* dq = nullptr;
    int* db = nullptr;
    double* dpart = nullptr;
    uint64_t* dres = nullptr;
    std::vector<size_t> ids;    // the requests of the batch in flight
    bool busy = false;
    double launched = 0;
};

struct cf_server {
    cf_tables* t = nullptr;
    int max_batch = 0, crt_threads = 1;
    std::vector<Slot> slots;
    std::unique_ptr<Crt> crt;
    int max_groups = 0, max_ranges = 0;
};

namespace {

int free_slot(cf_server* s, Slot& sl) {
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.hq) (void)hipHostFree(sl.hq);
    if (sl.hb) (void)hipHostFree(sl.hb);
    if (sl.hres) (void)hipHostFree(sl.hres);
    if (sl.dq) (void)hipFree(sl.dq);
    if (sl.db) (void)hipFree(sl.db);
    if (sl.dpart) (void)hipFree(sl.dpart);
    if (sl.dres) (void)hipFree(sl.dres);
    sl = Slot{};
    (void)s;
    return CF_OK;
}

}  // namespace

extern "C" {

int cf_tables_build(cf_engine* e, int d, int model, int nmax, int Mmax, const uint64_t* primes, size_t np,
                    cf_tables** out) {
    if (!e || !out) return CF_E_ARGS;
    *out = nullptr;
    e->err.clear();
    e->stats = cf_stats{};
    const auto t_start = std::chrono::steady_clock::now();
    if (d < 1 || (model != 0 && model != 1) || nmax < 1 || Mmax <= nmax || (np && !primes) || !np)
        return fail(e, CF_E_ARGS, "bad table set arguments");
    if (std::max(nmax, Mmax) + 1 > kServeMaxF1)
        return fail(e, CF_E_UNSUPPORTED, "max(nmax, Mmax) = " + std::to_string(std::max(nmax, Mmax)) +
                                             " exceeds the query kernel's LDS table of 1/i (at most " +
                                             std::to_string(kServeMaxF1 - 1) + ")");
    HIP_TRY(e, hipSetDevice(e->device), "hipSetDevice");
    auto t = std::make_unique<cf_tables>();
    t->e = e;
    t->device = e->device;
    t->d = d;
    t->model = model;
    t->nmax = nmax;
    t->Mmax = Mmax;
    t->F1 = std::max(nmax, Mmax) + 1;
    t->v0 = model ? 1 : d;
    t->primes.assign(primes, primes + np);
    t->stride = (keep_doubles(t->F1, nmax) + 31) & ~(int64_t)31;
    std::vector<TaskPlan> tasks;
    for (size_t b = 0; b < np; ++b) {
        const uint64_t p = primes[b];
        if (p < 3 || !(p & 1) || p >= (uint64_t(1) << 50) || (uint64_t)Mmax >= p)
            return fail(e, CF_E_ARGS, "prime " + std::to_string(p) + " is not an odd modulus in (Mmax, 2^50)");
        TaskPlan tp;
        tp.p = p;
        tp.dd = d;
        tp.model = model;
        tp.nm = nmax;
        tp.F1 = t->F1;
        tp.lg = engine_log2_length(nmax, p);
        if (tp.lg < 0 || tp.lg > kMaxLog2L)
            return fail(e, CF_E_UNSUPPORTED, "n = " + std::to_string(nmax) + " modulo " + std::to_string(p) +
                                                 " needs a transform this build does not run");
        tp.bytes = task_bytes(plan_for(e, nmax), t->F1, 0, tp.lg);
        tp.cost = 1.0;
        tasks.push_back(std::move(tp));
    }
    // admission: the resident store must fit beside the largest build launch's arena
    const size_t store_bytes = sizeof(double) * (size_t)t->stride * np;
    if (store_bytes + tasks[0].bytes > usable_bytes(e))
        return fail(e, CF_E_NOMEM, "the table set needs " + std::to_string(store_bytes) + " bytes of device memory, " +
                                       std::to_string(usable_bytes(e)) + " usable");
    hipError_t st = hipMalloc(&t->store, store_bytes);
    if (st != hipSuccess) {
        (void)hipGetLastError();
        return fail(e, CF_E_NOMEM, "cannot allocate the table set (" + std::to_string(store_bytes) + " bytes)");
    }
    std::vector<ServePrime> hp(np);
    for (size_t b = 0; b < np; ++b) {
        tasks[b].keep = t->store + t->stride * (int64_t)b;
        hp[b] = ServePrime{(double)primes[b], 1.0 / (double)primes[b]};
    }
    HIP_TRY(e, hipMalloc(&t->dprimes, sizeof(ServePrime) * np), "hipMalloc");
    HIP_TRY(e, hipMemcpy(t->dprimes, hp.data(), sizeof(ServePrime) * np, hipMemcpyHostToDevice), "upload");
    HIP_TRY(e, hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking), "hipStreamCreate");
    const int rc = execute(e, tasks, 0, nullptr, false);
    if (rc) {
        cf_tables_free(t.release());
        return rc;
    }
    e->stats.device_bytes += store_bytes;
    e->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    *out = t.release();
    return CF_OK;
}

void cf_tables_free(cf_tables* t) {
    if (!t) return;
    (void)hipSetDevice(t->device);
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    if (t->store) (void)hipFree(t->store);
    if (t->dprimes) (void)hipFree(t->dprimes);
    if (t->scratch) (void)hipFree(t->scratch);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
}

size_t cf_tables_bytes(const cf_tables* t) { return t ? sizeof(double) * (size_t)t->stride * t->primes.size() : 0; }

long cf_tables_need(const cf_tables* t, int n, int M) { return t ? need_of(t, n, M) : -1; }

}  // extern "C"

namespace {

// The per-device lock of cf_query_noalloc: calls on one device share its fixed scratch, so calls
// from several threads of this process run one after the other.
std::mutex g_noalloc_mu[kQueryScratchDevices];

// cf_query and cf_query_noalloc: validation, the batch plan, the launch and the result order are
// the same; `fixed` takes the device scratch from the static array (and caps the range groups so
// the partial sums fit it) instead of the table set's grown buffer.
int run_query(cf_tables* t, const int32_t* n, const int32_t* M, size_t J, const int64_t* need, uint64_t* res,
              bool fixed) {
    if (!t) return CF_E_ARGS;
    cf_engine* e = t->e;
    e->err.clear();
    e->stats = cf_stats{};
    const auto t_start = std::chrono::steady_clock::now();
    const size_t np = t->primes.size();
    if (J && (!n || !M || !res)) return fail(e, CF_E_ARGS, "bad arguments");
    Batch bt;
    bt.q.resize(J);
    for (size_t j = 0; j < J; ++j) {
        if (n[j] < 1 || M[j] <= n[j] || n[j] > t->nmax || M[j] > t->Mmax)
            return fail(e, CF_E_ARGS, "query " + std::to_string(n[j]) + " " + std::to_string(M[j]) +
                                          " outside the table set (n <= " + std::to_string(t->nmax) +
                                          ", M <= " + std::to_string(t->Mmax) + ")");
        const int64_t k = need ? need[j] : (int64_t)np;
        if (k < 0 || k > (int64_t)np) return fail(e, CF_E_ARGS, "need outside the table set's primes");
        bt.q[j] = ServeQuery{n[j], M[j], (int)k, 0};
    }
    if (!J) return CF_OK;
    HIP_TRY(e, hipSetDevice(e->device), "hipSetDevice");
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    // scratch layout: queries, bounds, partial sums, results
    auto layout_bytes = [&](int groups) {
        const int R = groups * (kServeThreads / 64);
        return up(sizeof(ServeQuery) * J) + up(sizeof(int) * bounds_count(R, J)) + up(part_bytes(np, R, J)) +
               sizeof(uint64_t) * np * J;
    };
    int max_groups = 1 << 20;
    if (fixed) {
        max_groups = groups_for((int)np, t->nmax);
        while (max_groups > 1 && layout_bytes(max_groups) > kQueryScratchBytes) --max_groups;
        if (layout_bytes(max_groups) > kQueryScratchBytes)
            return fail(e, CF_E_ARGS, std::to_string(J) + " queries over " + std::to_string(np) + " primes need " +
                                          std::to_string(layout_bytes(1)) + " bytes of scratch; the fixed scratch holds " +
                                          std::to_string(kQueryScratchBytes));
    }
    plan_batch((int)np, max_groups, bt);
    const int R = bt.ranges;
    const size_t o_q = 0, o_b = up(sizeof(ServeQuery) * J), o_p = o_b + up(sizeof(int) * bt.bounds.size());
    const size_t o_r = o_p + up(part_bytes(np, R, J)), total = o_r + sizeof(uint64_t) * np * J;
    std::unique_lock<std::mutex> lk;
    char* sc = nullptr;
    if (fixed) {
        sc = (char*)query_scratch(e->device);
        if (!sc) return fail(e, CF_E_DEVICE, "the fixed query scratch was not attached on this device");
        if (total > kQueryScratchBytes) return fail(e, CF_E_ARGS, "query scratch layout exceeds the fixed scratch");
        lk = std::unique_lock<std::mutex>(g_noalloc_mu[e->device]);
    } else {
        const int rc = ensure(e, &t->scratch, &t->scratch_cap, total, "the query scratch");
        if (rc) return rc;
        sc = (char*)t->scratch;
    }
    HIP_TRY(e, hipMemcpyAsync(sc + o_q, bt.q.data(), sizeof(ServeQuery) * J, hipMemcpyHostToDevice, t->stream), "upload");
    HIP_TRY(e, hipMemcpyAsync(sc + o_b, bt.bounds.data(), sizeof(int) * bt.bounds.size(), hipMemcpyHostToDevice, t->stream),
            "upload");
    ServeArgs a = args_of(t);
    a.q = (const ServeQuery*)(sc + o_q);
    a.J = (int)J;
    a.bounds = (const int*)(sc + o_b);
    a.ranges = R;
    a.passes = bt.passes;
    a.groups = bt.groups;
    a.part = (double*)(sc + o_p);
    std::vector<uint64_t> sorted(np * J);
    HIP_TRY(e, hipEventRecord(e->ev0, t->stream), "hipEventRecord");
    HIP_TRY(e, launch_serve(a, (uint64_t*)(sc + o_r), t->stream), "query launch");
    HIP_TRY(e, hipEventRecord(e->ev1, t->stream), "hipEventRecord");
    HIP_TRY(e, hipMemcpyAsync(sorted.data(), sc + o_r, sizeof(uint64_t) * np * J, hipMemcpyDeviceToHost, t->stream), "download");
    HIP_TRY(e, hipStreamSynchronize(t->stream), "query execution");
    if (lk.owns_lock()) lk.unlock();
    for (size_t b = 0; b < np; ++b)
        for (size_t k = 0; k < J; ++k) res[b * J + bt.order[k]] = sorted[b * J + k];
    float ms = 0;
    HIP_TRY(e, hipEventElapsedTime(&ms, e->ev0, e->ev1), "hipEventElapsedTime");
    e->stats.kernel_ms = ms;
    e->stats.launches = 1;
    e->stats.tasks = (int)np;
    e->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return CF_OK;
}

}  // namespace

extern "C" {

int cf_query(cf_tables* t, const int32_t* n, const int32_t* M, size_t J, const int64_t* need, uint64_t* res) {
    return run_query(t, n, M, J, need, res, false);
}

int cf_query_noalloc(cf_tables* t, const int32_t* n, const int32_t* M, size_t J, const int64_t* need, uint64_t* res) {
    return run_query(t, n, M, J, need, res, true);
}

double cf_crt_log2(const uint64_t* primes, const uint64_t* res, size_t k) {
    if (!k || !primes || !res) return -INFINITY;
    const Crt c(std::vector<uint64_t>(primes, primes + k));
    return c.log2x(res, k);
}

int cf_server_open(cf_tables* t, int slots, int max_batch, int crt_threads, cf_server** out) {
    if (!t || !out || slots < 1 || max_batch < 1 || crt_threads < 1) return CF_E_ARGS;
    *out = nullptr;
    cf_engine* e = t->e;
    e->err.clear();
    HIP_TRY(e, hipSetDevice(e->device), "hipSetDevice");
    auto s = std::make_unique<cf_server>();
    s->t = t;
    s->max_batch = max_batch;
    s->crt_threads = crt_threads;
    s->crt = std::make_unique<Crt>(t->primes);
    const size_t np = t->primes.size();
    s->max_groups = groups_for((int)np, t->nmax);
    s->max_ranges = s->max_groups * (kServeThreads / 64);
    // admission of the batch slots by free memory: as many as fit, at least one
    const size_t per = sizeof(ServeQuery) * max_batch + sizeof(int) * bounds_count(s->max_ranges, max_batch) +
                       part_bytes(np, s->max_ranges, max_batch) + sizeof(uint64_t) * np * max_batch + 4096;
    const size_t fit = usable_bytes(e) > e->arena_cap ? (usable_bytes(e) - e->arena_cap) / per : 0;
    const int ns = (int)std::min<size_t>(slots, fit);
    if (ns < 1) return fail(e, CF_E_NOMEM, "no device memory for a batch slot");
    s->slots.resize(ns);
    for (Slot& sl : s->slots) {
        HIP_TRY(e, hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreate");
        HIP_TRY(e, hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
        HIP_TRY(e, hipHostMalloc((void**)&sl.hq, sizeof(ServeQuery) * max_batch, hipHostMallocDefault), "hipHostMalloc");
        HIP_TRY(e, hipHostMalloc((void**)&sl.hb, sizeof(int) * bounds_count(s->max_ranges, max_batch), hipHostMallocDefault),
                "hipHostMalloc");
        HIP_TRY(e, hipHostMalloc((void**)&sl.hres, sizeof(uint64_t) * np * max_batch, hipHostMallocDefault), "hipHostMalloc");
        HIP_TRY(e, hipMalloc((void**)&sl.dq, sizeof(ServeQuery) * max_batch), "hipMalloc");
        HIP_TRY(e, hipMalloc((void**)&sl.db, sizeof(int) * bounds_count(s->max_ranges, max_batch)), "hipMalloc");
        HIP_TRY(e, hipMalloc((void**)&sl.dpart, part_bytes(np, s->max_ranges, max_batch)), "hipMalloc");
        HIP_TRY(e, hipMalloc((void**)&sl.dres, sizeof(uint64_t) * np * max_batch), "hipMalloc");
    }
    *out = s.release();
    return CF_OK;
}

void cf_server_close(cf_server* s) {
    if (!s) return;
    (void)hipSetDevice(s->t->device);
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        free_slot(s, sl);
    }
    delete s;
}

int cf_server_slots(const cf_server* s) { return s ? (int)s->slots.size() : 0; }

int cf_server_replay(cf_server* s, size_t count, const int32_t* n, const int32_t* M, const double* arrival_s,
                     double* latency_ms, double* log2_pfail, uint64_t* res, cf_serve_stats* stats) {
    if (!s) return CF_E_ARGS;
    cf_tables* t = s->t;
    cf_engine* e = t->e;
    e->err.clear();
    if (count && (!n || !M || !arrival_s)) return fail(e, CF_E_ARGS, "bad arguments");
    const size_t np = t->primes.size();
    std::vector<long> nd(count);
    for (size_t i = 0; i < count; ++i) {
        if (n[i] < 1 || M[i] <= n[i] || n[i] > t->nmax || M[i] > t->Mmax)
            return fail(e, CF_E_ARGS, "request " + std::to_string(i) + " outside the table set");
        nd[i] = need_of(t, n[i], M[i]);
        if (nd[i] > (long)np) return fail(e, CF_E_PRIMES, "request " + std::to_string(i) + " needs " +
                                                          std::to_string(nd[i]) + " primes, the set has " + std::to_string(np));
        if (i && arrival_s[i] < arrival_s[i - 1]) return fail(e, CF_E_ARGS, "arrivals must be in order");
    }
    HIP_TRY(e, hipSetDevice(e->device), "hipSetDevice");
    std::vector<uint64_t> resid(count * np);  // per request, its residues (b < need)
    std::vector<double> done_at(count, 0.0), l2(count, 0.0);

    // CRT workers
    std::mutex mu;
    std::condition_variable cv;
    std::deque<size_t> crt_q;
    bool closing = false;
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto now_s = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
    // each worker takes up to kCrtGroup finished requests at a time and combines them together
    constexpr size_t kCrtGroup = 8;
    auto worker = [&] {
        std::vector<size_t> ids;
        std::vector<const uint64_t*> rs;
        std::vector<size_t> ks;
        std::vector<double> lx;
        for (;;) {
            ids.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return closing || !crt_q.empty(); });
                if (crt_q.empty()) return;
                while (!crt_q.empty() && ids.size() < kCrtGroup) {
                    ids.push_back(crt_q.front());
                    crt_q.pop_front();
                }
            }
            rs.clear(), ks.clear();
            for (size_t i : ids) rs.push_back(resid.data() + i * np), ks.push_back((size_t)nd[i]);
            lx.resize(ids.size());
            s->crt->log2x_many(rs.data(), ks.data(), ids.size(), lx.data());
            const double tn = now_s();
            for (size_t g = 0; g < ids.size(); ++g) {
                const size_t i = ids[g];
                const double lq = t->model ? log2_q_replacement(M[i], t->d) : log2_q(M[i], t->d);
                l2[i] = lx[g] - n[i] * lq;
                done_at[i] = tn;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int k = 0; k < s->crt_threads; ++k) pool.emplace_back(worker);

    size_t next = 0, finished_gpu = 0;
    std::deque<size_t> waiting;
    long batches = 0;
    double batched = 0;
    int rc = CF_OK;
    while (finished_gpu < count) {
        const double tn = now_s();
        while (next < count && arrival_s[next] <= tn) waiting.push_back(next++);
        bool progress = false;
        // completions
        for (Slot& sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                rc = hip_fail(e, q, "query batch");
                break;
            }
            const size_t J = sl.ids.size();
            for (size_t jj = 0; jj < J; ++jj) {
                const size_t i = sl.ids[jj];
                for (size_t b = 0; b < np; ++b) resid[i * np + b] = sl.hres[b * J + jj];
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                for (size_t i : sl.ids) crt_q.push_back(i);
            }
            cv.notify_all();
            finished_gpu += J;
            sl.busy = false;
            sl.ids.clear();
            progress = true;
        }
        if (rc) break;
        // admissions: every free slot takes the requests waiting now (up to max_batch)
        for (Slot& sl : s->slots) {
            if (sl.busy || waiting.empty()) continue;
            const size_t J = std::min(waiting.size(), (size_t)s->max_batch);
            Batch bt;
            for (size_t jj = 0; jj < J; ++jj) {
                const size_t i = waiting.front();
                waiting.pop_front();
                bt.q.push_back(ServeQuery{n[i], M[i], (int)nd[i], 0});
                bt.order.push_back(i);
            }
            plan_batch((int)np, s->max_groups, bt);
            sl.ids = bt.order;
            std::copy(bt.q.begin(), bt.q.end(), sl.hq);
            std::copy(bt.bounds.begin(), bt.bounds.end(), sl.hb);
            const int R = bt.ranges;
            ServeArgs a = args_of(t);
            a.q = sl.dq;
            a.J = (int)J;
            a.bounds = sl.db;
            a.ranges = R;
            a.passes = bt.passes;
            a.groups = bt.groups;
            a.part = sl.dpart;
            hipError_t st = hipMemcpyAsync(sl.dq, sl.hq, sizeof(ServeQuery) * J, hipMemcpyHostToDevice, sl.stream);
            if (st == hipSuccess)
                st = hipMemcpyAsync(sl.db, sl.hb, sizeof(int) * bt.bounds.size(), hipMemcpyHostToDevice, sl.stream);
            if (st == hipSuccess) st = launch_serve(a, sl.dres, sl.stream);
            if (st == hipSuccess)
                st = hipMemcpyAsync(sl.hres, sl.dres, sizeof(uint64_t) * np * J, hipMemcpyDeviceToHost, sl.stream);
            if (st == hipSuccess) st = hipEventRecord(sl.done, sl.stream);
            if (st != hipSuccess) {
                rc = hip_fail(e, st, "query batch launch");
                break;
            }
            sl.busy = true;
            sl.launched = tn;
            ++batches;
            batched += (double)J;
            progress = true;
        }
        if (rc) break;
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        closing = true;
    }
    cv.notify_all();
    for (std::thread& th : pool) th.join();
    for (Slot& sl : s->slots)
        if (sl.busy) {
            (void)hipStreamSynchronize(sl.stream);
            sl.busy = false;
            sl.ids.clear();
        }
    if (rc) return rc;
    std::vector<double> lat(count);
    double last = 0;
    for (size_t i = 0; i < count; ++i) {
        lat[i] = (done_at[i] - arrival_s[i]) * 1e3;
        last = std::max(last, done_at[i]);
        if (latency_ms) latency_ms[i] = lat[i];
        if (log2_pfail) log2_pfail[i] = l2[i];
    }
    if (res) std::copy(resid.begin(), resid.end(), res);
    if (stats) {
        *stats = cf_serve_stats{};
        std::vector<double> sl = lat;
        std::sort(sl.begin(), sl.end());
        auto pct = [&](double f) { return sl.empty() ? 0.0 : sl[std::min(sl.size() - 1, (size_t)(f * (sl.size() - 1) + 0.5))]; };
        stats->wall_s = last - (count ? arrival_s[0] : 0);
        stats->queries_per_s = stats->wall_s > 0 ? count / stats->wall_s : 0;
        stats->p50_ms = pct(0.50);
        stats->p99_ms = pct(0.99);
        stats->batches = batches;
        stats->mean_batch = batches ? batched / batches : 0;
    }
    return CF_OK;
}

}  // extern "C"
