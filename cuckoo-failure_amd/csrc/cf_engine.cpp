// cf_engine.cpp -- the C ABI (include/cuckoo_amd.h): planning, device memory and launches.
//
// A call to cf_run_fused turns every (prime, its masked-in queries) into one task: one
// workgroup that builds the prime's suffix table and answers its queries. Tasks are grouped by
// transform length (one kernel instance each) and packed into launches that fit the device memory,
// which the engine keeps allocated between calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cuckoo_amd.h"
#include "cf_host.hpp"
#include "cf_kernels.hpp"

using namespace cf;

struct cf_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    cf_stats stats{};
    // device buffers, grown on demand and kept between calls
    void* arena = nullptr;
    size_t arena_cap = 0;
    void* small = nullptr;  // tasks, queries, slot tables, results
    size_t small_cap = 0;
    std::map<int, SlotPlan> plans;  // by nm
};

namespace {

int fail(cf_engine* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    return code;
}

int hip_fail(cf_engine* e, hipError_t st, const char* what) {
    return fail(e, CF_E_DEVICE, std::string("HIP error in ") + what + ": " + hipGetErrorString(st));
}

#define HIP_TRY(e, x, what)                              \
    do {                                                 \
        hipError_t st_ = (x);                            \
        if (st_ != hipSuccess) return hip_fail(e, st_, what); \
    } while (0)

const SlotPlan& plan_for(cf_engine* e, int nm) {
    auto it = e->plans.find(nm);
    if (it == e->plans.end()) it = e->plans.emplace(nm, plan_slots(nm)).first;
    return it->second;
}

int ensure(cf_engine* e, void** buf, size_t* cap, size_t bytes, const char* what) {
    if (bytes <= *cap) return CF_OK;
    if (*buf) {
        HIP_TRY(e, hipFree(*buf), "hipFree");
        *buf = nullptr, *cap = 0;
    }
    hipError_t st = hipMalloc(buf, bytes);
    if (st != hipSuccess) {
        (void)hipGetLastError();
        *buf = nullptr;
        return fail(e, CF_E_NOMEM, std::string("cannot allocate ") + std::to_string(bytes) + " bytes for " + what);
    }
    *cap = bytes;
    return CF_OK;
}

struct TaskPlan {
    size_t b;                  // prime index in the call
    int nm, F1, lg;
    std::vector<int> queries;  // query indices, sorted by n descending
    size_t bytes;
};

// Free device memory that a batch may use: what is free now plus the arena we already hold,
// minus a reserve for the small buffers and the runtime.
size_t usable_bytes(cf_engine* e) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    const size_t have = fr + e->arena_cap;
    const size_t reserve = (size_t)2 << 30;
    return have > reserve ? have - reserve : 0;
}

}  // namespace

extern "C" {

const char* cf_strerror(int code) {
    switch (code) {
        case CF_OK: return "ok";
        case CF_E_DEVICE: return "device error";
        case CF_E_ARGS: return "invalid arguments";
        case CF_E_NOMEM: return "not enough device memory";
        case CF_E_PRIMES: return "prime list exhausted";
        case CF_E_UNSUPPORTED: return "shape not supported by this build";
        default: return "unknown error";
    }
}

const char* cf_last_error(const cf_engine* e) { return e ? e->err.c_str() : "no engine"; }

int cf_last_stats(const cf_engine* e, cf_stats* out) {
    if (!e || !out) return CF_E_ARGS;
    *out = e->stats;
    return CF_OK;
}

int cf_open(int device, cf_engine** out) {
    if (!out) return CF_E_ARGS;
    *out = nullptr;
    auto e = std::make_unique<cf_engine>();
    e->device = device;
    int n = 0;
    HIP_TRY(e.get(), hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) return CF_E_DEVICE;
    HIP_TRY(e.get(), hipSetDevice(device), "hipSetDevice");
    HIP_TRY(e.get(), hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "hipStreamCreate");
    HIP_TRY(e.get(), hipEventCreate(&e->ev0), "hipEventCreate");
    HIP_TRY(e.get(), hipEventCreate(&e->ev1), "hipEventCreate");
    *out = e.release();
    return CF_OK;
}

void cf_close(cf_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->arena) (void)hipFree(e->arena);
    if (e->small) (void)hipFree(e->small);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int cf_device_name(const cf_engine* e, char* buf, size_t len) {
    if (!e || !buf || !len) return CF_E_ARGS;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, e->device) != hipSuccess) return CF_E_DEVICE;
    snprintf(buf, len, "%s %s", pr.name, pr.gcnArchName);
    return CF_OK;
}

size_t cf_mem_per_prime(int nmax, int Mmax, size_t J) {
    if (nmax < 1) return 0;
    int lg = kMinLog2L;
    while ((1 << lg) < nmax) ++lg;
    return task_bytes(plan_slots(nmax), std::max(nmax, Mmax) + 1, (int)J, lg);
}

long cf_max_batch(cf_engine* e, int nmax, int Mmax, size_t J) {
    if (!e || nmax < 1) return 0;
    (void)hipSetDevice(e->device);
    int lg = kMinLog2L;
    while ((1 << lg) < nmax) ++lg;
    const size_t per = task_bytes(plan_for(e, nmax), std::max(nmax, Mmax) + 1, (int)J, lg);
    return (long)(usable_bytes(e) / per);
}

int cf_run_fused(cf_engine* e, int d, int model, const int32_t* n, const int32_t* M, size_t J, const uint64_t* primes,
                 size_t np, const uint8_t* mask, uint64_t* res) {
    if (!e) return CF_E_ARGS;
    const auto t_start = std::chrono::steady_clock::now();
    e->err.clear();
    e->stats = cf_stats{};
    if (d < 1 || (model != 0 && model != 1) || (J && (!n || !M)) || (np && !primes) || (np && J && !res))
        return fail(e, CF_E_ARGS, "bad arguments");
    for (size_t j = 0; j < J; ++j)
        if (n[j] < 1 || M[j] <= n[j]) return fail(e, CF_E_ARGS, "bad query " + std::to_string(n[j]) + " " + std::to_string(M[j]));
    if (np * J) std::fill(res, res + np * J, 0);
    HIP_TRY(e, hipSetDevice(e->device), "hipSetDevice");

    // ---- one task per prime that has any query ----
    std::vector<TaskPlan> tasks;
    for (size_t b = 0; b < np; ++b) {
        const uint64_t p = primes[b];
        TaskPlan tp;
        tp.b = b;
        tp.nm = 0;
        int Mmax = 0;
        for (size_t j = 0; j < J; ++j)
            if (!mask || mask[b * J + j]) {
                tp.queries.push_back((int)j);
                tp.nm = std::max(tp.nm, (int)n[j]);
                Mmax = std::max(Mmax, (int)M[j]);
            }
        if (tp.queries.empty()) continue;
        if (p < 3 || !(p & 1) || p >= (uint64_t(1) << 50) || (uint64_t)Mmax >= p)
            return fail(e, CF_E_ARGS, "prime " + std::to_string(p) + " is not an odd modulus in (M, 2^50)");
        std::stable_sort(tp.queries.begin(), tp.queries.end(), [&](int a, int c) { return n[a] > n[c]; });
        tp.F1 = std::max(tp.nm, Mmax) + 1;
        tp.lg = engine_log2_length(tp.nm, p);
        if (tp.lg < 0)
            return fail(e, CF_E_UNSUPPORTED, "p - 1 of " + std::to_string(p) + " has too few factors of two for n = " +
                                                 std::to_string(tp.nm));
        if (tp.lg > kMaxLog2L)
            return fail(e, CF_E_UNSUPPORTED, "n = " + std::to_string(tp.nm) + " needs a transform longer than 2^" +
                                                 std::to_string(kMaxLog2L) + ", which this build does not run");
        tp.bytes = task_bytes(plan_for(e, tp.nm), tp.F1, (int)tp.queries.size(), tp.lg);
        tasks.push_back(std::move(tp));
    }
    // largest transforms first, then by table size, so a launch holds tasks of one length
    std::stable_sort(tasks.begin(), tasks.end(), [](const TaskPlan& a, const TaskPlan& b) {
        return a.lg != b.lg ? a.lg > b.lg : a.nm > b.nm;
    });

    const size_t budget = usable_bytes(e);
    size_t at = 0;
    float kernel_ms_total = 0;
    while (at < tasks.size()) {
        // ---- one launch: tasks of one length that fit in memory together, as whole multiples of
        //      the workgroups the device keeps resident (a partial extra wave would double the
        //      launch), and balanced over the launches this length class needs ----
        const int lg = tasks[at].lg;
        size_t end = at, bytes = 0;
        while (end < tasks.size() && tasks[end].lg == lg && bytes + tasks[end].bytes <= budget) bytes += tasks[end++].bytes;
        if (end == at)
            return fail(e, CF_E_NOMEM, "one prime needs " + std::to_string(tasks[at].bytes) + " bytes of device memory, " +
                                           std::to_string(budget) + " usable");
        size_t cls_end = end;
        while (cls_end < tasks.size() && tasks[cls_end].lg == lg) ++cls_end;
        const size_t fit = end - at, left = cls_end - at, cap = (size_t)std::max(resident_tasks(lg, e->device), 1);
        size_t take = fit >= cap ? fit / cap * cap : fit;
        if (take < left) {
            const size_t launches = (left + take - 1) / take;
            const size_t even = (left + launches - 1) / launches;
            take = std::min(take, (even + cap - 1) / cap * cap);
        }
        end = at + std::min(take, left);
        bytes = 0;
        for (size_t k = at; k < end; ++k) bytes += tasks[k].bytes;
        const int nt = (int)(end - at);
        std::vector<TaskDesc> td(nt);
        std::vector<QueryDesc> qd;
        std::vector<int> slots;
        std::map<int, int> slot_off;
        int64_t arena_off = 0;
        for (int k = 0; k < nt; ++k) {
            const TaskPlan& tp = tasks[at + k];
            TaskDesc& t = td[k];
            memset(&t, 0, sizeof t);
            task_constants(primes[tp.b], lg, t);
            const SlotPlan& sp = plan_for(e, tp.nm);
            auto so = slot_off.find(tp.nm);
            if (so == slot_off.end()) {
                so = slot_off.emplace(tp.nm, (int)slots.size()).first;
                slots.insert(slots.end(), sp.segs.begin(), sp.segs.end());
            }
            t.arena = arena_off;
            t.nm = tp.nm;
            t.F1 = tp.F1;
            t.q_off = (int)qd.size();
            t.q_cnt = (int)tp.queries.size();
            t.slot_off = so->second;
            t.pool = sp.pool;
            t.NB = sp.NB;
            t.NSEG = sp.NSEG;
            t.NS = sp.NS;
            t.NT = (1 << lg) / 16;
            for (int j : tp.queries) qd.push_back(QueryDesc{n[j], M[j], (int64_t)(tp.b * J + j)});
            arena_off += (int64_t)(tp.bytes / 8);
        }
        // ---- device buffers ----
        int rc = ensure(e, &e->arena, &e->arena_cap, bytes, "the task tables");
        if (rc) return rc;
        const size_t o_td = 0, o_qd = o_td + ((sizeof(TaskDesc) * nt + 255) & ~(size_t)255);
        const size_t o_sl = o_qd + ((sizeof(QueryDesc) * qd.size() + 255) & ~(size_t)255);
        const size_t o_rs = o_sl + ((sizeof(int) * slots.size() + 255) & ~(size_t)255);
        const size_t small = o_rs + sizeof(uint64_t) * std::max<size_t>(np * J, 1);
        rc = ensure(e, &e->small, &e->small_cap, small, "the task descriptors");
        if (rc) return rc;
        char* sm = (char*)e->small;
        HIP_TRY(e, hipMemcpyAsync(sm + o_td, td.data(), sizeof(TaskDesc) * nt, hipMemcpyHostToDevice, e->stream), "upload");
        HIP_TRY(e, hipMemcpyAsync(sm + o_qd, qd.data(), sizeof(QueryDesc) * qd.size(), hipMemcpyHostToDevice, e->stream), "upload");
        HIP_TRY(e, hipMemcpyAsync(sm + o_sl, slots.data(), sizeof(int) * slots.size(), hipMemcpyHostToDevice, e->stream), "upload");
        // ---- launch ----
        HIP_TRY(e, hipEventRecord(e->ev0, e->stream), "hipEventRecord");
        HIP_TRY(e, launch_tasks(lg, nt, (const TaskDesc*)(sm + o_td), (const QueryDesc*)(sm + o_qd), (const int*)(sm + o_sl),
                                (double*)e->arena, (uint64_t*)(sm + o_rs), model, d, e->stream),
                "kernel launch");
        HIP_TRY(e, hipEventRecord(e->ev1, e->stream), "hipEventRecord");
        HIP_TRY(e, hipStreamSynchronize(e->stream), "kernel execution");
        float ms = 0;
        HIP_TRY(e, hipEventElapsedTime(&ms, e->ev0, e->ev1), "hipEventElapsedTime");
        kernel_ms_total += ms;
        // ---- results: the kernel wrote only the masked-in entries of this launch ----
        std::vector<uint64_t> out(np * J);
        HIP_TRY(e, hipMemcpy(out.data(), sm + o_rs, sizeof(uint64_t) * np * J, hipMemcpyDeviceToHost), "download");
        for (const QueryDesc& q : qd) res[q.out] = out[q.out];
        e->stats.launches++;
        e->stats.tasks += nt;
        e->stats.log2_length = lg;
        e->stats.device_bytes = std::max(e->stats.device_bytes, bytes);
        at = end;
    }
    e->stats.batches = e->stats.launches;
    e->stats.kernel_ms = kernel_ms_total;
    e->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return CF_OK;
}

// ---- host logic, no device ----
long cf_prime_range(int prime_bits, long first, long count, uint64_t* out) {
    if (prime_bits < 21 || prime_bits > 62 || first < 0 || count < 0 || (count && !out)) return -1;
    const std::vector<uint64_t> v = prime_range(prime_bits, first, count);
    std::copy(v.begin(), v.end(), out);
    return (long)v.size();
}

long cf_need(int n, int M, int d, int prime_bits) { return need(n, M, d, prime_bits); }

size_t cf_transform_length(size_t nmax) { return transform_length(nmax); }

int cf_ntt_primes32(int count, int min_pow2, uint32_t* out) {
    if (count < 0 || (count && !out)) return -1;
    try {
        const std::vector<uint32_t> v = ntt_primes32(count, min_pow2);
        std::copy(v.begin(), v.end(), out);
        return (int)v.size();
    } catch (...) {
        return -1;
    }
}

int cf_engine_log2_length(int nmax, uint64_t p) {
    const int lg = engine_log2_length(nmax, p);
    return lg > kMaxLog2L ? -1 : lg;
}

}  // extern "C"
