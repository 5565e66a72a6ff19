// cf_engine.cpp -- the C ABI (include/cuckoo_amd.h): planning, device memory and launches.
//
// A call to cf_run_fused turns every (prime, its masked-in queries) into one task: one
// workgroup that builds the prime's suffix table and answers its queries. Tasks are grouped by
// transform length (one kernel instance each) and packed into launches that fit the device memory,
// which the engine keeps allocated between calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cuckoo_amd.h"
#include "cf_host.hpp"
#include "cf_kernels.hpp"
#include "cf_serve.hpp"
#include "cf_engine.hpp"

using namespace cf;

namespace cf_int {

int fail(cf_engine* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    return code;
}

int hip_fail(cf_engine* e, hipError_t st, const char* what) {
    return fail(e, CF_E_DEVICE, std::string("HIP error in ") + what + ": " + hipGetErrorString(st));
}

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

// Free device memory that a batch may use: what is free now plus the arena we already hold,
// minus a reserve for the small buffers and the runtime.
size_t usable_bytes(cf_engine* e) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    const size_t have = fr + e->arena_cap;
    const size_t reserve = (size_t)2 << 30;
    return have > reserve ? have - reserve : 0;
}


// One task: prime p, its queries (nq of them, query(k) giving each), checked and costed.
template <class Q>
int plan_task(cf_engine* e, uint64_t p, int d, int model, size_t nq, Q query, TaskPlan& tp) {
    tp.p = p;
    tp.dd = d;
    tp.model = model;
    tp.nm = 0;
    int Mmax = 0;
    for (size_t k = 0; k < nq; ++k) {
        tp.queries.push_back(query(k));
        tp.nm = std::max(tp.nm, tp.queries.back().n);
        Mmax = std::max(Mmax, tp.queries.back().M);
    }
    if (p < 3 || !(p & 1) || p >= (uint64_t(1) << 50) || (uint64_t)Mmax >= p)
        return fail(e, CF_E_ARGS, "prime " + std::to_string(p) + " is not an odd modulus in (M, 2^50)");
    std::stable_sort(tp.queries.begin(), tp.queries.end(), [](const QueryDesc& a, const QueryDesc& c) { return a.n > c.n; });
    tp.F1 = std::max(tp.nm, Mmax) + 1;
    tp.lg = engine_log2_length(tp.nm, p);
    if (tp.lg < 0)
        return fail(e, CF_E_UNSUPPORTED,
                    "p - 1 of " + std::to_string(p) + " has too few factors of two for n = " + std::to_string(tp.nm));
    if (tp.lg > kWideMaxLog2 || (tp.lg > kMaxLog2L && tp.lg < kWideLog2))
        return fail(e, CF_E_UNSUPPORTED, "n = " + std::to_string(tp.nm) + " needs a transform longer than 2^" +
                                             std::to_string(kWideMaxLog2) + ", which this build does not run");
    tp.bytes = task_bytes(plan_for(e, tp.nm), tp.F1, (int)tp.queries.size(), tp.lg);
    // two convolutions of length L per row, the tile traffic, and each query's row sums
    tp.cost = 2.0 * tp.nm * ((double)(1 << tp.lg) * (tp.lg + 1)) + 8.0 * tp.nm * (double)tp.nm;
    for (const QueryDesc& q : tp.queries) tp.cost += 1.5 * (double)q.n * (double)q.n;
    return CF_OK;
}

// Descriptors of tasks[at, end) (arena offsets from arena0, consecutive) and their queries.
void describe(cf_engine* e, const std::vector<TaskPlan>& tasks, size_t at, size_t end, bool shared_arena,
              std::vector<TaskDesc>& td, std::vector<QueryDesc>& qd, std::vector<int>& slots) {
    std::map<int, int> slot_off;
    int64_t arena_off = 0;
    td.assign(end - at, TaskDesc{});
    qd.clear();
    slots.clear();
    for (size_t k = at; k < end; ++k) {
        const TaskPlan& tp = tasks[k];
        TaskDesc& t = td[k - at];
        memset(&t, 0, sizeof t);
        if (tp.lg >= kWideLog2) wide_constants(tp.p, tp.lg, t);
        else task_constants(tp.p, tp.lg, t);
        const SlotPlan& sp = plan_for(e, tp.nm);
        auto so = slot_off.find(tp.nm);
        if (so == slot_off.end()) {
            so = slot_off.emplace(tp.nm, (int)slots.size()).first;
            slots.insert(slots.end(), sp.segs.begin(), sp.segs.end());
        }
        t.arena = shared_arena ? 0 : arena_off;
        t.nm = tp.nm;
        t.F1 = tp.F1;
        t.q_off = (int)qd.size();
        t.q_cnt = (int)tp.queries.size();
        t.slot_off = so->second;
        t.pool = sp.pool;
        t.NB = sp.NB;
        t.NSEG = sp.NSEG;
        t.NS = sp.NS;
        t.NT = (1 << tp.lg) / 16;
        t.lg = tp.lg;
        t.dd = tp.dd;
        t.model = tp.model;
        t.keep = tp.keep;
        qd.insert(qd.end(), tp.queries.begin(), tp.queries.end());
        arena_off += (int64_t)(tp.bytes / 8);
    }
}

// Uploads descriptors, queries and slot tables (and a zero queue position) into the small
// buffer; returns the device pointers.
struct Uploaded {
    const QueueJob* jobs;
    const TaskDesc* td;
    const QueryDesc* qd;
    const int* slots;
    uint64_t* res;
    unsigned* next;
    unsigned* sync;
};

int upload(cf_engine* e, const std::vector<TaskDesc>& td, const std::vector<QueryDesc>& qd, const std::vector<int>& slots,
           size_t nres, Uploaded& u, const std::vector<QueueJob>* jobs = nullptr, size_t nsync = 0) {
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t njobs = jobs ? jobs->size() : 0;
    const size_t o_jb = 0, o_td = o_jb + up(sizeof(QueueJob) * njobs), o_qd = o_td + up(sizeof(TaskDesc) * td.size());
    const size_t o_sl = o_qd + up(sizeof(QueryDesc) * qd.size());
    const size_t o_nx = o_sl + up(sizeof(int) * slots.size());
    const size_t o_sy = o_nx + 256;
    const size_t o_rs = o_sy + up(sizeof(unsigned) * nsync);
    const size_t small = o_rs + sizeof(uint64_t) * std::max<size_t>(nres, 1);
    int rc = ensure(e, &e->small, &e->small_cap, small, "the task descriptors");
    if (rc) return rc;
    char* sm = (char*)e->small;
    static const unsigned zero = 0;
    if (njobs)
        HIP_TRY(e, hipMemcpyAsync(sm + o_jb, jobs->data(), sizeof(QueueJob) * njobs, hipMemcpyHostToDevice, e->stream), "upload");
    HIP_TRY(e, hipMemcpyAsync(sm + o_td, td.data(), sizeof(TaskDesc) * td.size(), hipMemcpyHostToDevice, e->stream), "upload");
    HIP_TRY(e, hipMemcpyAsync(sm + o_qd, qd.data(), sizeof(QueryDesc) * qd.size(), hipMemcpyHostToDevice, e->stream), "upload");
    HIP_TRY(e, hipMemcpyAsync(sm + o_sl, slots.data(), sizeof(int) * slots.size(), hipMemcpyHostToDevice, e->stream), "upload");
    HIP_TRY(e, hipMemcpyAsync(sm + o_nx, &zero, sizeof zero, hipMemcpyHostToDevice, e->stream), "upload");
    if (nsync) HIP_TRY(e, hipMemsetAsync(sm + o_sy, 0, sizeof(unsigned) * nsync, e->stream), "hipMemsetAsync");
    u.sync = (unsigned*)(sm + o_sy);
    u.jobs = (const QueueJob*)(sm + o_jb);
    u.td = (const TaskDesc*)(sm + o_td);
    u.qd = (const QueryDesc*)(sm + o_qd);
    u.slots = (const int*)(sm + o_sl);
    u.next = (unsigned*)(sm + o_nx);
    u.res = (uint64_t*)(sm + o_rs);
    return CF_OK;
}

// Times one launch (HIP events on the engine's stream), then copies back the results of its
// queries.
template <class L>
int timed_launch(cf_engine* e, L launch, const std::vector<QueryDesc>& qd, const Uploaded& u, size_t nres, uint64_t* res,
                 float& ms) {
    HIP_TRY(e, hipEventRecord(e->ev0, e->stream), "hipEventRecord");
    HIP_TRY(e, launch(), "kernel launch");
    HIP_TRY(e, hipEventRecord(e->ev1, e->stream), "hipEventRecord");
    HIP_TRY(e, hipStreamSynchronize(e->stream), "kernel execution");
    HIP_TRY(e, hipEventElapsedTime(&ms, e->ev0, e->ev1), "hipEventElapsedTime");
    std::vector<uint64_t> out(nres);
    HIP_TRY(e, hipMemcpy(out.data(), u.res, sizeof(uint64_t) * nres, hipMemcpyDeviceToHost), "download");
    for (const QueryDesc& q : qd) res[q.out] = out[q.out];
    return CF_OK;
}

// ---- graph mode: the call's launches recorded, then run as one captured hipGraph ----
struct Rec {
    int kind;  // 0: k_tasks<lg>, 1: the work-queue kernel
    int lg, grid, njobs, variant;
    int64_t stride;  // queue: arena doubles per workgroup
    size_t o_jb, o_td, o_qd, o_sl, o_nx;
};

struct Recorder {
    std::vector<char> stage;  // host image of every launch's descriptors, 256-byte aligned regions
    std::vector<Rec> recs;
    std::vector<QueryDesc> outs;
    size_t put(const void* p, size_t n) {
        const size_t o = (stage.size() + 255) & ~(size_t)255;
        stage.resize(o + std::max<size_t>(n, 1), 0);
        if (n) memcpy(stage.data() + o, p, n);
        return o;
    }
    void add(int kind, int lg, int grid, int variant, int64_t stride, const std::vector<TaskDesc>& td,
             const std::vector<QueryDesc>& qd, const std::vector<int>& slots, const std::vector<QueueJob>* jobs) {
        static const unsigned zero = 0;  // the queue position, reset by every replay's upload
        Rec r{kind, lg, grid, jobs ? (int)jobs->size() : 0, variant, stride, 0, 0, 0, 0, 0};
        if (jobs) r.o_jb = put(jobs->data(), sizeof(QueueJob) * jobs->size());
        r.o_td = put(td.data(), sizeof(TaskDesc) * td.size());
        r.o_qd = put(qd.data(), sizeof(QueryDesc) * qd.size());
        r.o_sl = put(slots.data(), sizeof(int) * slots.size());
        r.o_nx = put(&zero, sizeof zero);
        recs.push_back(r);
        outs.insert(outs.end(), qd.begin(), qd.end());
    }
};

void drop_graph(cf_engine* e) {
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
    e->graph_sig.clear();
}

int ensure_pinned(cf_engine* e, void** buf, size_t* cap, size_t bytes) {
    if (bytes <= *cap) return CF_OK;
    drop_graph(e);  // the graph holds the old address
    if (*buf) (void)hipHostFree(*buf);
    *buf = nullptr, *cap = 0;
    HIP_TRY(e, hipHostMalloc(buf, bytes, hipHostMallocDefault), "hipHostMalloc");
    *cap = bytes;
    return CF_OK;
}

// The captured chain: staging upload -> every launch in order -> result download.
int capture(cf_engine* e, const Recorder& R, size_t o_res, size_t nres) {
    HIP_TRY(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    char* sm = (char*)e->small;
    hipError_t st = hipMemcpyAsync(sm, e->pin_stage, R.stage.size(), hipMemcpyHostToDevice, e->stream);
    for (size_t k = 0; st == hipSuccess && k < R.recs.size(); ++k) {
        const Rec& r = R.recs[k];
        uint64_t* res = (uint64_t*)(sm + o_res);
        if (r.kind == 1)
            st = launch_queue(r.variant, r.grid, r.njobs, (const QueueJob*)(sm + r.o_jb), (const TaskDesc*)(sm + r.o_td),
                              (const QueryDesc*)(sm + r.o_qd), (const int*)(sm + r.o_sl), (double*)e->arena, r.stride, res,
                              (unsigned*)(sm + r.o_nx), e->stream);
        else
            st = launch_tasks(r.lg, r.grid, (const TaskDesc*)(sm + r.o_td), (const QueryDesc*)(sm + r.o_qd),
                              (const int*)(sm + r.o_sl), (double*)e->arena, res, e->stream, r.variant);
    }
    if (st == hipSuccess)
        st = hipMemcpyAsync(e->pin_res, sm + o_res, sizeof(uint64_t) * std::max<size_t>(nres, 1), hipMemcpyDeviceToHost,
                            e->stream);
    hipGraph_t g = nullptr;
    const hipError_t st_end = hipStreamEndCapture(e->stream, &g);
    if (st != hipSuccess || st_end != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(e, st != hipSuccess ? st : st_end, "graph capture");
    }
    hipGraphExec_t x = nullptr;
    st = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (st != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, st, "hipGraphInstantiate");
    }
    e->graph = g;
    e->graph_exec = x;
    return CF_OK;
}

// Runs the recorded launches as one graph (captured now, or replayed when the signature of the
// plan and of the buffers it points into is the one already captured).
int run_graph(cf_engine* e, const Recorder& R, size_t nres, uint64_t* res, float& ms) {
    const size_t o_res = (R.stage.size() + 255) & ~(size_t)255;
    int rc = ensure(e, &e->small, &e->small_cap, o_res + sizeof(uint64_t) * std::max<size_t>(nres, 1), "the task descriptors");
    if (rc) return rc;
    if ((rc = ensure_pinned(e, (void**)&e->pin_stage, &e->pin_stage_cap, R.stage.size()))) return rc;
    if ((rc = ensure_pinned(e, (void**)&e->pin_res, &e->pin_res_cap, sizeof(uint64_t) * std::max<size_t>(nres, 1))))
        return rc;
    std::vector<int64_t> sig = {(int64_t)R.stage.size(), (int64_t)nres, (int64_t)(uintptr_t)e->small,
                                (int64_t)(uintptr_t)e->arena, (int64_t)(uintptr_t)e->pin_stage, (int64_t)(uintptr_t)e->pin_res};
    for (const Rec& r : R.recs)
        sig.insert(sig.end(), {r.kind, r.lg, r.grid, r.njobs, r.variant, r.stride, (int64_t)r.o_jb, (int64_t)r.o_td, (int64_t)r.o_qd,
                               (int64_t)r.o_sl, (int64_t)r.o_nx});
    memcpy(e->pin_stage, R.stage.data(), R.stage.size());
    if (sig != e->graph_sig || !e->graph_exec) {
        drop_graph(e);
        if ((rc = capture(e, R, o_res, nres))) return rc;
        e->graph_sig = sig;
        e->stats.graph = 1;
    } else {
        e->stats.graph = 2;
    }
    HIP_TRY(e, hipEventRecord(e->ev0, e->stream), "hipEventRecord");
    HIP_TRY(e, hipGraphLaunch(e->graph_exec, e->stream), "hipGraphLaunch");
    HIP_TRY(e, hipEventRecord(e->ev1, e->stream), "hipEventRecord");
    HIP_TRY(e, hipStreamSynchronize(e->stream), "graph execution");
    HIP_TRY(e, hipEventElapsedTime(&ms, e->ev0, e->ev1), "hipEventElapsedTime");
    for (const QueryDesc& q : R.outs) res[q.out] = e->pin_res[q.out];
    return CF_OK;
}

// XCD-local group barriers for the 2^16 kernel, unless CF_WIDE_DEVICE_BARRIERS=1 asks for
// device-scope ones everywhere.
int wide_local() {
    const char* v = getenv("CF_WIDE_DEVICE_BARRIERS");
    return !(v && v[0] == '1');
}

// Tasks per launch of the 16-workgroup kernel: as many as the device keeps resident, at most
// kWideMaxTasks (one workgroup per CU: two per CU, where registers and LDS would allow them,
// measured slower); CF_WIDE_TASKS=k caps it lower (for measurements).
size_t wide_task_cap(int lg, int device) {
    size_t cap = (size_t)std::min(kWideMaxTasks, std::max(wide_resident(lg, device), 0) / kWideMembers);
    if (const char* v = getenv("CF_WIDE_TASKS")) {
        const int k = atoi(v);
        if (k > 0) cap = std::min(cap, (size_t)k);
    }
    return cap;
}

// Runs every task. With queue, the tasks whose transform fits the work-queue kernel go to one
// persistent launch (longest first); the rest, and everything without queue, run as launches
// of one transform length each, packed into device memory.
int execute(cf_engine* e, std::vector<TaskPlan>& tasks, size_t nres, uint64_t* res, bool queue) {
    const size_t budget = usable_bytes(e);
    float kernel_ms_total = 0;
    // graph mode records the launches and runs them as one graph at the end (not with the
    // cooperative 2^16 launches)
    const bool record = e->graphs && std::none_of(tasks.begin(), tasks.end(), [](const TaskPlan& t) { return t.lg >= kWideLog2; });
    Recorder R;
    std::vector<TaskDesc> td;
    std::vector<QueryDesc> qd;
    std::vector<int> slots;
    if (queue) {
        auto mid = std::stable_partition(tasks.begin(), tasks.end(), [](const TaskPlan& t) { return t.lg <= kQueueMaxLog2L; });
        const size_t nq = (size_t)(mid - tasks.begin());
        // longest first; tasks of one shape adjacent, so they can share a job
        auto shape = [](const TaskPlan& t) {
            std::vector<int> k = {t.lg, t.nm, t.dd, t.model, (int)t.queries.size()};
            for (const QueryDesc& q : t.queries) k.push_back(q.n);
            return k;
        };
        std::stable_sort(tasks.begin(), mid, [&](const TaskPlan& a, const TaskPlan& b) {
            return a.cost != b.cost ? a.cost > b.cost : shape(a) < shape(b);
        });
        std::vector<QueueJob> jobs;
        size_t stride = 0;
        const int qv = e->queue_variant, qthreads = queue_variant_info(qv).threads;
        for (size_t k = 0; k < nq;) {
            const int lg = tasks[k].lg, G = queue_slots(lg, qthreads);
            const std::vector<int> key = shape(tasks[k]);
            size_t j = k + 1;
            while (j < nq && (int)(j - k) < G && shape(tasks[j]) == key) ++j;
            jobs.push_back(QueueJob{(int)k, (int)(j - k), lg, 0});
            for (size_t i = k; i < j; ++i) stride = std::max(stride, (size_t)G * ((tasks[i].bytes + 255) & ~(size_t)255));
            k = j;
        }
        if (nq) {
            stride = (stride + 1023) & ~(size_t)1023;  // every slot stride a multiple of 256 bytes
            const size_t wg = std::min({jobs.size(), (size_t)std::max(queue_resident(e->device, qv), 1), budget / stride});
            if (!wg)
                return fail(e, CF_E_NOMEM, "one job needs " + std::to_string(stride) + " bytes of device memory, " +
                                               std::to_string(budget) + " usable");
            int rc = ensure(e, &e->arena, &e->arena_cap, wg * stride, "the task tables");
            if (rc) return rc;
            describe(e, tasks, 0, nq, true, td, qd, slots);
            if (record) {
                R.add(1, tasks[0].lg, (int)wg, qv, (int64_t)(stride / 8), td, qd, slots, &jobs);
            } else {
                Uploaded u;
                rc = upload(e, td, qd, slots, nres, u, &jobs);
                if (rc) return rc;
                float ms = 0;
                rc = timed_launch(e, [&] {
                    return launch_queue(qv, (int)wg, (int)jobs.size(), u.jobs, u.td, u.qd, u.slots, (double*)e->arena,
                                        (int64_t)(stride / 8), u.res, u.next, e->stream);
                }, qd, u, nres, res, ms);
                if (rc) return rc;
                kernel_ms_total += ms;
            }
            e->stats.launches++;
            e->stats.tasks += (int)nq;
            e->stats.log2_length = std::max(e->stats.log2_length, tasks[0].lg);
            e->stats.device_bytes = std::max(e->stats.device_bytes, wg * stride);
        }
        tasks.erase(tasks.begin(), mid);
    }
    // largest transforms first, then by table size, so a launch holds tasks of one length
    std::stable_sort(tasks.begin(), tasks.end(), [](const TaskPlan& a, const TaskPlan& b) {
        return a.lg != b.lg ? a.lg > b.lg : a.nm > b.nm;
    });
    size_t at = 0;
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
        const bool wide = lg >= kWideLog2;
        const int variant = wide ? 0 : e->variant[lg];
        const size_t cap = wide ? wide_task_cap(lg, e->device)
                                : (size_t)std::max(resident_tasks(lg, e->device, variant), 1);
        if (!cap) return fail(e, CF_E_UNSUPPORTED, "the device cannot hold the 16 workgroups of a 2^16 transform");
        const size_t fit = std::min(end - at, wide ? cap : end - at), left = cls_end - at;
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
        describe(e, tasks, at, end, false, td, qd, slots);
        int rc = ensure(e, &e->arena, &e->arena_cap, bytes, "the task tables");
        if (rc) return rc;
        e->stats.launches++;
        e->stats.tasks += nt;
        e->stats.log2_length = std::max(e->stats.log2_length, lg);
        e->stats.device_bytes = std::max(e->stats.device_bytes, bytes);
        if (record) {
            R.add(0, lg, nt, variant, 0, td, qd, slots, nullptr);
            at = end;
            continue;
        }
        Uploaded u;
        rc = upload(e, td, qd, slots, nres, u, nullptr, wide ? wide_sync_words(nt) : 0);
        if (rc) return rc;
        float ms = 0;
        rc = timed_launch(e, [&] {
            if (wide) return launch_wide(lg, nt, u.td, u.qd, u.slots, (double*)e->arena, u.res, u.sync, wide_local(), e->stream);
            return launch_tasks(lg, nt, u.td, u.qd, u.slots, (double*)e->arena, u.res, e->stream, variant);
        }, qd, u, nres, res, ms);
        if (rc) return rc;
        if (wide) {
            unsigned aborted = 0;
            HIP_TRY(e, hipMemcpy(&aborted, u.sync + 2 * nt, sizeof aborted, hipMemcpyDeviceToHost), "download");
            if (aborted) return fail(e, CF_E_DEVICE, "a group barrier of the 2^16 transform timed out");
            // the XCD each workgroup ran on: a task whose members differ used device-scope barriers
            std::vector<unsigned> xcc((size_t)kWideMembers * nt);
            HIP_TRY(e, hipMemcpy(xcc.data(), u.sync + 2 * nt + 2, sizeof(unsigned) * xcc.size(), hipMemcpyDeviceToHost),
                    "download");
            for (int k = 0; k < nt; ++k)
                e->stats.wide_split += !std::all_of(xcc.begin() + kWideMembers * k, xcc.begin() + kWideMembers * (k + 1),
                                                    [&](unsigned x) { return x == xcc[(size_t)kWideMembers * k]; });
        }
        kernel_ms_total += ms;
        at = end;
    }
    if (record && !R.recs.empty()) {
        const int rc = run_graph(e, R, nres, res, kernel_ms_total);
        if (rc) return rc;
    }
    e->stats.batches = e->stats.launches;
    e->stats.kernel_ms = kernel_ms_total;
    return CF_OK;
}

}  // namespace cf_int

using namespace cf_int;

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
    // the device address of cf_query_noalloc's fixed scratch, looked up once per device
    if (device < cf::kQueryScratchDevices) HIP_TRY(e.get(), cf::attach_query_scratch(device), "hipGetSymbolAddress");
    HIP_TRY(e.get(), hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "hipStreamCreate");
    HIP_TRY(e.get(), hipEventCreate(&e->ev0), "hipEventCreate");
    HIP_TRY(e.get(), hipEventCreate(&e->ev1), "hipEventCreate");
    *out = e.release();
    return CF_OK;
}

int cf_set_graphs(cf_engine* e, int on) {
    if (!e) return CF_E_ARGS;
    e->graphs = on != 0;
    if (!e->graphs) drop_graph(e);
    return CF_OK;
}

void cf_close(cf_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    if (e->pin_stage) (void)hipHostFree(e->pin_stage);
    if (e->pin_res) (void)hipHostFree(e->pin_res);
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
    char name[256] = {0};
    if (hipDeviceGetName(name, sizeof name, e->device) != hipSuccess || !name[0])
        snprintf(name, sizeof name, "%s", pr.name[0] ? pr.name : "AMD GPU");
    snprintf(buf, len, "%s %s (%d CUs)", name, pr.gcnArchName, pr.multiProcessorCount);
    return CF_OK;
}

// Tables beyond the 2^17 transform are not run; their memory is estimated (the two tile pools
// hold about nm^2 / 4 entries each) without building a slot plan of that size.
static size_t beyond_engine_bytes(int nmax, int Mmax, size_t J) {
    const size_t nm = (size_t)nmax, F1 = (size_t)std::max(nmax, Mmax) + 1;
    return 8 * (nm * nm / 2 + 64 * (nm + 1) + 4 * F1 + 2 * J);
}

size_t cf_mem_per_prime(int nmax, int Mmax, size_t J) {
    if (nmax < 1) return 0;
    if (nmax > (1 << kWideMaxLog2)) return beyond_engine_bytes(nmax, Mmax, J);
    const int lg = engine_log2_length(nmax, 1 + (uint64_t(1) << 49));
    return task_bytes(plan_slots(nmax), std::max(nmax, Mmax) + 1, (int)J, lg);
}

long cf_max_batch(cf_engine* e, int nmax, int Mmax, size_t J) {
    if (!e || nmax < 1) return 0;
    (void)hipSetDevice(e->device);
    if (nmax > (1 << kWideMaxLog2)) return (long)(usable_bytes(e) / beyond_engine_bytes(nmax, Mmax, J));
    const int lg = engine_log2_length(nmax, 1 + (uint64_t(1) << 49));
    const size_t per = task_bytes(plan_for(e, nmax), std::max(nmax, Mmax) + 1, (int)J, lg);
    return (long)(usable_bytes(e) / per);
}

size_t cf_mem_per_prime62(int nmax, int Mmax, size_t J) { return nmax < 1 ? 0 : p62_task_bytes(nmax, Mmax, J); }

long cf_max_batch62(cf_engine* e, int nmax, int Mmax, size_t J) {
    if (!e || nmax < 1) return 0;
    (void)hipSetDevice(e->device);
    return (long)(usable_bytes(e) / p62_task_bytes(nmax, Mmax, J));
}

}  // extern "C"

namespace {

// The queries of primes whose p - 1 lacks the engine's shortest transform, one cubic-algorithm
// call per query over those primes (cf_check.hip); the stats of the whole call are kept.
int run_short_primes(cf_engine* e, int d, int model, const int32_t* n, const int32_t* M, size_t J, const uint64_t* primes,
                     const uint8_t* mask, const std::vector<size_t>& idx, uint64_t* res) {
    const cf_stats before = e->stats;
    double kernel_ms = 0;
    int launches = 0;
    std::vector<uint64_t> ps, out;
    std::vector<size_t> at;
    for (size_t j = 0; j < J; ++j) {
        ps.clear();
        at.clear();
        for (size_t b : idx)
            if (!mask || mask[b * J + j]) ps.push_back(primes[b]), at.push_back(b);
        if (ps.empty()) continue;
        out.assign(ps.size(), 0);
        const int rc = cf_run_algo(e, CF_ALGO_CUBIC, d, model, n[j], M[j], ps.data(), ps.size(), out.data());
        if (rc) return rc;
        kernel_ms += e->stats.kernel_ms;
        launches += e->stats.launches;
        for (size_t k = 0; k < at.size(); ++k) res[at[k] * J + j] = out[k];
    }
    e->stats = before;
    e->stats.kernel_ms += kernel_ms;
    e->stats.launches += launches;
    e->stats.tasks += (int)idx.size();
    return CF_OK;
}

}  // namespace

extern "C" {

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
    std::vector<size_t> short_primes, big_primes;
    for (size_t b = 0; b < np; ++b) {
        if (primes[b] >= (uint64_t(1) << 50)) {  // beyond the FP64 product: the 64-bit Montgomery kernel
            big_primes.push_back(b);
            continue;
        }
        std::vector<int> qi;
        int nm = 0;
        for (size_t j = 0; j < J; ++j)
            if (!mask || mask[b * J + j]) qi.push_back((int)j), nm = std::max(nm, (int)n[j]);
        if (qi.empty()) continue;
        // p - 1 without the engine's shortest transform (2^10) but with the table's own power of
        // two next_pow2(nm) (a 32-bit NTT prime with 2^8 | p - 1, say): the table has at most 512
        // rows, and the O(n^3) device algorithm gives its residues
        const int lg_own = nm > 1 ? 32 - __builtin_clz((unsigned)(nm - 1)) : 1;
        if (primes[b] < (uint64_t(1) << 50) && engine_log2_length(nm, primes[b]) < 0 &&
            two_adic_valuation(primes[b] - 1) >= lg_own && nm <= 1024) {
            short_primes.push_back(b);
            continue;
        }
        TaskPlan tp;
        int rc = plan_task(e, primes[b], d, model, qi.size(), [&](size_t k) {
            const int j = qi[k];
            return QueryDesc{n[j], M[j], (int64_t)(b * J + j)};
        }, tp);
        if (rc) return rc;
        tasks.push_back(std::move(tp));
    }
    int rc = execute(e, tasks, np * J, res, false);
    if (rc == CF_OK && !short_primes.empty()) rc = run_short_primes(e, d, model, n, M, J, primes, mask, short_primes, res);
    if (rc == CF_OK && !big_primes.empty()) rc = run_p62(e, d, model, n, M, J, primes, big_primes, mask, res);
    e->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return rc;
}

int cf_run_batch(cf_engine* e, const cf_instance* inst, size_t k, const uint64_t* primes, uint64_t* res) {
    if (!e) return CF_E_ARGS;
    const auto t_start = std::chrono::steady_clock::now();
    e->err.clear();
    e->stats = cf_stats{};
    if (k && (!inst || !primes || !res)) return fail(e, CF_E_ARGS, "bad arguments");
    // the instances' prime ranges (and so their result ranges) must not overlap
    std::vector<std::pair<int64_t, int64_t>> rng;
    int64_t nres = 0;
    for (size_t i = 0; i < k; ++i) {
        const cf_instance& x = inst[i];
        if (x.d < 1 || (x.model != 0 && x.model != 1) || x.n < 1 || x.M <= x.n || x.prime_off < 0 || x.prime_cnt < 0)
            return fail(e, CF_E_ARGS, "bad instance " + std::to_string(i));
        if (x.prime_cnt) rng.emplace_back(x.prime_off, x.prime_off + x.prime_cnt);
        nres = std::max(nres, x.prime_off + x.prime_cnt);
    }
    std::sort(rng.begin(), rng.end());
    for (size_t i = 1; i < rng.size(); ++i)
        if (rng[i].first < rng[i - 1].second) return fail(e, CF_E_ARGS, "instances share prime/result entries");
    for (const auto& r : rng) std::fill(res + r.first, res + r.second, 0);
    HIP_TRY(e, hipSetDevice(e->device), "hipSetDevice");
    // the tables: one per (d, model, prime) with every query on it when sharing, else one per
    // (instance, prime). A shared table has the rows of its largest n; a smaller query reads the
    // first rows of it, which are those of its own table (row v of V and S depends only on rows
    // and columns below v, and the cyclic convolutions of length L >= nm never wrap into them).
    std::map<std::tuple<int, int, uint64_t>, std::vector<QueryDesc>> shared;
    std::vector<TaskPlan> tasks;
    for (size_t i = 0; i < k; ++i) {
        const cf_instance& x = inst[i];
        for (int64_t b = x.prime_off; b < x.prime_off + x.prime_cnt; ++b) {
            if (e->batch_share) {
                shared[std::make_tuple(x.d, x.model, primes[b])].push_back(QueryDesc{x.n, x.M, b});
                continue;
            }
            TaskPlan tp;
            int rc = plan_task(e, primes[b], x.d, x.model, 1, [&](size_t) { return QueryDesc{x.n, x.M, b}; }, tp);
            if (rc) return rc;
            tasks.push_back(std::move(tp));
        }
    }
    for (const auto& kv : shared) {
        const std::vector<QueryDesc>& qs = kv.second;
        TaskPlan tp;
        // (a shared table has the transform length of its largest query, so it is refused only
        // where that query's own table would be)
        int rc = plan_task(e, std::get<2>(kv.first), std::get<0>(kv.first), std::get<1>(kv.first), qs.size(),
                           [&](size_t j) { return qs[j]; }, tp);
        if (rc) return rc;
        tasks.push_back(std::move(tp));
    }
    const int rc = execute(e, tasks, (size_t)nres, res, true);
    e->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return rc;
}

int cf_set_batch_sharing(cf_engine* e, int on) {
    if (!e) return CF_E_ARGS;
    e->batch_share = on != 0;
    return CF_OK;
}

int cf_set_batch_variant(cf_engine* e, int variant) {
    if (!e || variant < 0 || variant >= kQueueVariants) return CF_E_ARGS;
    e->queue_variant = variant;
    return CF_OK;
}

int cf_autotune_batch(cf_engine* e, const cf_instance* inst, size_t k, const uint64_t* primes, size_t sample_tables,
                      cf_batch_tune* out) {
    if (!e || !out || (k && (!inst || !primes))) return CF_E_ARGS;
    memset(out, 0, sizeof *out);
    e->err.clear();
    // the sample: every instance with the same share of its primes (at least one), so the mix of
    // shapes and lengths is the batch's own
    int64_t total = 0;
    for (size_t i = 0; i < k; ++i) total += std::max<int64_t>(inst[i].prime_cnt, 0);
    if (!total) return fail(e, CF_E_ARGS, "the batch has no tables");
    const double f = sample_tables ? std::min(1.0, (double)sample_tables / (double)total) : 1.0;
    std::vector<cf_instance> sub;
    std::vector<uint64_t> sp;
    for (size_t i = 0; i < k; ++i) {
        if (inst[i].prime_cnt <= 0) continue;
        const int64_t c = std::max<int64_t>(1, (int64_t)std::ceil(f * (double)inst[i].prime_cnt));
        cf_instance x = inst[i];
        x.prime_off = (int64_t)sp.size();
        x.prime_cnt = c;
        sp.insert(sp.end(), primes + inst[i].prime_off, primes + inst[i].prime_off + c);
        sub.push_back(x);
    }
    const int saved = e->queue_variant;
    std::vector<uint64_t> first, res(sp.size());
    for (int v = 0; v < kQueueVariants; ++v) {
        e->queue_variant = v;
        // code load and first touch of this variant on one instance, then the timed sample
        int rc = cf_run_batch(e, sub.data(), 1, sp.data(), res.data());
        if (!rc) rc = cf_run_batch(e, sub.data(), sub.size(), sp.data(), res.data());
        if (rc) {
            e->queue_variant = saved;
            return rc;
        }
        const QueueVariant qv = queue_variant_info(v);
        out->threads[v] = qv.threads;
        out->per_cu[v] = qv.per_cu;
        out->resident[v] = queue_resident(e->device, v);
        out->ms[v] = e->stats.kernel_ms;
        if (v == 0) first = res;
        else if (res != first) {
            e->queue_variant = saved;
            return fail(e, CF_E_DEVICE, "work-queue variant " + std::to_string(v) + " gave residues different from variant 0");
        }
    }
    int best = 0;
    for (int v = 1; v < kQueueVariants; ++v)
        if (out->ms[v] < out->ms[best]) best = v;
    e->queue_variant = best;
    out->variants = kQueueVariants;
    out->chosen = best;
    out->tables = (int)sp.size();
    return CF_OK;
}

int cf_set_variant(cf_engine* e, int log2_length, int variant) {
    if (!e || log2_length < kMinLog2L || log2_length > kMaxLog2L) return CF_E_ARGS;
    if (variant < 0) variant = 0;
    if (variant >= variant_count(log2_length)) return CF_E_ARGS;
    e->variant[log2_length] = variant;
    return CF_OK;
}

int cf_autotune(cf_engine* e, int log2_length, int n_cal, cf_tune* out) {
    if (!e || !out || log2_length < kMinLog2L || log2_length > kMaxLog2L) return CF_E_ARGS;
    const int L = 1 << log2_length;
    if (n_cal && (n_cal > L || (log2_length > kMinLog2L && n_cal <= L / 2) || n_cal < 1)) return CF_E_ARGS;
    memset(out, 0, sizeof *out);
    e->err.clear();
    HIP_TRY(e, hipSetDevice(e->device), "hipSetDevice");
    const int lg = log2_length, nv = std::min(variant_count(lg), 4);
    // calibration: tables of n rows (the caller's, else the class's smallest, L/2 + 1), M = 2n,
    // d = 3, two waves of the most resident variant's workgroups (every variant at full occupancy)
    const int n = n_cal ? n_cal : L / 2 + 1, M = 2 * n;
    int most = 1;
    for (int v = 0; v < nv; ++v) most = std::max(most, resident_tasks(lg, e->device, v));
    const size_t per = task_bytes(plan_for(e, n), M + 1, 1, lg);
    const size_t count = std::max<size_t>(1, std::min<size_t>(2 * (size_t)most, usable_bytes(e) / per));
    const std::vector<uint64_t> primes = prime_range(50, 0, (long)count);
    const int saved = e->variant[lg];
    std::vector<uint64_t> first;
    auto run = [&](int v, size_t np, std::vector<uint64_t>& res) -> int {
        e->variant[lg] = v;
        res.assign(np, 0);
        std::vector<TaskPlan> tasks(np);
        for (size_t b = 0; b < np; ++b) {
            int rc = plan_task(e, primes[b], 3, 0, 1, [&](size_t) { return QueryDesc{n, M, (int64_t)b}; }, tasks[b]);
            if (rc) return rc;
        }
        e->stats = cf_stats{};
        const bool g = e->graphs;
        e->graphs = false;
        const int rc = execute(e, tasks, np, res.data(), false);
        e->graphs = g;
        return rc;
    };
    for (int v = 0; v < nv; ++v) {
        std::vector<uint64_t> res;
        int rc = run(v, 1, res);  // code load and first-touch of this variant
        if (!rc) rc = run(v, count, res);
        if (rc) {
            e->variant[lg] = saved;
            return rc;
        }
        const KernelVariant kv = variant_info(lg, v);
        out->waves[v] = kv.waves;
        out->deferred[v] = kv.deferred;
        out->resident[v] = resident_tasks(lg, e->device, v);
        out->ms[v] = e->stats.kernel_ms;
        if (v == 0) first = res;
        else if (res != first) {
            e->variant[lg] = saved;
            return fail(e, CF_E_DEVICE, "kernel variant " + std::to_string(v) + " of 2^" + std::to_string(lg) +
                                            " gave residues different from variant 0");
        }
        if (out->ms[v] < out->ms[out->chosen]) out->chosen = v;
    }
    out->log2_length = lg;
    out->variants = nv;
    out->tables = (int)count;
    out->n = n;
    e->variant[lg] = out->chosen;
    e->stats = cf_stats{};
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
    return lg > kWideMaxLog2 || (lg > kMaxLog2L && lg < kWideLog2) ? -1 : lg;
}

}  // extern "C"
