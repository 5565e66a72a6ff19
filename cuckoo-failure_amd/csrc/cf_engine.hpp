// cf_engine.hpp -- the engine's internals shared by the C ABI's source files: the engine
// object, task plans and the executor that runs them.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/cuckoo_amd.h"
#include "cf_core.hpp"
#include "cf_host.hpp"

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
    std::map<int, cf::SlotPlan> plans;  // by nm
    int variant[32] = {};               // task-kernel variant per log2 transform length (cf_autotune)
    int queue_variant = 0;              // work-queue kernel variant (cf_autotune_batch)
    bool batch_share = true;            // cf_run_batch: one table per (d, model, prime)
    // graph mode (cf_set_graphs): a call's uploads, launches and downloads are one hipGraph,
    // captured once per launch-plan signature and replayed while the signature repeats
    bool graphs = false;
    std::vector<int64_t> graph_sig;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    char* pin_stage = nullptr;  // pinned host staging of the descriptors
    size_t pin_stage_cap = 0;
    uint64_t* pin_res = nullptr;  // pinned host copy of the results
    size_t pin_res_cap = 0;
};

namespace cf_int {

using cf::QueryDesc;

struct TaskPlan {
    uint64_t p;
    int dd, model;
    int nm, F1, lg;
    std::vector<QueryDesc> queries;  // sorted by n descending; out = index into the result array
    size_t bytes;
    double cost;                     // estimated run time (arbitrary units), for ordering
    double* keep = nullptr;          // retain mode: the prime's resident table store
};

int fail(cf_engine* e, int code, const std::string& msg);
int hip_fail(cf_engine* e, hipError_t st, const char* what);

#define HIP_TRY(e, x, what)                                   \
    do {                                                      \
        hipError_t st_ = (x);                                 \
        if (st_ != hipSuccess) return cf_int::hip_fail(e, st_, what); \
    } while (0)

const cf::SlotPlan& plan_for(cf_engine* e, int nm);
// Grows *buf (device memory) to at least bytes.
int ensure(cf_engine* e, void** buf, size_t* cap, size_t bytes, const char* what);
// Device memory a call may use: free now plus the engine's arena, minus a reserve.
size_t usable_bytes(cf_engine* e);
// Runs every task (see cf_engine.cpp); res gets each query's residue at its out index.
int execute(cf_engine* e, std::vector<TaskPlan>& tasks, size_t nres, uint64_t* res, bool queue);
// The queries of primes above 2^50 (up to 2^63), primes[b] for b in idx, on the 64-bit Montgomery
// kernel (cf_p62_host.cpp): res[b J + j] as cf_run_fused.
int run_p62(cf_engine* e, int d, int model, const int32_t* n, const int32_t* M, size_t J, const uint64_t* primes,
            const std::vector<size_t>& idx, const uint8_t* mask, uint64_t* res);
// Device bytes of one prime on that kernel.
size_t p62_task_bytes(int nmax, int Mmax, size_t J);

}  // namespace cf_int
