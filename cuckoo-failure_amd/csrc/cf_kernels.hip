// cf_kernels.hip -- the gfx950 task kernel: one workgroup runs one prime's whole suffix-table
// build and query reduction (cf_task.hpp), L/16 threads with 16 transform entries each, the
// transform buffer, twiddle and power tables in LDS.
#include <hip/hip_runtime.h>

#include "cf_kernels.hpp"
#include "cf_task_kernel.hpp"
#include "cf_wide.hpp"

namespace cf {

// ---- the work-queue kernel: a persistent grid over jobs of transform lengths 2^10 .. 2^12 ----
// A job is up to THREADS/256 tasks of L = 2^12, THREADS/128 of 2^11 or THREADS/64 of 2^10 that
// the workgroup runs side by side, each on its own T = L/16 threads (whole waves), LDS slice and
// arena slice. The host puts only tasks of identical shape (nm, d, model, query sizes) into one
// job, so their phase sequences are the same and they share the workgroup barriers; a job with
// fewer tasks than slots repeats its first task in the spare slots and discards those results.
// Workgroups take jobs from the queue in the host's order (longest first) until it is empty.
template <int THREADS>
union QueueLds {
    Shared<12> s12[THREADS / 256];
    Shared<11> s11[THREADS / 128];
    Shared<10> s10[THREADS / 64];
};

template <int LG, int THREADS>
__device__ __forceinline__ void queue_job(QueueLds<THREADS>& lds, const QueueJob& jb, const TaskDesc* tasks, const QueryDesc* qs,
                                          const int* slots, double* arena, int64_t slot_stride, uint64_t* res) {
    constexpr int T = Plan<LG>::T, G = THREADS / T;
    static_assert(G * sizeof(Shared<LG>) <= sizeof(QueueLds<THREADS>), "LDS");
    const int g = (int)threadIdx.x / T;  // uniform over each wave, since T is a multiple of 64
    Shared<LG>& sh = reinterpret_cast<Shared<LG>*>(&lds)[g];
    const int k = __builtin_amdgcn_readfirstlane(jb.first + (g < jb.cnt ? g : 0));
    DevRunner<1, T> run;
    run.t = (int)threadIdx.x - g * T;
    const Ctx c = make_ctx<LG>(tasks + k, qs, slots, arena + slot_stride * g);
    Task<LG>::run_task(run, sh, c);
    if (g < jb.cnt)
        for (int j = run.t; j < c.d->q_cnt; j += T) res[qs[c.d->q_off + j].out] = (uint64_t)Task<LG>::result(j, c);
}

template <int THREADS, int PER_CU>
__global__ void __launch_bounds__(THREADS, PER_CU)
    k_queue(const QueueJob* __restrict__ jobs, int njobs, const TaskDesc* __restrict__ tasks, const QueryDesc* __restrict__ qs,
            const int* __restrict__ slots, double* __restrict__ mem, int64_t arena_stride, uint64_t* __restrict__ res,
            unsigned* __restrict__ next) {
    __shared__ QueueLds<THREADS> lds;
    __shared__ int take;
    double* arena = mem + arena_stride * blockIdx.x;
    for (;;) {
        if (threadIdx.x == 0) take = (int)atomicAdd(next, 1u);
        __syncthreads();
        const int k = take;
        __syncthreads();  // every thread has read take before thread 0 writes the next one
        if (k >= njobs) break;
        const QueueJob jb = jobs[k];
        switch (jb.lg) {
            case 10: queue_job<10, THREADS>(lds, jb, tasks, qs, slots, arena, arena_stride / (THREADS / 64), res); break;
            case 11: queue_job<11, THREADS>(lds, jb, tasks, qs, slots, arena, arena_stride / (THREADS / 128), res); break;
            default: queue_job<12, THREADS>(lds, jb, tasks, qs, slots, arena, arena_stride / (THREADS / 256), res); break;
        }
        __syncthreads();  // the results are written before the next job reuses the arena
    }
}

// variant 0: 256 threads, compiled for 2 workgroups per CU (the built-in choice); 1: the same
// workgroup at one per CU (the register budget of a whole SIMD, no spills); 2: 512 threads (two
// 2^12 tasks, four of 2^11, eight of 2^10 side by side) at one per CU
QueueVariant queue_variant_info(int variant) {
    switch (variant) {
        case 1: return {256, 1};
        case 2: return {512, 1};
        default: return {256, 2};
    }
}

template <int THREADS, int PER_CU>
int queue_resident_t(int device) {
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, device) != hipSuccess) return 0;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_queue<THREADS, PER_CU>, THREADS, 0) != hipSuccess) return 0;
    return pr.multiProcessorCount * (per > 0 ? per : 1);
}

int queue_resident(int device, int variant) {
    switch (variant) {
        case 1: return queue_resident_t<256, 1>(device);
        case 2: return queue_resident_t<512, 1>(device);
        default: return queue_resident_t<256, 2>(device);
    }
}

hipError_t launch_queue(int variant, int workgroups, int njobs, const QueueJob* jobs, const TaskDesc* tasks, const QueryDesc* qs,
                        const int* slots, double* mem, int64_t arena_stride, uint64_t* res, unsigned* next, hipStream_t stream) {
    if (njobs <= 0 || workgroups <= 0) return hipSuccess;
    const dim3 grid((unsigned)workgroups);
    switch (variant) {
        case 1: k_queue<256, 1><<<grid, 256, 0, stream>>>(jobs, njobs, tasks, qs, slots, mem, arena_stride, res, next); break;
        case 2: k_queue<512, 1><<<grid, 512, 0, stream>>>(jobs, njobs, tasks, qs, slots, mem, arena_stride, res, next); break;
        default: k_queue<256, 2><<<grid, 256, 0, stream>>>(jobs, njobs, tasks, qs, slots, mem, arena_stride, res, next); break;
    }
    return hipGetLastError();
}

// ---- tasks spread over kWideG workgroups (transforms of 2^16 and 2^17) ----
// Group barrier of the workgroups of one task: each arrives on a monotone counter and waits until
// all have arrived. The wait is bounded: a waiter that gives up sets the launch's abort flag,
// which releases every other waiter, and the task ends with an error instead of hanging.
constexpr long kWideSpinLimit = 1L << 24;
#ifndef CF_WIDE_POLL_SLEEP
#define CF_WIDE_POLL_SLEEP 2  // s_sleep between polls of the arrival counter (units of 64 clocks)
#endif
static_assert(kWideMembers == kWideG, "workgroups per wide task");

template <int T>
struct WideRunner : DevRunner<1, T> {
    unsigned* cnt;    // this task's arrival counter
    unsigned* abort;  // the launch's abort flag
    int* flag;        // LDS: the outcome for the whole workgroup
    unsigned target = 0;
    bool local = false;  // all workgroups of the task share one XCD (and L2)
    __device__ __forceinline__ bool gsync() {
        // Every wave's loads and stores are complete (its stores in this XCD's L2) before it
        // arrives: the workgroup barrier waits for LDS operations only, and the fence below runs
        // in one wave, so without this another wave's stores could miss the write-back and its
        // loads could still be reading what the next phase's writers overwrite.
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (threadIdx.x == 0) {
            target += kWideG;
            if (!local) __threadfence();  // this workgroup's writes (all complete in L2 by now) to the device
            __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int bad = 0;
            long spins = 0;
            while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    bad = 1;
                    break;
                }
                if (++spins > kWideSpinLimit) {
                    __hip_atomic_fetch_or(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    bad = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(CF_WIDE_POLL_SLEEP);
            }
            // the other workgroups' writes, not stale cache lines: within one XCD only this CU's
            // L1 can be stale (the L2 is shared); otherwise the device-scope fence. Its invalidate
            // must be complete before the workgroup barrier below releases the other waves: their
            // first loads (often of lines this CU read a step earlier) may not hit lines it has
            // not yet reached.
            if (local) asm volatile("buffer_inv sc0" ::: "memory");
            else __threadfence();
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            *flag = bad;
        }
        __syncthreads();
        return *flag != 0;
    }
};

__device__ __forceinline__ unsigned xcc_id() {
    unsigned x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return x;
}

// Workgroup -> (task, member): workgroups are dealt to the 8 XCDs round robin, so the kWideG
// members of a task get numbers 8 apart and share one XCD (and its L2) for their exchanges. Each
// workgroup records the XCD it runs on; after the first (device-scope) group barrier every member
// checks the record, and only when all kWideG share one XCD (and `local` is allowed) do the
// remaining barriers use the XCD-local form.
// sync: [2 ntasks] arrival counters (zero at launch), the abort flag, then kWideG XCD ids per task.
template <int INNER>
__global__ void __launch_bounds__(Wide<INNER>::T)
    k_wide(const TaskDesc* __restrict__ tasks, int ntasks, const QueryDesc* __restrict__ qs, const int* __restrict__ slots,
           double* __restrict__ mem, uint64_t* __restrict__ res, unsigned* __restrict__ sync, int local) {
    using W = Wide<INNER>;
    __shared__ typename W::Sh sh;
    __shared__ int flag;
    const int L = (int)blockIdx.x, xcd = L & 7, i = L >> 3;
    const int task = xcd + 8 * (i / kWideG), g = i % kWideG;
    if (task >= ntasks) return;
    const TaskDesc* d = tasks + task;
    WideRunner<W::T> run;
    run.t = (int)threadIdx.x;
    run.cnt = sync + 2 * task;
    run.abort = sync + 2 * ntasks;
    run.flag = &flag;
    unsigned* xcc = sync + 2 * ntasks + 2 + kWideG * task;
    if (threadIdx.x == 0) xcc[g] = xcc_id();
    if (run.gsync()) return;
    if (local) {
        bool same = true;
        for (int h = 1; h < kWideG; ++h) same &= xcc[h] == xcc[0];
        run.local = same;  // the same data for every member: one decision per task
    }
    const Ctx c = W::ctx(d, qs, slots, mem, g, d->model, d->dd);
    const WideArena w = wide_arena_of<INNER>(mem, *d);
    const bool ok = W::run_task(run, sh, c, w, g);
    if (ok && g == 0)
        for (int j = threadIdx.x; j < d->q_cnt; j += W::T) res[qs[d->q_off + j].out] = (uint64_t)W::TK::result(j, c);
}

template <int INNER>
int wide_resident_t(int device) {
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, device) != hipSuccess) return 0;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_wide<INNER>, Wide<INNER>::T, 0) != hipSuccess) return 0;
    return pr.multiProcessorCount * per;
}

int wide_resident(int log2L, int device) {
    return log2L == 17 ? wide_resident_t<13>(device) : log2L == 16 ? wide_resident_t<12>(device) : 0;
}

template <int INNER>
hipError_t launch_wide_t(int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots, double* mem,
                         uint64_t* res, unsigned* sync, int local, hipStream_t stream) {
    const unsigned grid = (unsigned)(((ntasks + 7) / 8) * 8 * kWideG);
    void* args[] = {(void*)&tasks, (void*)&ntasks, (void*)&qs, (void*)&slots, (void*)&mem, (void*)&res, (void*)&sync,
                    (void*)&local};
    return hipLaunchCooperativeKernel((const void*)k_wide<INNER>, dim3(grid), dim3(Wide<INNER>::T), args, 0, stream);
}

hipError_t launch_wide(int log2L, int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots, double* mem,
                       uint64_t* res, unsigned* sync, int local, hipStream_t stream) {
    if (ntasks <= 0) return hipSuccess;
    if (log2L == 16) return launch_wide_t<12>(ntasks, tasks, qs, slots, mem, res, sync, local, stream);
    if (log2L == 17) return launch_wide_t<13>(ntasks, tasks, qs, slots, mem, res, sync, local, stream);
    return hipErrorInvalidValue;
}

// Workgroups of the log2L kernel that the device runs at the same time (CUs x blocks per CU).
// Variant 0 of each length is the built-in choice: one virtual thread per hardware thread up to
// 2^13, two at 2^14; the tile flush deferred wherever the transform has the carrier rounds.
// The other variants live in cf_variants.hip.
int variant_count(int log2L) {
    switch (log2L) {
        case 10: return 1;
        case 11: case 12: return 2;
        case 13: return 3;
        case 14: return 2;
        default: return 0;
    }
}

KernelVariant variant_info(int log2L, int variant) {
    const int T = (1 << log2L) / 16;
    static const int vt[5][3] = {{1, 0, 0}, {1, 2, 0}, {1, 2, 0}, {1, 2, 1}, {2, 2, 0}};
    static const int df[5][3] = {{-1, 0, 0}, {-1, -1, 0}, {-1, -1, 0}, {-1, -1, 0}, {-1, 0, 0}};
    if (log2L < kMinLog2L || log2L > kMaxLog2L || variant < 0 || variant >= variant_count(log2L)) return {0, 0};
    const int i = log2L - kMinLog2L;
    return {T / vt[i][variant] / 64, df[i][variant] != 0 && log2L >= 13};
}

int resident_tasks(int log2L, int device, int variant) {
    if (variant) return resident_variant(log2L, variant, device);
    switch (log2L) {
        case 10: return resident_k<10, 1, -1>(device);
        case 11: return resident_k<11, 1, -1>(device);
        case 12: return resident_k<12, 1, -1>(device);
        case 13: return resident_k<13, 1, -1>(device);
        case 14: return resident_k<14, 2, -1>(device);
        default: return 0;
    }
}

hipError_t launch_tasks(int log2L, int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots, double* mem,
                        uint64_t* res, hipStream_t stream, int variant) {
    if (ntasks <= 0) return hipSuccess;
    if (variant) return launch_variant(log2L, variant, ntasks, tasks, qs, slots, mem, res, stream);
    switch (log2L) {
        case 10: return launch_k<10, 1, -1>(ntasks, tasks, qs, slots, mem, res, stream);
        case 11: return launch_k<11, 1, -1>(ntasks, tasks, qs, slots, mem, res, stream);
        case 12: return launch_k<12, 1, -1>(ntasks, tasks, qs, slots, mem, res, stream);
        case 13: return launch_k<13, 1, -1>(ntasks, tasks, qs, slots, mem, res, stream);
        case 14: return launch_k<14, 2, -1>(ntasks, tasks, qs, slots, mem, res, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cf
