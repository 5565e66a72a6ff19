// cf_task_kernel.hpp -- the per-prime task kernel template (one workgroup runs one prime's whole
// suffix-table build and query reduction, cf_task.hpp) and the thread runner it shares with the
// work-queue and 2^16 kernels. Instantiated in cf_kernels.hip (each length's default variant)
// and cf_variants.hip (the other variants the autotuner times).
#pragma once
#include <hip/hip_runtime.h>

#include "cf_kernels.hpp"
#include "cf_task.hpp"

namespace cf {

// A double moved between lanes by a DPP pattern (both halves), all rows and banks enabled.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Sum over the 64 lanes of a wavefront (every lane active), reduced after every level (each level
// at most doubles). Within a row of 16 lanes the partners come from DPP (quad_perm [1,0,3,2] and
// [2,3,0,1], then row_half_mirror and row_mirror: after each level every lane of the group holds
// the group's sum), across rows from two lane shuffles; every lane ends with the total.
__device__ __forceinline__ double wave_sum(double x, const Mod& m) {
    x = reduce(x + dpp_f64<0xB1>(x), m);   // quad_perm [1,0,3,2]
    x = reduce(x + dpp_f64<0x4E>(x), m);   // quad_perm [2,3,0,1]
    x = reduce(x + dpp_f64<0x141>(x), m);  // row_half_mirror: lane i with 7 - i of its half-row
    x = reduce(x + dpp_f64<0x140>(x), m);  // row_mirror: lane i with 15 - i of its row
    x = reduce(x + __shfl_xor(x, 16, 64), m);
    x = reduce(x + __shfl_xor(x, 32, 64), m);
    return x;
}

// A workgroup of T / VT hardware threads runs the T threads of the task plan, VT per hardware
// thread (a "virtual thread" is one Thread state; hardware thread h runs virtual threads
// h, h + T/VT, ...). With VT = 2 the L = 2^14 kernel has 512 hardware threads and a 256-register
// budget, so each wave carries twice the independent work between barriers.
template <int VT, int TP>
struct DevRunner {
    Thread th[VT];
    int t;  // hardware thread
    // The thread index is passed through an empty asm statement, so the index arithmetic of a
    // phase is recomputed there instead of being hoisted out of the v-loop and kept live (in
    // spilled registers) across all steps. All virtual threads' indices are formed first, so the
    // scheduler can start the second one's loads under the first one's arithmetic.
    template <class F>
    __device__ __forceinline__ void phase(F f) {
        int tt[VT];
#pragma unroll
        for (int i = 0; i < VT; ++i) {
            tt[i] = t + i * TP;
            asm volatile("" : "+v"(tt[i]));
        }
#pragma unroll
        for (int i = 0; i < VT; ++i) f(th[i], tt[i]);
        __syncthreads();
    }
    // Inside a phase: this wave's sums of the cnt partials, written by lane 0 to red[wave][8].
    template <class F>
    __device__ __forceinline__ void wave_partial8(double* red, Thread& th_, int t_, F f, const Mod& m) {
        double acc[8];
        const int cnt = f(th_, t_, acc);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j < cnt) acc[j] = wave_sum(acc[j], m);  // cnt is uniform: whole waves skip
        if ((t_ & 63) == 0)
#pragma unroll
            for (int j = 0; j < 8; ++j) red[(t_ >> 6) * 8 + j] = acc[j];
    }
    // A phase of its own: the wave sums of f's partials for every virtual thread.
    template <class F>
    __device__ __forceinline__ void wave_reduce8(double* red, F f, const Mod& m) {
#pragma unroll
        for (int i = 0; i < VT; ++i) wave_partial8(red, th[i], t + i * TP, f, m);
        __syncthreads();
    }
};

// The task's context from its descriptor; arena_base is where its arena offset counts from.
template <int LG>
__device__ __forceinline__ Ctx make_ctx(const TaskDesc* d, const QueryDesc* qs, const int* slots, double* arena_base) {
    Ctx c;
    c.d = d;
    c.m.p = d->p;
    c.m.pinv = d->pinv;
    c.a = arena_of(arena_base, *d);
    c.slots = slots + d->slot_off;
    c.q = qs;
    c.model = d->model;
    c.dd = d->dd;
    c.v0 = d->model ? 1 : d->dd;
    c.es = Plan<LG>::T;
    c.eb = 0;
    c.L = Plan<LG>::L;
    return c;
}

// The task kernel: VT virtual threads per hardware thread (Plan<LG>::T / VT threads, so
// Plan<LG>::T / VT / 64 waves per transform), DF the tile-flush choice of Task.
template <int LG, int VT, int DF>
__global__ void __launch_bounds__(Plan<LG>::T / VT)
    k_tasks(const TaskDesc* __restrict__ tasks, const QueryDesc* __restrict__ qs, const int* __restrict__ slots,
            double* __restrict__ mem, uint64_t* __restrict__ res) {
    constexpr int TP = Plan<LG>::T / VT;
    __shared__ Shared<LG> sh;
    DevRunner<VT, TP> run;
    run.t = (int)threadIdx.x;
    const Ctx c = make_ctx<LG>(tasks + blockIdx.x, qs, slots, mem);
    Task<LG, LG, DF>::run_task(run, sh, c);
    for (int j = threadIdx.x; j < c.d->q_cnt; j += TP)
        res[qs[c.d->q_off + j].out] = (uint64_t)Task<LG, LG, DF>::result(j, c);
}

template <int LG, int VT, int DF>
hipError_t launch_k(int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots, double* mem, uint64_t* res,
                    hipStream_t stream) {
    k_tasks<LG, VT, DF><<<dim3((unsigned)ntasks), Plan<LG>::T / VT, 0, stream>>>(tasks, qs, slots, mem, res);
    return hipGetLastError();
}

template <int LG, int VT, int DF>
int resident_k(int device) {
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, device) != hipSuccess) return 0;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_tasks<LG, VT, DF>, Plan<LG>::T / VT, 0) != hipSuccess) return 0;
    return pr.multiProcessorCount * (per > 0 ? per : 1);
}

}  // namespace cf
