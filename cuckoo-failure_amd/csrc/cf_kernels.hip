// cf_kernels.hip -- the gfx950 task kernel: one workgroup runs one prime's whole suffix-table
// build and query reduction (cf_task.hpp), L/16 threads with 16 transform entries each, the
// transform buffer, twiddle and power tables in LDS.
#include <hip/hip_runtime.h>

#include "cf_kernels.hpp"
#include "cf_task.hpp"

namespace cf {

// Sum over the 64 lanes of a wavefront, reduced after every level (each level at most doubles).
__device__ __forceinline__ double wave_sum(double x, const Mod& m) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x = reduce(x + __shfl_xor(x, o, 64), m);
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
    // spilled registers) across all steps.
    template <class F>
    __device__ __forceinline__ void phase(F f) {
#pragma unroll
        for (int i = 0; i < VT; ++i) {
            int tt = t + i * TP;
            asm volatile("" : "+v"(tt));
            f(th[i], tt);
        }
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
    // Inside a phase: this wave's sum of x, written by lane 0 to red[wave].
    __device__ __forceinline__ void wave_partial1(double* red, Thread&, int t_, double x, const Mod& m) {
        x = wave_sum(x, m);
        if ((t_ & 63) == 0) red[t_ >> 6] = x;
    }
    // A phase of its own: the wave sums of f's partials for every virtual thread.
    template <class F>
    __device__ __forceinline__ void wave_reduce8(double* red, F f, const Mod& m) {
#pragma unroll
        for (int i = 0; i < VT; ++i) wave_partial8(red, th[i], t + i * TP, f, m);
        __syncthreads();
    }
};

template <int LG>
constexpr int kVirtual = LG >= 14 ? 2 : 1;

template <int LG>
__global__ void __launch_bounds__(Plan<LG>::T / kVirtual<LG>)
    k_tasks(const TaskDesc* __restrict__ tasks, const QueryDesc* __restrict__ qs, const int* __restrict__ slots,
            double* __restrict__ mem, uint64_t* __restrict__ res, int model, int dd) {
    constexpr int VT = kVirtual<LG>, TP = Plan<LG>::T / VT;
    __shared__ Shared<LG> sh;
    DevRunner<VT, TP> run;
    run.t = (int)threadIdx.x;
    Ctx c;
    c.d = tasks + blockIdx.x;
    c.m.p = c.d->p;
    c.m.pinv = c.d->pinv;
    c.a = arena_of(mem, *c.d);
    c.slots = slots + c.d->slot_off;
    c.q = qs;
    c.model = model;
    c.dd = dd;
    c.v0 = model ? 1 : dd;
    Task<LG>::run_task(run, sh, c);
    for (int j = threadIdx.x; j < c.d->q_cnt; j += TP) res[qs[c.d->q_off + j].out] = (uint64_t)Task<LG>::result(j, c);
}

// Workgroups of the log2L kernel that the device runs at the same time (CUs x blocks per CU).
int resident_tasks(int log2L, int device) {
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, device) != hipSuccess) return 0;
    int per = 0;
    hipError_t st = hipErrorInvalidValue;
    switch (log2L) {
        case 10: st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_tasks<10>, Plan<10>::T / kVirtual<10>, 0); break;
        case 11: st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_tasks<11>, Plan<11>::T / kVirtual<11>, 0); break;
        case 12: st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_tasks<12>, Plan<12>::T / kVirtual<12>, 0); break;
        case 13: st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_tasks<13>, Plan<13>::T / kVirtual<13>, 0); break;
        case 14: st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_tasks<14>, Plan<14>::T / kVirtual<14>, 0); break;
        default: break;
    }
    if (st != hipSuccess) return 0;
    return pr.multiProcessorCount * (per > 0 ? per : 1);
}

hipError_t launch_tasks(int log2L, int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots, double* mem,
                        uint64_t* res, int model, int dd, hipStream_t stream) {
    if (ntasks <= 0) return hipSuccess;
    const dim3 grid((unsigned)ntasks);
    switch (log2L) {
        case 10: k_tasks<10><<<grid, Plan<10>::T / kVirtual<10>, 0, stream>>>(tasks, qs, slots, mem, res, model, dd); break;
        case 11: k_tasks<11><<<grid, Plan<11>::T / kVirtual<11>, 0, stream>>>(tasks, qs, slots, mem, res, model, dd); break;
        case 12: k_tasks<12><<<grid, Plan<12>::T / kVirtual<12>, 0, stream>>>(tasks, qs, slots, mem, res, model, dd); break;
        case 13: k_tasks<13><<<grid, Plan<13>::T / kVirtual<13>, 0, stream>>>(tasks, qs, slots, mem, res, model, dd); break;
        case 14: k_tasks<14><<<grid, Plan<14>::T / kVirtual<14>, 0, stream>>>(tasks, qs, slots, mem, res, model, dd); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cf
