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

struct DevRunner {
    Thread* th;
    int t;
    template <class F>
    __device__ __forceinline__ void phase(F f) {
        f(*th, t);
        __syncthreads();
    }
    template <class F>
    __device__ __forceinline__ void wave_reduce8(double* red, F f, const Mod& m) {
        double acc[8];
        f(*th, t, acc);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = wave_sum(acc[j], m);
        if ((t & 63) == 0)
#pragma unroll
            for (int j = 0; j < 8; ++j) red[(t >> 6) * 8 + j] = acc[j];
        __syncthreads();
    }
};

template <int LG>
__global__ void __launch_bounds__(Plan<LG>::T) k_tasks(const TaskDesc* __restrict__ tasks, const QueryDesc* __restrict__ qs,
                                                       const int* __restrict__ slots, double* __restrict__ mem,
                                                       uint64_t* __restrict__ res, int model, int dd) {
    __shared__ Shared<LG> sh;
    Thread th;
    DevRunner run{&th, (int)threadIdx.x};
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
    for (int j = threadIdx.x; j < c.d->q_cnt; j += Plan<LG>::T) res[qs[c.d->q_off + j].out] = (uint64_t)Task<LG>::result(j, c);
}

hipError_t launch_tasks(int log2L, int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots, double* mem,
                        uint64_t* res, int model, int dd, hipStream_t stream) {
    if (ntasks <= 0) return hipSuccess;
    const dim3 grid((unsigned)ntasks);
    switch (log2L) {
        case 10: k_tasks<10><<<grid, Plan<10>::T, 0, stream>>>(tasks, qs, slots, mem, res, model, dd); break;
        case 11: k_tasks<11><<<grid, Plan<11>::T, 0, stream>>>(tasks, qs, slots, mem, res, model, dd); break;
        case 12: k_tasks<12><<<grid, Plan<12>::T, 0, stream>>>(tasks, qs, slots, mem, res, model, dd); break;
        case 13: k_tasks<13><<<grid, Plan<13>::T, 0, stream>>>(tasks, qs, slots, mem, res, model, dd); break;
        case 14: k_tasks<14><<<grid, Plan<14>::T, 0, stream>>>(tasks, qs, slots, mem, res, model, dd); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cf
