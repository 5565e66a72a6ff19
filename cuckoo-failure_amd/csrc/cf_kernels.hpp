// cf_kernels.hpp -- host-visible launcher of the task kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "cf_core.hpp"

namespace cf {

constexpr int kMinLog2L = 10, kMaxLog2L = 14;

// Compiled variants of the task kernel per transform length (variant 0 the built-in choice):
// waves per transform (virtual threads per hardware thread) and whether the tile flush of a
// column block is deferred into the next block's transform rounds.
struct KernelVariant {
    int waves;
    int deferred;
};
int variant_count(int log2L);
KernelVariant variant_info(int log2L, int variant);

// Workgroups of the log2L kernel (variant) resident on the device at once (0 if unknown).
int resident_tasks(int log2L, int device, int variant = 0);

// One workgroup per task; log2L selects the kernel instance (transform length 2^log2L).
hipError_t launch_tasks(int log2L, int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots, double* mem,
                        uint64_t* res, hipStream_t stream, int variant = 0);
// The variants other than 0 (cf_variants.hip).
int resident_variant(int log2L, int variant, int device);
hipError_t launch_variant(int log2L, int variant, int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots,
                          double* mem, uint64_t* res, hipStream_t stream);

// The work-queue kernel: jobs of tasks with log2 L in [kMinLog2L, kQueueMaxLog2L], any mix,
// ordered as they should start. A job is up to queue_slots(lg, threads) consecutive tasks of
// identical shape that one workgroup runs side by side, each on L/16 of its threads.
// `workgroups` persistent workgroups; workgroup w uses the arena at mem + w * arena_stride
// (doubles), slot g of an lg job its part g * arena_stride / queue_slots(lg, threads). *next must
// be 0 at launch (the queue position). Compiled variants (the autotuner of the batched form picks
// one, cf_autotune_batch; variant 0 is the built-in choice) differ in the workgroup's threads and
// the workgroups per CU its register budget is compiled for.
constexpr int kQueueMaxLog2L = 12;
constexpr int kQueueVariants = 3;
struct QueueVariant {
    int threads;  // per workgroup: one 2^12 task per 256
    int per_cu;   // workgroups per CU the register budget is compiled for
};
QueueVariant queue_variant_info(int variant);
inline int queue_slots(int lg, int threads) { return threads / ((1 << lg) / 16); }
struct QueueJob {
    int first, cnt, lg, pad_;
};
int queue_resident(int device, int variant = 0);
hipError_t launch_queue(int variant, int workgroups, int njobs, const QueueJob* jobs, const TaskDesc* tasks, const QueryDesc* qs,
                        const int* slots, double* mem, int64_t arena_stride, uint64_t* res, unsigned* next, hipStream_t stream);

// Tasks with transforms of 2^16 or 2^17 (log2L), each spread over 16 workgroups (of 256 or 512
// threads) that meet at group barriers: one cooperative launch of ntasks tasks (all workgroups
// resident at once). sync holds
// wide_sync_words(ntasks) zeros at launch; afterwards sync[2 ntasks] != 0 means a barrier timed
// out (results invalid), and sync[2 ntasks + 2 + 16 k + g] is the XCD that workgroup g of task k
// ran on. local != 0 lets a task whose workgroups share one XCD use XCD-local barriers (its L2 is
// then the common cache level).
constexpr int kWideLog2 = 16, kWideMaxLog2 = 17, kWideMaxTasks = 16, kWideMembers = 16;
constexpr size_t wide_sync_words(int ntasks) { return 2 * (size_t)ntasks + 2 + kWideMembers * (size_t)ntasks; }
int wide_resident(int log2L, int device);  // workgroups of the wide kernel the device holds at once
hipError_t launch_wide(int log2L, int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots, double* mem, uint64_t* res,
                       unsigned* sync, int local, hipStream_t stream);

}  // namespace cf
