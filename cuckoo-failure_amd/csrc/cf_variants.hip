// cf_variants.hip -- the task-kernel variants other than each length's built-in choice, timed by
// the autotuner (cf_autotune): two virtual threads per hardware thread (half the waves per
// transform, twice the independent work per wave) at 2^11 .. 2^13, and the immediate tile flush
// at 2^13 and 2^14. Same arithmetic, same results; only the schedule differs.
#include <hip/hip_runtime.h>

#include "cf_task_kernel.hpp"

namespace cf {

int resident_variant(int log2L, int variant, int device) {
    switch (log2L * 4 + variant) {
        case 11 * 4 + 1: return resident_k<11, 2, -1>(device);
        case 12 * 4 + 1: return resident_k<12, 2, -1>(device);
        case 13 * 4 + 1: return resident_k<13, 2, -1>(device);
        case 13 * 4 + 2: return resident_k<13, 1, 0>(device);
        case 14 * 4 + 1: return resident_k<14, 2, 0>(device);
        default: return 0;
    }
}

hipError_t launch_variant(int log2L, int variant, int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots,
                          double* mem, uint64_t* res, hipStream_t stream) {
    switch (log2L * 4 + variant) {
        case 11 * 4 + 1: return launch_k<11, 2, -1>(ntasks, tasks, qs, slots, mem, res, stream);
        case 12 * 4 + 1: return launch_k<12, 2, -1>(ntasks, tasks, qs, slots, mem, res, stream);
        case 13 * 4 + 1: return launch_k<13, 2, -1>(ntasks, tasks, qs, slots, mem, res, stream);
        case 13 * 4 + 2: return launch_k<13, 1, 0>(ntasks, tasks, qs, slots, mem, res, stream);
        case 14 * 4 + 1: return launch_k<14, 2, 0>(ntasks, tasks, qs, slots, mem, res, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cf
