// cf_kernels.hpp -- host-visible launcher of the task kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "cf_core.hpp"

namespace cf {

constexpr int kMinLog2L = 10, kMaxLog2L = 14;

// Workgroups of the log2L kernel resident on the device at once (0 if unknown).
int resident_tasks(int log2L, int device);

// One workgroup per task; log2L selects the kernel instance (transform length 2^log2L).
hipError_t launch_tasks(int log2L, int ntasks, const TaskDesc* tasks, const QueryDesc* qs, const int* slots, double* mem,
                        uint64_t* res, int model, int dd, hipStream_t stream);

}  // namespace cf
