// cf_serve.hpp -- launcher of the query kernels on resident tables (cf_serve.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cf {

constexpr int kServeThreads = 512;  // 8 waves per workgroup, one row range each
constexpr int kServeJ = 32;         // queries per pass (one accumulator each per lane)
// The query kernel keeps 1/i for i < F1 = max(nmax, Mmax) + 1, behind 512 zero entries, in LDS
// (160 KB per CU): table sets with a larger F1 are refused when they are built.
constexpr int kServeLdsPad = 512;
constexpr int kServeMaxF1 = 163840 / 8 - kServeLdsPad;

struct ServePrime {
    double p, pinv;
};

struct ServeQuery {
    int n, M;
    int need;  // primes 0 .. need-1 of the table set answer this query
    int pad_;
};

struct ServeArgs {
    const double* store;     // resident stores, stride doubles apart
    int64_t stride;
    const ServePrime* primes;
    int np, F1;              // primes; factorial tables 0 .. F1-1
    int dd, model, v0;       // the table set's d and model; rows v <= v0 are zero
    const ServeQuery* q;     // the batch
    int J;
    const int* bounds;       // per pass z: row ranges (bounds[z (ranges+1) + r], ... + r + 1]
    int ranges;              // ranges per pass: groups x the waves per workgroup
    int passes, groups;      // passes = ceil(J / kServeJ)
    double* part;            // np x ranges x J partial sums
};

// res[b J + j] = N_bad(n_j, M_j) mod p_b for b < need_j, else 0. The queries must be sorted by
// n, largest first.
hipError_t launch_serve(const ServeArgs& a, uint64_t* res, hipStream_t stream);

// The fixed device scratch of cf_query_noalloc: one statically sized array per device, holding a
// call's query descriptors, row-range bounds, partial sums and result block. Every part a call
// reads, it has written first in the same call, so nothing needs clearing between calls.
constexpr size_t kQueryScratchBytes = size_t(32) << 20;
constexpr int kQueryScratchDevices = 64;
// Looks up the array's address on `device`, which must be the current device; once per device
// (later calls return at once). Called when an engine is opened.
hipError_t attach_query_scratch(int device);
// The address found for `device`, or nullptr if it was never attached.
void* query_scratch(int device);

}  // namespace cf
