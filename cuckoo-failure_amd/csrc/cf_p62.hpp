// cf_p62.hpp -- the suffix-table recurrence modulo primes of up to 63 bits (cf_p62.hip): the
// task descriptor shared by the host runner (cf_p62_host.cpp) and the kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cf_core.hpp"

namespace cf {

constexpr int kP62Threads = 1024;  // bound on the kernel's workgroup size (it uses at most 512 threads)
constexpr int kP62MaxLog2L = 17;   // tables of up to 131072 rows (transforms in LDS up to 2^14, in HBM beyond)

// Montgomery arithmetic modulo an odd p < 2^63 with R = 2^64.
struct Mod62 {
    uint64_t p;
    uint64_t pinv;  // p^-1 mod 2^64
    uint64_t one;   // R mod p, the Montgomery form of 1
    uint64_t r2;    // R^2 mod p
};

struct Task62 {
    Mod62 m;
    uint64_t linv;       // 1/L, Montgomery form
    int64_t arena;       // offset (words) of the task's memory
    int nm, F1, L, NS;   // table rows 0..nm, factorials 0..F1-1, transform length, staging rows
    int q_off, q_cnt;    // queries, sorted by n (largest first)
    int dd, model;
    int slot_off, NSEG;  // the tile-slot plan of nm (plan_slots): offset in the launch's slot array, segments per row block
    int64_t pool;        // tiles per table
};

// Words of one task's arena: five factorial-range tables, three of length L, 32 Shoup constants, the two tile pools
// (pool tiles of 256 words each), the staging of both tables (32 columns of NS rows), the query
// state and the transform's buffer when it does not fit in LDS (L > 2^14).
int64_t arena62_words(int NS, int64_t pool, int F1, int L, int q_cnt);
hipError_t launch_p62(int log2L, int ntasks, const Task62* tasks, const QueryDesc* qs, const int* slots, uint64_t* mem,
                      uint64_t* res, hipStream_t stream);

}  // namespace cf
