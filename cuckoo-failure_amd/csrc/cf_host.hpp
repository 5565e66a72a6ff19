// cf_host.hpp -- host-side logic of the engine: the reference's prime lists, need() and transform
// length, per-prime constants for the device, and the static tile-slot plan of the suffix tables.
// Pure C++ (no HIP): shared by the engine library and its CPU emulation build.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "cf_core.hpp"

namespace cf {

// ---- the reference's host logic (fast/cuckoo_fast.cu:391-472; include/cuckoo/modint.hpp:137-196) ----
uint64_t mulmod_u64(uint64_t a, uint64_t b, uint64_t p);
uint64_t powmod_u64(uint64_t a, uint64_t e, uint64_t p);
bool is_prime_u64(uint64_t n);
// primes number first .. first+count-1 of the list "i-th largest c 2^20 + 1 below 2^bits"
std::vector<uint64_t> prime_range(int prime_bits, long first, long count);
double log2_q(int M, int d);                  // log2 C(M, d), long double sum
double log2_q_replacement(int M, int d);      // log2 M^d
long need(int n, int M, int d, int prime_bits);
size_t transform_length(size_t nmax);         // max(16, next_pow2(nmax))
std::vector<uint32_t> ntt_primes32(int count, int min_pow2);  // throws if the list runs out
int two_adic_valuation(uint64_t x);

// ---- device plan ----
// log2 of the engine's transform length for a table of nm rows modulo p: the smallest power of
// two >= max(nm, 1024), with 2^16 in place of 2^15; -1 when p - 1 lacks that power of two.
int engine_log2_length(int nm, uint64_t p);
// Constants of one prime for the device (roots, 1/L, the DFT16 roots), balanced doubles.
void task_constants(uint64_t p, int log2L, TaskDesc& d);
// Constants of a task with transforms of 2^log2L (16 or 17) spread over 16 workgroups (inner
// length 2^(log2L - 4)).
void wide_constants(uint64_t p, int log2L, TaskDesc& d);

// Static assignment of the suffix-table tiles to pool memory. A tile holds 16 rows x 16 columns
// (row-major). Tile (Y, X), X < Y, holds rows 16Y..16Y+15 and columns 16X..16X+15; it is written
// when column block X is flushed (end of step 16X + 15) and dead once row block Y has been read
// (end of step 16Y + 15). Tiles are grouped into segments of 16 consecutive column blocks of one
// row block: segment (Y, s) holds tiles (Y, 16s .. 16s+15) contiguously (32 KB), is allocated
// when its first tile is written and freed with its row block. Freed segments are reused (last
// freed first), so the pool holds about nm^2/4 entries per table instead of nm^2/2, a row of the
// table lies in at most NSEG segments, and the segment numbers of one row block (NSEG ints) fit
// in LDS.
struct SlotPlan {
    int nm = 0, NB = 0, NS = 0, NSEG = 0;
    int pool = 0;            // tiles (16 per segment)
    std::vector<int> segs;   // segment slot of (Y, s) at Y * NSEG + s, -1 if never allocated
};
SlotPlan plan_slots(int nm);

// Device bytes of one task (both tables, staging, factorial tables, query state, spectrum).
size_t task_bytes(const SlotPlan& sp, int F1, int q_cnt, int log2L);

}  // namespace cf
