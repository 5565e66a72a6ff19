/*
 * oracle.h -- CPU restatement of the reference's one-table suffix-table algorithm.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls this library; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only as the
 * checker. Arithmetic is deliberately different from the product: canonical residues in [0, p)
 * with a 128-bit product and a hardware division (a * b % p), textbook radix-2 NTTs with a
 * bit-reversal permutation and a transform long enough that nothing wraps.
 *
 * Every function cites the reference file:line whose behaviour it restates (paths relative to
 * the reference root).
 */
#ifndef CF_ORACLE_H
#define CF_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Deterministic Miller-Rabin for n < 2^64 (fast/cuckoo_fast.cu:403-421). */
int or_is_prime_u64(uint64_t n);
/* Primes number first .. first+count-1 of the fixed list: the i-th largest p = c*2^20 + 1 below
 * 2^prime_bits (fast/cuckoo_fast.cu:423-432). Returns how many were written. */
long or_prime_range(int prime_bits, long first, long count, uint64_t* out);
/* log2 C(M, d) as the reference computes it (long double sum, rounded to double) and the number
 * of primes needed (fast/cuckoo_fast.cu:452-459). */
double or_log2_q(int M, int d);
long or_need(int n, int M, int d, int prime_bits);
/* max(16, next_pow2(nmax)) (fast/cuckoo_fast.cu:465-472). */
size_t or_transform_length(size_t nmax);

/* 32-bit NTT-friendly primes, largest first, 2-adic valuation exactly k for k = 31 .. min_pow2
 * (include/cuckoo/modint.hpp:183-196). Returns count, or -1 if the list ran out. */
int or_ntt_primes32(int count, int min_pow2, uint32_t* out);

/* Suffix table V[p][w], 0 <= w < p <= nm, as a packed triangle (row p at p(p-1)/2), by the
 * O(n^2 log n) two-stage convolution (include/cuckoo/single_table.hpp:178-229).
 * rep = 0: g(v) = C(v, d); rep = 1: g(v) = v^d (single_table.hpp:16-23). 0 on success. */
int or_suffix_table_ntt(int nm, int d, int rep, uint64_t p, uint64_t* V);
/* The same table by the O(n^3) shared-inner-sum recurrence (single_table.hpp:135-166). */
int or_suffix_table_cubic(int nm, int d, int rep, uint64_t p, uint64_t* V);

/* N_bad mod p for each query from a table built at nm = max n, evaluated with the paper's double
 * sum (single_table.hpp:103-119). algo 0: ntt table, 1: cubic table. */
int or_nbad(int algo, int d, int rep, uint64_t p, int nq, const int* n, const int* M, uint64_t* out);
/* N_bad(n, M) mod p by Algorithm 1, the O(n^4) forward recurrence over block boundaries
 * (include/cuckoo/single_table.hpp:25-31, 44-78). */
int or_nbad_forward(int d, int rep, uint64_t p, int n, int M, uint64_t* out);
/* Separate tables (include/cuckoo/separate_tables.hpp): d tables of lengths m[0..d), one slot per
 * table per item. or_separate_nbad builds the suffix table V[p][v] over v >= 1, |v| < p <= n from
 * its defining sums (:175-186) and evaluates N_bad (:190-213); or_separate_forward is Algorithm 4,
 * the forward recurrence over (l, b) (:27-108). */
int or_separate_nbad(int n, const int* m, int d, uint64_t p, uint64_t* out);
int or_separate_forward(int n, const int* m, int d, uint64_t p, uint64_t* out);
/* N_bad mod p evaluated with the fast engine's query rewrite and Horner finish
 * (fast/cuckoo_fast.cu:306-323, 877-890) on the ntt table. */
int or_nbad_horner(int d, int rep, uint64_t p, int nq, const int* n, const int* M, uint64_t* out);

/* Brute force over multisets of candidate sets with Kuhn matching
 * (include/cuckoo/brute_force.hpp:22-42, 70-146). M <= 62, Q^n < 2^64. 0 on success. */
int or_brute_force(int n, int M, int d, int rep, unsigned long long* nbad, unsigned long long* total);

/* Cyclic convolution of a and k (length L each) mod p by the schoolbook O(L^2) sum. */
void or_cyclic_conv_schoolbook(uint64_t p, size_t L, const uint64_t* a, const uint64_t* k, uint64_t* out);
/* The same by a textbook NTT (requires L | p - 1). 0 on success. */
int or_cyclic_conv_ntt(uint64_t p, size_t L, const uint64_t* a, const uint64_t* k, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif
