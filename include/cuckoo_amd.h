/*
 * cuckoo_amd.h -- C ABI of the MI355X (gfx950) one-table failure-count engine.
 *
 * The reference (zcy05331/Cuckoo-Failure) has no library API for this path: its seams are the
 * in-source back end of fast/cuckoo_fast.cu (launch / dev_alloc / to_dev / from_dev / dev_sync /
 * dev_free_bytes at :325-386, convolve at :633-646), the v-loop and query reduction driven from
 * main (:787-900), and the 32-bit library calls build_one_table_suffix_ntt(nmax, model) and
 * OneTableSuffix::nbad(n, M) (include/cuckoo/single_table.hpp:97-120, 178-229). Each entry point
 * below names the reference code it replaces.
 *
 * Conventions: plain C types; caller-owned host buffers; the engine owns all device memory.
 * Residues cross the ABI in canonical form [0, p). Integer return codes map to the reference's
 * exit codes (CF_E_ARGS -> 2, CF_E_NOMEM -> 3, CF_E_PRIMES -> 4, CF_E_DEVICE -> 1); no exception
 * crosses the ABI. One engine per host thread.
 */
#ifndef CUCKOO_AMD_H
#define CUCKOO_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CF_OK 0
#define CF_E_DEVICE 1      /* HIP runtime or kernel failure (cuda_check, fast/cuckoo_fast.cu:329-334) */
#define CF_E_ARGS 2        /* invalid arguments (fast/cuckoo_fast.cu:654-695) */
#define CF_E_NOMEM 3       /* not enough device memory for one prime (fast/cuckoo_fast.cu:751-753) */
#define CF_E_PRIMES 4      /* prime list exhausted (fast/cuckoo_fast.cu:729-732) */
#define CF_E_UNSUPPORTED 5 /* a shape this build does not run (transform length beyond the LDS kernels,
                              or p - 1 without the needed power of two) */

typedef struct cf_engine cf_engine;

/* Timing and size of the last cf_run_fused call. kernel_ms is measured with HIP events on the
 * engine's stream around the task kernels only; wall_ms covers the whole call. */
typedef struct {
    double kernel_ms;
    double wall_ms;
    int launches;
    int tasks;
    int batches;
    int log2_length; /* log2 of the longest transform of the call */
    size_t device_bytes;
    int graph;       /* 0: launched one by one; 1: captured into a hipGraph and run; 2: graph replayed */
    int wide_split;  /* tables of more than 16384 rows whose 16 workgroups ran on more than one XCD
                        (device-scope group barriers instead of XCD-local ones) */
} cf_stats;

/* Replaces the CUDA back end's device set-up (fast/cuckoo_fast.cu:328-359): selects a HIP device
 * and creates the engine's stream. */
int cf_open(int device, cf_engine** out);
void cf_close(cf_engine* e);
const char* cf_strerror(int code);
/* Human-readable detail of the last error on this engine. */
const char* cf_last_error(const cf_engine* e);
int cf_last_stats(const cf_engine* e, cf_stats* out);
/* Device name and architecture, e.g. "AMD Instinct MI355X gfx950". */
int cf_device_name(const cf_engine* e, char* buf, size_t len);
/* Graph mode for cf_run_fused / cf_run_batch (off by default). On, a call's descriptor upload, its
 * kernel launches and the result download are captured into one hipGraph (no memset nodes; the
 * queue counter is reset by the upload) and the graph is replayed by later calls whose launch
 * plan has the same signature (same shapes and counts, new primes), with only the pinned staging
 * rewritten: the reference's per-batch host-precompute -> build -> query -> finish chain
 * (fast/cuckoo_fast.cu:787-898) as one submission. Calls with 2^16 transforms (cooperative
 * launches) always run one launch at a time. */
int cf_set_graphs(cf_engine* e, int on);

/* Builds the suffix table for each prime and answers the queries on it: the reference's per-batch
 * work of fast/cuckoo_fast.cu:787-896 (host precompute, kernel spectrum, the v-loop of stage-1 /
 * stage-2 convolutions, the query reduction and the Horner finish), and equally the 32-bit path's
 * build_one_table_suffix_ntt + nbad (include/cuckoo/single_table.hpp:103-119, 178-229).
 *   d        candidate slots per item (d >= 1)
 *   model    0: uniform d-subset, g(v) = C(v, d), Q = C(M, d); 1: d draws with replacement,
 *            g(v) = v^d, Q = M^d (single_table.hpp:16-23)
 *   n, M     J queries (1 <= n < M)
 *   primes   np odd primes p < 2^63 (each with M < p); those above 2^50 (the reference's 62-bit
 *            list of its -DFP_RB=64 build) run on a 64-bit Montgomery kernel, n <= 131072
 *   mask     np x J bytes, row-major; query j is evaluated for prime b iff mask[b J + j] != 0.
 *            NULL means all.
 *   res      np x J canonical residues N_bad(n_j, M_j) mod p_b; 0 where the mask is 0.
 * Each prime's table has as many rows as the largest n it is asked for. Returns CF_OK or an error
 * code; cf_last_error() has the detail. */
int cf_run_fused(cf_engine* e, int d, int model, const int32_t* n, const int32_t* M, size_t J,
                 const uint64_t* primes, size_t np, const uint8_t* mask, uint64_t* res);

/* One instance of an uneven batch: the query N_bad(n, M) of a d-choice table, modulo each of
 * primes[prime_off .. prime_off + prime_cnt), with results at the same indices of res. */
typedef struct {
    int32_t d;      /* candidate slots per item (d >= 1) */
    int32_t model;  /* 0: uniform d-subset; 1: d draws with replacement */
    int32_t n, M;   /* 1 <= n < M */
    int64_t prime_off, prime_cnt;
} cf_instance;

/* Uneven batch: k instances of different n, M and d in one persistent work-queue launch
 * (transforms up to 2^12, i.e. tables of up to 4096 rows; longer ones in launches of their own),
 * longest table first. Instances with the same d and model that name the same prime share that
 * prime's table: one table of max n rows answers all their queries, as one cuckoo_fast run
 * answers every line of its query file from one table per prime (fast/cuckoo_fast.cu:787-876;
 * rows below n do not depend on the rows beyond). The reference runs such a batch as one
 * cuckoo_fast process per instance (fast/cuckoo_fast.cu:653-898 with D, QUERIES = "n M",
 * FIRST = 0, COUNT = need); the residues are the same, res[i] = N_bad mod primes[i]. Instances
 * must not share prime/result indices; indices no instance covers are left untouched. */
int cf_run_batch(cf_engine* e, const cf_instance* inst, size_t k, const uint64_t* primes, uint64_t* res);
/* Table sharing in cf_run_batch: 1 (the default) as above, 0 one table per (instance, prime).
 * The residues do not depend on it. */
int cf_set_batch_sharing(cf_engine* e, int on);

/* The reference's direct one-table algorithms on the device, one workgroup per prime, as
 * cross-checks of the suffix-table path: res[b] = N_bad(n, M) mod primes[b] for model 0 (d-subsets)
 * or 1 (with replacement). CF_ALGO_FORWARD is one_table_forward, the O(n^4) boundary recurrence
 * (include/cuckoo/single_table.hpp:44-78), n <= 256; CF_ALGO_CUBIC is
 * build_one_table_suffix_cubic + OneTableSuffix::nbad (single_table.hpp:135-166, 104-120),
 * n <= 1024 (CF_E_UNSUPPORTED beyond). Selected by `cuckoo_fail single --algo forward|cubic`
 * (src/cuckoo_fail.cpp:60-87). Primes odd, M < p < 2^50. */
#define CF_ALGO_FORWARD 1
#define CF_ALGO_CUBIC 2
int cf_run_algo(cf_engine* e, int algo, int d, int model, int n, int M, const uint64_t* primes, size_t np,
                uint64_t* res);

/* The separate-tables model (include/cuckoo/separate_tables.hpp): d tables of lengths m[0..d),
 * each item one slot per table, Q = prod m_t. res[b] = N_bad(n, m) mod primes[b], one workgroup
 * per prime. CF_ALGO_SUFFIX: Algorithm 5, the suffix table over the states v >= 1, |v| < p <= n,
 * with the inner sums by Horner's rule in sigma = x_1 + .. + x_d (separate_tables.hpp:175-291,
 * build_separate_suffix + SeparateSuffix::nbad), up to C(n-1, d) <= 4e6 states; CF_ALGO_FORWARD:
 * Algorithm 4, the recurrence over the boundary states (separate_tables.hpp:27-108,
 * separate_forward), up to 2^20 of them. Selected by `cuckoo_fail separate` (src/cuckoo_fail.cpp:
 * 83-99). 1 <= d <= 8, m_t >= 1, n < M = sum m_t; primes odd, M < p < 2^50. */
#define CF_ALGO_SUFFIX 3
int cf_run_separate(cf_engine* e, int algo, int n, const int32_t* m, int d, const uint64_t* primes, size_t np,
                    uint64_t* res);

/* The autotuner the batched forms use (no counterpart in the reference, whose CUDA build has one
 * fixed schedule per length, fast/cuckoo_fast.cu:510-646). Each transform-length class 2^10 ..
 * 2^14 has compiled variants of the per-prime kernel that differ in waves per transform (one or
 * two virtual threads per hardware thread) and in the tile flush (immediate, or deferred into the
 * next column block's rounds); variant 0 is the built-in choice. cf_autotune times every variant
 * of one class on a calibration batch (tables of n rows, M = 2n, d = 3, two waves of workgroups;
 * n = n_cal, which must lie in the class (L/2, L], or L/2 + 1 for n_cal = 0), checks that all give
 * identical residues, and makes the fastest the engine's choice for that class in later calls.
 * The best variant can depend on n within a class: tune on the sizes that will run. The
 * work-queue kernel of cf_run_batch has variants of its own (cf_autotune_batch). */
typedef struct {
    int log2_length;  /* the class */
    int variants;     /* how many variants were timed */
    int chosen;       /* the fastest, now used by this engine */
    int tables;       /* calibration tables per variant */
    int n;            /* their rows */
    int waves[4];     /* waves per transform */
    int deferred[4];  /* 1: deferred tile flush */
    int resident[4];  /* workgroups resident on the device at once */
    double ms[4];     /* kernel time of the calibration batch */
} cf_tune;
int cf_autotune(cf_engine* e, int log2_length, int n_cal, cf_tune* out);
/* Selects a class's variant by hand (0 = the built-in choice). */
int cf_set_variant(cf_engine* e, int log2_length, int variant);

/* The autotuner of the batched form (cf_run_batch; no counterpart in the reference, which runs one
 * process per instance with one fixed schedule, fast/cuckoo_fast.cu:510-646). The work-queue
 * kernel is compiled in variants that differ in threads per workgroup (how many tasks of one length
 * share a workgroup's barriers) and in the workgroups per CU its register budget is compiled for;
 * variant 0 is the built-in choice. cf_autotune_batch times every variant on a sample of the
 * caller's own batch (each instance with the same share of its primes, at least one, about
 * sample_tables tables in all; 0 = the whole batch), requires identical residues from all of
 * them, and keeps the fastest for this engine's later cf_run_batch calls. */
typedef struct {
    int variants;     /* how many were timed */
    int chosen;       /* the fastest, now used by this engine */
    int tables;       /* tables in the sample */
    int threads[4];   /* threads per workgroup */
    int per_cu[4];    /* workgroups per CU the register budget is compiled for */
    int resident[4];  /* workgroups resident on the device at once */
    double ms[4];     /* kernel time of the sample */
} cf_batch_tune;
int cf_autotune_batch(cf_engine* e, const cf_instance* inst, size_t k, const uint64_t* primes, size_t sample_tables,
                      cf_batch_tune* out);
/* Selects the work-queue variant by hand (0 = the built-in choice). */
int cf_set_batch_variant(cf_engine* e, int variant);

/* ---- resident tables and serving ---- */
typedef struct cf_tables cf_tables;

/* Builds, and keeps in device memory, the one-table suffix tables of every prime (rows
 * v = 0 .. nmax, factorials up to max(nmax, Mmax)): the reference's build_one_table_suffix_ntt
 * (include/cuckoo/single_table.hpp:103-119) / the v-loop of fast/cuckoo_fast.cu:787-857, kept for
 * later queries instead of being consumed. The set is admitted only if it fits in the free device
 * memory (CF_E_NOMEM otherwise). Requires Mmax > nmax >= 1 and Mmax < every prime < 2^50. */
int cf_tables_build(cf_engine* e, int d, int model, int nmax, int Mmax, const uint64_t* primes, size_t np,
                    cf_tables** out);
/* Frees a set. Servers on it must be closed first; its engine must still be open while the set
 * is queried (errors and stats are reported through the engine). */
void cf_tables_free(cf_tables* t);
/* Device bytes of the resident store. */
size_t cf_tables_bytes(const cf_tables* t);
/* The reference's need(n, M) for the set's d and model (fast/cuckoo_fast.cu:452-459). */
long cf_tables_need(const cf_tables* t, int n, int M);

/* Batched queries on a resident set: OneTableSuffix::nbad (single_table.hpp:178-229) /
 * QueryAll with the host Horner finish (fast/cuckoo_fast.cu:306-323, 877-890) for J queries at
 * once, each table row read once for the batch. res[b J + j] = N_bad(n_j, M_j) mod p_b for
 * b < need[j] (need NULL: every prime), else 0. n_j <= nmax, n_j < M_j <= Mmax. */
int cf_query(cf_tables* t, const int32_t* n, const int32_t* M, size_t J, const int64_t* need, uint64_t* res);

/* cf_query without any device allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, checks and results as cf_query; the query descriptors, row
 * ranges, partial sums and results of a call live in one fixed 32 MiB array per device, compiled
 * into the library (held until the process ends), instead of the set's grown buffer. Its device
 * address is looked up when an engine is opened (devices 0..63). CF_E_ARGS when a call would need
 * more than the 32 MiB: about 110 KB per query at 1536 primes, so up to ~290 queries per call
 * there. The array keeps the last call's contents, and every part a call reads it writes first.
 * Calls on one device share the array: this library runs them one after the other (a per-device
 * lock held until the call's results are back), so only another process could race them. */
int cf_query_noalloc(cf_tables* t, const int32_t* n, const int32_t* M, size_t J, const int64_t* need, uint64_t* res);

/* log2 of the x in [0, prod p) with x = res[i] mod primes[i] (host CRT by Garner's mixed radix,
 * as fast/post_gmp.py's CRT merge, in long double from the leading digits); -inf for x = 0. */
double cf_crt_log2(const uint64_t* primes, const uint64_t* res, size_t k);

/* ---- the exact combiner: fast/post_gmp.py's columns without GMP (host only) ---- */
typedef struct {
    char log2_pfail[48];  /* log2(N_bad / Q^n) correctly rounded to 15 decimals; "-inf" for N_bad = 0 */
    int32_t le[3];        /* 1 iff N_bad * 2^lambda <= Q^n, lambda = 40, 64, 128 (exact) */
    char p_fail[64];      /* N_bad / Q^n truncated to 40 significant digits ("d.ddd...e-X"), or "0" */
    int64_t nbad_bits;    /* bit lengths of N_bad and of Q^n = C(M, d)^n */
    int64_t qn_bits;
    char error[128];      /* why a call failed */
} cf_exact;
/* One row of post_gmp.py (fast/post_gmp.py:51-60 crt, :63-76 sci, :79-91 the row): N_bad is the
 * x in [0, prod primes) with x = res[i] mod primes[i] (odd, pairwise coprime moduli < 2^63), and
 * Q = C(M, d). CF_E_PRIMES when prod primes <= 4 Q^n, CF_E_ARGS when x > Q^n or the moduli are
 * unusable (post_gmp.py:84's assertion and gmpy2.invert's error). With nbad_dec, also writes
 * N_bad in decimal (dec_len >= 20 k + 2 always suffices). */
int cf_exact_combine(int d, int n, int M, const uint64_t* primes, const uint64_t* res, size_t k, cf_exact* out,
                     char* nbad_dec, size_t dec_len);
/* post_gmp.py's need(n, M) in double precision (fast/post_gmp.py:46-48): how many primes above
 * 2^pbits its CRT uses. It can differ from cf_need(), the engine's long-double need. -1 for M < d. */
long cf_post_need(int n, int M, int d, int pbits);

/* The serving loop over one resident set. Requests are admitted into up to `slots` batches in
 * flight at once (each on its own stream; fewer if device memory runs short), a free slot taking
 * every request waiting at that moment (at most max_batch); finished requests are combined by CRT
 * on crt_threads host threads. No counterpart in the reference, which has no serving mode. */
typedef struct cf_server cf_server;
typedef struct {
    double wall_s;         /* first arrival to last completion */
    double queries_per_s;
    double p50_ms, p99_ms; /* arrival to completion (CRT included) */
    double mean_batch;
    long batches;
} cf_serve_stats;
int cf_server_open(cf_tables* t, int slots, int max_batch, int crt_threads, cf_server** out);
void cf_server_close(cf_server* s);
int cf_server_slots(const cf_server* s);
/* Replays `count` requests (n[i], M[i]) arriving at arrival_s[i] seconds (non-decreasing) after
 * the call starts, in real time, and returns when all are answered. Per request (any output may
 * be NULL): latency_ms, log2_pfail = log2(N_bad / Q^n) over its need() primes, and res[i np + b]
 * (N_bad mod p_b for b < need, else 0). CF_E_PRIMES if a request needs more primes than the set. */
int cf_server_replay(cf_server* s, size_t count, const int32_t* n, const int32_t* M, const double* arrival_s,
                     double* latency_ms, double* log2_pfail, uint64_t* res, cf_serve_stats* stats);

/* How many primes of a table with nmax rows and factorials up to Mmax, answering J queries, fit
 * in the device memory now free (the reference's batch sizing, fast/cuckoo_fast.cu:742-750). */
long cf_max_batch(cf_engine* e, int nmax, int Mmax, size_t J);
/* Device bytes one such prime takes. */
size_t cf_mem_per_prime(int nmax, int Mmax, size_t J);
/* The same for primes above 2^50 (the reference's -DFP_RB=64 build with 62-bit primes,
 * fast/cuckoo_fast.cu:48-60, batch sizing :742-750): cf_run_fused runs such primes (odd, up to
 * 2^63, tables of up to 131072 rows) on a 64-bit Montgomery kernel, in tile pools of about
 * 4 n^2 bytes per prime. */
long cf_max_batch62(cf_engine* e, int nmax, int Mmax, size_t J);
size_t cf_mem_per_prime62(int nmax, int Mmax, size_t J);

/* ---- host logic of the reference, exported for drivers and tests (no device needed) ---- */
/* Primes number first .. first+count-1 of the fixed list: the i-th largest p = c 2^20 + 1 below
 * 2^prime_bits (fast/cuckoo_fast.cu:423-432). Returns how many were written to out. */
long cf_prime_range(int prime_bits, long first, long count, uint64_t* out);
/* Number of primes whose product exceeds Q^n, Q = C(M, d) (fast/cuckoo_fast.cu:452-459). */
long cf_need(int n, int M, int d, int prime_bits);
/* max(16, next_pow2(nmax)) (fast/cuckoo_fast.cu:465-472). */
size_t cf_transform_length(size_t nmax);
/* count 32-bit primes with 2-adic valuation k = 31 .. min_pow2, largest first
 * (include/cuckoo/modint.hpp:183-196). Returns count, or -1 if the list runs out. */
int cf_ntt_primes32(int count, int min_pow2, uint32_t* out);
/* log2 of the engine's transform length for nmax rows modulo p, or -1 if this build cannot run
 * that shape. */
int cf_engine_log2_length(int nmax, uint64_t p);

#ifdef __cplusplus
}
#endif
#endif
