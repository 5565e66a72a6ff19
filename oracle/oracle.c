/*
 * oracle.c -- CPU restatement of the reference's one-table failure-count algorithm.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine, never part of the product. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Parity pins (see tests/test_oracle.py):
 *   - or_prime_range / or_need / or_nbad*: residue fixtures written by the reference's own host
 *     build (tests/golden/fast_*.txt), and N_bad from the reference's 32-bit CLI reduced mod the
 *     50-bit primes (cross-precision identity);
 *   - or_suffix_table_* / or_nbad: the reference test program's brute-force counts
 *     (tests/golden/check_ref.txt), the d = 1 closed form M^n - M!/(M-n)!, ntt = cubic, and the
 *     reference's 32-bit CLI output (tests/golden/fail_*.txt);
 *   - or_ntt_primes32: the prime lists implied by the reference CLI's product/primes lines;
 *   - or_nbad_forward: the reference CLI's --algo forward outputs (tests/golden/fail_algo_cases.json),
 *     the brute-force counts, and forward = cubic = ntt;
 *   - or_brute_force: the reference test program's brute-force counts;
 *   - or_cyclic_conv_*: each other (schoolbook vs NTT) on random inputs.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef unsigned __int128 u128;

static u64 mulm(u64 a, u64 b, u64 p) { return (u64)((u128)a * b % p); }
static u64 addm(u64 a, u64 b, u64 p) { u64 s = a + b; return (s >= p || s < a) ? s - p : s; }
static u64 subm(u64 a, u64 b, u64 p) { return a >= b ? a - b : a + (p - b); }
static u64 powm(u64 a, u64 e, u64 p) {
    u64 r = 1 % p;
    a %= p;
    while (e) {
        if (e & 1) r = mulm(r, a, p);
        a = mulm(a, a, p);
        e >>= 1;
    }
    return r;
}
static u64 invm(u64 a, u64 p) { return powm(a, p - 2, p); }

/* ---- primes, need, transform length (fast/cuckoo_fast.cu:403-472) ---- */

int or_is_prime_u64(u64 n) {
    static const u64 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; ++i)
        if (n % small[i] == 0) return n == small[i];
    u64 d = n - 1;
    int s = 0;
    while (!(d & 1)) d >>= 1, ++s;
    for (int i = 0; i < 12; ++i) {
        u64 x = powm(small[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int witness = 1;
        for (int r = 1; r < s; ++r) {
            x = mulm(x, x, n);
            if (x == n - 1) { witness = 0; break; }
        }
        if (witness) return 0;
    }
    return 1;
}

long or_prime_range(int prime_bits, long first, long count, u64* out) {
    long found = 0, idx = 0;
    for (u64 c = (((u64)1 << prime_bits) - 2) >> 20; c >= 1 && found < count; --c) {
        const u64 p = (c << 20) + 1;
        if (!or_is_prime_u64(p)) continue;
        if (idx++ >= first) out[found++] = p;
    }
    return found;
}

double or_log2_q(int M, int d) {
    long double b = 0;
    for (int i = 0; i < d; ++i) b += log2l((long double)(M - i)) - log2l((long double)(i + 1));
    return (double)b;
}

long or_need(int n, int M, int d, int prime_bits) {
    return (long)ceill(((long double)n * or_log2_q(M, d) + 3) / (prime_bits - 1)) + 1;
}

size_t or_transform_length(size_t nmax) {
    size_t L = 1;
    while (L < nmax) L <<= 1;
    return L < 16 ? 16 : L;
}

/* ---- 32-bit prime list (include/cuckoo/modint.hpp:137-196) ---- */

static int cmp_desc_u32(const void* a, const void* b) {
    const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? 1 : x > y ? -1 : 0;
}

int or_ntt_primes32(int count, int min_pow2, uint32_t* out) {
    int got = 0;
    for (int k = 31; k >= min_pow2 && got < count; --k) {
        long long c = (long long)((((u64)1 << 32) - 2) >> k);
        if (c % 2 == 0) --c;
        for (; c >= 1 && got < count; c -= 2) {
            const u64 p = ((u64)c << k) + 1;
            if (or_is_prime_u64(p)) out[got++] = (uint32_t)p;
        }
    }
    if (got < count) return -1;
    qsort(out, (size_t)got, sizeof(uint32_t), cmp_desc_u32);
    return got;
}

/* ---- factorials and the candidate-set model (single_table.hpp:16-23, modint.hpp:107-124) ---- */

typedef struct {
    int n;
    u64 *f, *fi;
} Fact;

static int fact_init(Fact* F, int n, u64 p) {
    F->n = n;
    F->f = (u64*)malloc(sizeof(u64) * (size_t)(n + 1));
    F->fi = (u64*)malloc(sizeof(u64) * (size_t)(n + 1));
    if (!F->f || !F->fi) return -1;
    F->f[0] = 1 % p;
    for (int i = 1; i <= n; ++i) F->f[i] = mulm(F->f[i - 1], (u64)i % p, p);
    F->fi[n] = invm(F->f[n], p);
    for (int i = n; i > 0; --i) F->fi[i - 1] = mulm(F->fi[i], (u64)i % p, p);
    return 0;
}
static void fact_free(Fact* F) { free(F->f), free(F->fi); }
static u64 binom(const Fact* F, long long a, long long b, u64 p) {
    if (a < 0 || b < 0 || b > a) return 0;
    return mulm(mulm(F->f[a], F->fi[b], p), F->fi[a - b], p);
}
static u64 model_g(const Fact* F, int v, int d, int rep, u64 p) {
    if (v < 0) return 0;
    return rep ? powm((u64)v % p, (u64)d, p) : binom(F, v, d, p);
}

/* ---- textbook NTT (bit reversal + iterative Cooley-Tukey) ---- */

static u64 root_of_order(u64 p, size_t L) {
    for (u64 x = 2;; ++x) {
        const u64 w = powm(x, (p - 1) / L, p);
        if (L == 1 || powm(w, L / 2, p) != 1) return w;
    }
}

static void ntt_inplace(u64* a, size_t L, u64 w, u64 p) {
    for (size_t i = 1, j = 0; i < L; ++i) {
        size_t bit = L >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { u64 t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (size_t len = 2; len <= L; len <<= 1) {
        const u64 wl = powm(w, L / len, p);
        for (size_t i = 0; i < L; i += len) {
            u64 x = 1;
            for (size_t k = 0; k < len / 2; ++k) {
                const u64 u = a[i + k], t = mulm(a[i + k + len / 2], x, p);
                a[i + k] = addm(u, t, p);
                a[i + k + len / 2] = subm(u, t, p);
                x = mulm(x, wl, p);
            }
        }
    }
}

void or_cyclic_conv_schoolbook(u64 p, size_t L, const u64* a, const u64* k, u64* out) {
    for (size_t i = 0; i < L; ++i) {
        u64 s = 0;
        for (size_t j = 0; j < L; ++j) s = addm(s, mulm(a[j] % p, k[(i + L - j) % L] % p, p), p);
        out[i] = s;
    }
}

int or_cyclic_conv_ntt(u64 p, size_t L, const u64* a, const u64* k, u64* out) {
    if (L == 0 || (L & (L - 1)) || (p - 1) % L) return -1;
    u64* A = (u64*)malloc(sizeof(u64) * L);
    u64* K = (u64*)malloc(sizeof(u64) * L);
    if (!A || !K) { free(A); free(K); return -1; }
    for (size_t i = 0; i < L; ++i) A[i] = a[i] % p, K[i] = k[i] % p;
    const u64 w = root_of_order(p, L), wi = invm(w, p), li = invm((u64)L % p, p);
    ntt_inplace(A, L, w, p);
    ntt_inplace(K, L, w, p);
    for (size_t i = 0; i < L; ++i) A[i] = mulm(A[i], K[i], p);
    ntt_inplace(A, L, wi, p);
    for (size_t i = 0; i < L; ++i) out[i] = mulm(A[i], li, p);
    free(A), free(K);
    return 0;
}

/* Convolution helper with a kernel spectrum prepared once. */
typedef struct {
    size_t L;
    u64 p, w, wi, li;
    u64* khat;
    u64* buf;
} Conv;

static int conv_init(Conv* C, u64 p, int nm, const Fact* F) {
    /* No wrap at all when L >= 2 nm (entries s + t <= 2 nm - 2); otherwise the shortest length
     * the reference admits (L >= nm, reading c[nm] at index nm mod L, single_table.hpp:168-175). */
    size_t L = 1;
    while (L < (size_t)(2 * nm)) L <<= 1;
    if ((p - 1) % L) {
        L = 1;
        while (L < (size_t)nm) L <<= 1;
        if (L < 2) L = 2;
        if ((p - 1) % L) return -1;
    }
    C->L = L, C->p = p;
    C->w = root_of_order(p, L), C->wi = invm(C->w, p), C->li = invm((u64)L % p, p);
    C->khat = (u64*)calloc(L, sizeof(u64));
    C->buf = (u64*)calloc(L, sizeof(u64));
    if (!C->khat || !C->buf) return -1;
    for (int t = 1; t < nm; ++t) C->khat[t] = mulm(F->fi[t], F->f[t - 1], p); /* 1/t */
    ntt_inplace(C->khat, L, C->w, p);
    return 0;
}
static void conv_free(Conv* C) { free(C->khat), free(C->buf); }
/* buf (length L, zero-padded by the caller) <- buf (*) (1/t kernel) */
static void conv_run(Conv* C) {
    ntt_inplace(C->buf, C->L, C->w, C->p);
    for (size_t i = 0; i < C->L; ++i) C->buf[i] = mulm(C->buf[i], C->khat[i], C->p);
    ntt_inplace(C->buf, C->L, C->wi, C->p);
    for (size_t i = 0; i < C->L; ++i) C->buf[i] = mulm(C->buf[i], C->li, C->p);
}

static size_t tri(size_t r) { return r * (r - 1) / 2; } /* row r >= 1 of a packed triangle */

/* ---- suffix table, O(n^2 log n) (single_table.hpp:80-96, 178-229) ---- */

int or_suffix_table_ntt(int nm, int d, int rep, u64 p, u64* V) {
    if (nm < 1 || d < 1) return -1;
    Fact F;
    Conv C;
    if (fact_init(&F, nm, p) || conv_init(&C, p, nm, &F)) return -1;
    memset(V, 0, sizeof(u64) * tri((size_t)nm + 1));
    /* S[v][q], 1 <= q <= v < nm, packed with row v at tri(v) + q - 1 */
    u64* S = (u64*)calloc(tri((size_t)nm) + 1, sizeof(u64));
    if (!S) return -1;
    for (int v = 0; v < nm; ++v) {
        if (v >= 1) { /* stage 1: column v of S from row v of V */
            int any = 0;
            memset(C.buf, 0, sizeof(u64) * C.L);
            for (int w = 0; w < v; ++w) {
                C.buf[w] = mulm(mulm(V[tri(v) + w], F.fi[w], p), F.fi[v - 1 - w], p);
                any |= C.buf[w] != 0;
            }
            if (any) {
                conv_run(&C);
                for (int vp = v; vp < nm; ++vp)
                    S[tri(vp) + v - 1] = mulm(mulm(F.f[vp], F.fi[vp - v], p), C.buf[(size_t)vp & (C.L - 1)], p);
            }
        }
        const u64 g = model_g(&F, v, d, rep, p);
        if (g == 0) continue; /* stage 2: column v of V from row v of S */
        const u64 gi = invm(g, p);
        int any = 0;
        memset(C.buf, 0, sizeof(u64) * C.L);
        u64 gq = 1;
        for (int q = 1; q <= v; ++q) {
            gq = mulm(gq, gi, p);
            C.buf[q] = mulm(mulm(mulm(S[tri(v) + q - 1], gq, p), F.fi[q], p), F.fi[v - q], p);
            any |= C.buf[q] != 0;
        }
        if (any) conv_run(&C);
        u64 gp = powm(g, (u64)v + 1, p);
        for (int pp = v + 1; pp <= nm; ++pp) {
            u64 val = binom(&F, pp - 1, v, p);
            if (any) val = subm(val, mulm(mulm(F.f[pp], F.fi[pp - v - 1], p), C.buf[(size_t)pp & (C.L - 1)], p), p);
            V[tri(pp) + v] = mulm(val, gp, p);
            gp = mulm(gp, g, p);
        }
    }
    free(S);
    conv_free(&C);
    fact_free(&F);
    return 0;
}

/* ---- suffix table, O(n^3) (single_table.hpp:135-166) ---- */

int or_suffix_table_cubic(int nm, int d, int rep, u64 p, u64* V) {
    if (nm < 1 || d < 1) return -1;
    Fact F;
    if (fact_init(&F, nm, p)) return -1;
    memset(V, 0, sizeof(u64) * tri((size_t)nm + 1));
    u64* S = (u64*)calloc(tri((size_t)nm) + 1, sizeof(u64));
    u64* gp = (u64*)malloc(sizeof(u64) * (size_t)(nm + 1)); /* g(v)^e for the current v */
    if (!S || !gp) return -1;
    for (int v = 0; v < nm; ++v) {
        if (v >= 1)
            for (int vp = v; vp < nm; ++vp) {
                u64 s = 0;
                for (int w = 0; w < v; ++w)
                    s = addm(s, mulm(mulm(binom(&F, vp - w - 1, vp - v, p), binom(&F, vp, w, p), p), V[tri(v) + w], p), p);
                S[tri(vp) + v - 1] = s;
            }
        const u64 g = model_g(&F, v, d, rep, p);
        gp[0] = 1 % p;
        for (int e = 1; e <= nm; ++e) gp[e] = mulm(gp[e - 1], g, p);
        for (int pp = v + 1; pp <= nm; ++pp) {
            u64 val = mulm(binom(&F, pp - 1, v, p), gp[pp], p);
            for (int q = 1; q <= v; ++q) {
                const u64 t = mulm(mulm(binom(&F, pp - q - 1, pp - v - 1, p), binom(&F, pp, q, p), p), gp[pp - q], p);
                val = subm(val, mulm(t, S[tri(v) + q - 1], p), p);
            }
            V[tri(pp) + v] = val;
        }
    }
    free(S), free(gp);
    fact_free(&F);
    return 0;
}

/* ---- Algorithm 1, the O(n^4) forward recurrence over block boundaries
 *      (single_table.hpp:25-31 phi, :44-78 one_table_forward), with the full M x (M+1) table ---- */

static u64 phi_ab(const Fact* F, long long a, long long b, long long r, u64 p) {
    if (b <= 0) return binom(F, a - 1, r, p);
    const u64 t = mulm(binom(F, a - b - 1, r - b, p), binom(F, r - 1, b - 1, p), p);
    return t ? p - t : 0;
}

int or_nbad_forward(int d, int rep, u64 p, int n, int M, u64* out) {
    if (n < 1 || M <= n || d < 1) return -1;
    Fact F;
    if (fact_init(&F, M, p)) return -1;
    const int z = M - n;
    const size_t HB = (size_t)(n + 1), FB = (size_t)(M + 1);
    u64* hpow = (u64*)malloc(sizeof(u64) * FB * HB); /* hpow[b][u] = g(M - b)^u */
    u64* Ft = (u64*)calloc((size_t)M * FB, sizeof(u64)); /* F[l][b] */
    if (!hpow || !Ft) return -1;
    for (int b = 0; b <= M; ++b) {
        const u64 h = model_g(&F, M - b, d, rep, p);
        hpow[(size_t)b * HB] = 1 % p;
        for (int u = 1; u <= n; ++u) hpow[(size_t)b * HB + u] = mulm(hpow[(size_t)b * HB + u - 1], h, p);
    }
#define NU(l) ((l) > z ? (l) - z : 0)
#define TAU(l, b, r, c)                                                                                     \
    mulm(mulm(mulm(phi_ab(&F, (c) - (l), (b) - (l), (r) - (l), p), binom(&F, M - (b), (c) - (b), p), p),    \
              binom(&F, n - NU(l), (r) - ((l) > z ? (l) : z), p), p),                                        \
         hpow[(size_t)(b) * HB + ((r) - ((l) > z ? (l) : z))], p)
    Ft[0] = 1 % p;
    for (int r = z; r < M; ++r)
        for (int c = r + 1; c <= M; ++c) {
            u64 s = mulm(Ft[0], TAU(0, 0, r, c), p);
            for (int l = z; l < r; ++l)
                for (int b = l + 1; b < c; ++b)
                    if (Ft[(size_t)l * FB + b]) s = addm(s, mulm(Ft[(size_t)l * FB + b], TAU(l, b, r, c), p), p);
            Ft[(size_t)r * FB + c] = s;
        }
    u64 total = 0;
    for (int l = 0; l < M; ++l)
        for (int b = 0; b <= M; ++b)
            if (Ft[(size_t)l * FB + b])
                total = addm(total, mulm(mulm(Ft[(size_t)l * FB + b], binom(&F, M - l - 1, M - b, p), p),
                                         hpow[(size_t)b * HB + n - NU(l)], p), p);
#undef TAU
#undef NU
    *out = total;
    free(hpow), free(Ft);
    fact_free(&F);
    return 0;
}

/* ---- separate tables (separate_tables.hpp): d tables of lengths m_1..m_d, g(v) = prod v_t ---- */

/* All v >= 1 (coordinatewise) with |v| <= smax, by increasing |v|; coordinates in *coord (d per
 * state), sums in *sum; idx (size (smax+2)^d) maps a vector to its index or -1. */
typedef struct {
    int d, smax, N;
    int *coord, *sum, *idx;
} States;

static long st_key(const States* S, const int* v) {
    long k = 0;
    for (int t = 0; t < S->d; ++t) k = k * (S->smax + 2) + v[t];
    return k;
}

static int states_init(States* S, int d, int smax) {
    S->d = d, S->smax = smax < 0 ? 0 : smax, S->N = 0;
    long cells = 1;
    for (int t = 0; t < d; ++t) cells *= S->smax + 2;
    S->idx = (int*)malloc(sizeof(int) * (size_t)cells);
    int cap = 1024;
    S->coord = (int*)malloc(sizeof(int) * (size_t)cap * d);
    S->sum = (int*)malloc(sizeof(int) * (size_t)cap);
    if (!S->idx || !S->coord || !S->sum) return -1;
    for (long c = 0; c < cells; ++c) S->idx[c] = -1;
    int v[8];
    for (int s = d; s <= S->smax; ++s) {
        /* every composition of s into d positive parts, odometer order */
        for (int t = 0; t < d; ++t) v[t] = 1;
        v[d - 1] = s - (d - 1);
        for (;;) {
            if (S->N == cap) {
                cap *= 2;
                S->coord = (int*)realloc(S->coord, sizeof(int) * (size_t)cap * d);
                S->sum = (int*)realloc(S->sum, sizeof(int) * (size_t)cap);
                if (!S->coord || !S->sum) return -1;
            }
            memcpy(S->coord + (size_t)S->N * d, v, sizeof(int) * d);
            S->sum[S->N] = s;
            S->idx[st_key(S, v)] = S->N;
            S->N++;
            /* the next composition in lexicographic order: the rightmost part t < d-1 whose tail
             * (parts t+1 .. d-1) can give up a unit grows by one, the tail restarts at 1, .., 1, rest */
            int rest = v[d - 1];
            int t = d - 2;
            while (t >= 0 && rest <= d - 1 - t) rest += v[t--];
            if (t < 0) break;
            v[t]++;
            rest--;
            for (int u = t + 1; u < d - 1; ++u) v[u] = 1, rest--;
            v[d - 1] = rest;
        }
    }
    return 0;
}
static void states_free(States* S) { free(S->coord), free(S->sum), free(S->idx); }

/* The suffix table by its defining sums (Algorithm 5's recurrence, separate_tables.hpp:175-186,
 * without the generating-function evaluation), then N_bad (:190-213). */
int or_separate_nbad(int n, const int* m, int d, u64 p, u64* out) {
    if (n < 1 || d < 1 || d > 8) return -1;
    int M = 0;
    for (int t = 0; t < d; ++t) {
        if (m[t] < 1) return -1;
        M += m[t];
    }
    if (M <= n) return -1;
    Fact F;
    if (fact_init(&F, M > n ? M : n, p)) return -1;
    States X;
    if (states_init(&X, d, n - 1)) return -1;
    const int N = X.N, P = n + 1;
    u64* g = (u64*)malloc(sizeof(u64) * (size_t)(N ? N : 1));
    u64* V = (u64*)calloc((size_t)P * (N ? N : 1), sizeof(u64));  /* V[p][i] */
    u64* S = (u64*)calloc((size_t)(N ? N : 1) * P, sizeof(u64));  /* S[i][q] */
    if (!g || !V || !S) return -1;
    for (int i = 0; i < N; ++i) {
        g[i] = 1 % p;
        for (int t = 0; t < d; ++t) g[i] = mulm(g[i], (u64)X.coord[(size_t)i * d + t] % p, p);
    }
    for (int pp = 1; pp <= n; ++pp) {
        const int q0 = pp - 1; /* row q0 of V is final: S[.][q0] from it */
        if (q0 >= 1)
            for (int i = 0; i < N; ++i) {
                if (X.sum[i] < q0) continue;
                const int* v = X.coord + (size_t)i * d;
                u64 acc = 0;
                for (int j = 0; j < N && X.sum[j] < q0; ++j) {
                    const int* w = X.coord + (size_t)j * d;
                    int le = 1;
                    for (int t = 0; t < d && le; ++t) le = w[t] <= v[t];
                    if (!le) continue;
                    u64 c = binom(&F, X.sum[i] - X.sum[j] - 1, X.sum[i] - q0, p);
                    for (int t = 0; t < d && c; ++t) c = mulm(c, binom(&F, v[t], w[t], p), p);
                    acc = addm(acc, mulm(c, V[(size_t)q0 * N + j], p), p);
                }
                S[(size_t)i * P + q0] = acc;
            }
        for (int i = 0; i < N && X.sum[i] < pp; ++i) {
            const int sv = X.sum[i];
            u64 val = mulm(binom(&F, pp - 1, sv, p), powm(g[i], (u64)pp, p), p);
            for (int q = 1; q <= sv; ++q) {
                const u64 c = mulm(mulm(binom(&F, pp - q - 1, pp - sv - 1, p), binom(&F, pp, q, p), p),
                                   powm(g[i], (u64)(pp - q), p), p);
                val = subm(val, mulm(c, S[(size_t)i * P + q], p), p);
            }
            V[(size_t)pp * N + i] = val;
        }
    }
    u64 Q = 1 % p, total = 0;
    for (int t = 0; t < d; ++t) Q = mulm(Q, (u64)m[t] % p, p);
    for (int q = 1; q <= n; ++q) {
        u64 inner = 0;
        for (int j = 0; j < N && X.sum[j] < q; ++j) {
            const int* w = X.coord + (size_t)j * d;
            u64 c = binom(&F, M - X.sum[j] - 1, M - q, p);
            for (int t = 0; t < d && c; ++t) c = mulm(c, binom(&F, m[t], w[t], p), p);
            inner = addm(inner, mulm(c, V[(size_t)q * N + j], p), p);
        }
        total = addm(total, mulm(mulm(binom(&F, n, q, p), powm(Q, (u64)(n - q), p), p), inner, p), p);
    }
    *out = total;
    free(g), free(V), free(S);
    states_free(&X);
    fact_free(&F);
    return 0;
}

/* Algorithm 4, the forward recurrence over (l, b) (separate_tables.hpp:27-108). */
int or_separate_forward(int n, const int* m, int d, u64 p, u64* out) {
    if (n < 1 || d < 1 || d > 8) return -1;
    int M = 0;
    for (int t = 0; t < d; ++t) {
        if (m[t] < 1) return -1;
        M += m[t];
    }
    if (M <= n) return -1;
    const int z = M - n;
    Fact F;
    if (fact_init(&F, M, p)) return -1;
    States X; /* v = m - b >= 1 with |v| <= M - 1 - l; enumerate all v >= 1 with |v| <= M - 1 once */
    if (states_init(&X, d, M - 1)) return -1;
    /* the boundary states in order of l: (0, 0), then for l = z .. M-1 every v <= m with |v| < M - l */
    int cap = 1024, K = 0;
    int* sl = (int*)malloc(sizeof(int) * (size_t)cap);
    int* sv = (int*)malloc(sizeof(int) * (size_t)cap); /* index into X, -1 for (0, 0) */
    if (!sl || !sv) return -1;
    sl[K] = 0, sv[K] = -1, K++;
    for (int l = z; l < M; ++l)
        for (int i = 0; i < X.N && X.sum[i] < M - l; ++i) {
            const int* v = X.coord + (size_t)i * d;
            int ok = 1;
            for (int t = 0; t < d && ok; ++t) ok = v[t] <= m[t];
            if (!ok) continue;
            if (K == cap) {
                cap *= 2;
                sl = (int*)realloc(sl, sizeof(int) * (size_t)cap);
                sv = (int*)realloc(sv, sizeof(int) * (size_t)cap);
                if (!sl || !sv) return -1;
            }
            sl[K] = l, sv[K] = i, K++;
        }
    int* b = (int*)malloc(sizeof(int) * (size_t)K * d);
    int* bs = (int*)malloc(sizeof(int) * (size_t)K);
    u64* h = (u64*)malloc(sizeof(u64) * (size_t)K);
    u64* Fw = (u64*)calloc((size_t)K, sizeof(u64));
    if (!b || !bs || !h || !Fw) return -1;
    for (int k = 0; k < K; ++k) {
        bs[k] = 0, h[k] = 1 % p;
        for (int t = 0; t < d; ++t) {
            const int vt = sv[k] < 0 ? m[t] : X.coord[(size_t)sv[k] * d + t];
            b[(size_t)k * d + t] = m[t] - vt;
            bs[k] += m[t] - vt;
            h[k] = mulm(h[k], (u64)vt % p, p);
        }
    }
#define NU(l) ((l) > z ? (l) - z : 0)
    Fw[0] = 1 % p;
    for (int i = 1; i < K; ++i) {
        u64 s = 0;
        for (int j = 0; j < i && sl[j] < sl[i]; ++j) {
            if (!Fw[j] || bs[j] >= bs[i]) continue;
            int le = 1;
            for (int t = 0; t < d && le; ++t) le = b[(size_t)j * d + t] <= b[(size_t)i * d + t];
            if (!le) continue;
            u64 tau = phi_ab(&F, bs[i] - sl[j], bs[j] - sl[j], sl[i] - sl[j], p);
            for (int t = 0; t < d && tau; ++t)
                tau = mulm(tau, binom(&F, m[t] - b[(size_t)j * d + t], b[(size_t)i * d + t] - b[(size_t)j * d + t], p), p);
            const int u = sl[i] - (sl[j] > z ? sl[j] : z);
            tau = mulm(tau, mulm(binom(&F, n - NU(sl[j]), u, p), powm(h[j], (u64)u, p), p), p);
            s = addm(s, mulm(Fw[j], tau, p), p);
        }
        Fw[i] = s;
    }
    u64 total = 0;
    for (int i = 0; i < K; ++i)
        if (Fw[i])
            total = addm(total, mulm(mulm(Fw[i], binom(&F, M - sl[i] - 1, M - bs[i], p), p),
                                     powm(h[i], (u64)(n - NU(sl[i])), p), p), p);
#undef NU
    *out = total;
    free(sl), free(sv), free(b), free(bs), free(h), free(Fw);
    states_free(&X);
    fact_free(&F);
    return 0;
}

/* ---- queries ---- */

static int max_n(int nq, const int* n) {
    int m = 0;
    for (int j = 0; j < nq; ++j) m = n[j] > m ? n[j] : m;
    return m;
}

/* The paper's double sum (single_table.hpp:103-119). */
int or_nbad(int algo, int d, int rep, u64 p, int nq, const int* n, const int* M, u64* out) {
    const int nm = max_n(nq, n);
    if (nm < 1) return -1;
    u64* V = (u64*)malloc(sizeof(u64) * tri((size_t)nm + 1));
    if (!V) return -1;
    const int rc = algo == 0 ? or_suffix_table_ntt(nm, d, rep, p, V) : or_suffix_table_cubic(nm, d, rep, p, V);
    if (rc) { free(V); return rc; }
    for (int j = 0; j < nq; ++j) {
        if (n[j] < 1 || M[j] <= n[j]) { free(V); return -1; }
        Fact F;
        if (fact_init(&F, M[j], p)) { free(V); return -1; }
        const u64 Q = model_g(&F, M[j], d, rep, p);
        u64 total = 0, Qp = 1 % p;
        /* Horner-free form: sum_q C(n,q) Q^(n-q) inner(q), accumulated from q = n down */
        for (int q = n[j]; q >= 1; --q) {
            u64 inner = 0;
            for (int w = 0; w < q; ++w) {
                const u64 x = V[tri(q) + w];
                if (!x) continue;
                inner = addm(inner, mulm(mulm(binom(&F, M[j] - w - 1, M[j] - q, p), binom(&F, M[j], w, p), p), x, p), p);
            }
            total = addm(total, mulm(mulm(binom(&F, n[j], q, p), Qp, p), inner, p), p);
            Qp = mulm(Qp, Q, p);
        }
        out[j] = total;
        fact_free(&F);
    }
    free(V);
    return 0;
}

/* The fast engine's query form (cuckoo_fast.cu:306-323) and Horner finish (877-890). */
int or_nbad_horner(int d, int rep, u64 p, int nq, const int* n, const int* M, u64* out) {
    const int nm = max_n(nq, n);
    if (nm < 1) return -1;
    u64* V = (u64*)malloc(sizeof(u64) * tri((size_t)nm + 1));
    if (!V || or_suffix_table_ntt(nm, d, rep, p, V)) { free(V); return -1; }
    for (int j = 0; j < nq; ++j) {
        if (n[j] < 1 || M[j] <= n[j]) { free(V); return -1; }
        Fact F;
        if (fact_init(&F, M[j], p)) { free(V); return -1; }
        const u64 Q = model_g(&F, M[j], d, rep, p);
        u64 total = 0;
        for (int k = 1; k <= n[j]; ++k) {
            u64 o = 0; /* out[k] = sum_w U[k][w] / (M - w) */
            for (int w = 0; w < k; ++w) {
                const u64 U = mulm(mulm(V[tri(k) + w], F.fi[w], p), F.fi[k - 1 - w], p);
                o = addm(o, mulm(U, mulm(F.fi[M[j] - w], F.f[M[j] - w - 1], p), p), p);
            }
            const u64 term = mulm(mulm(binom(&F, n[j], k, p), F.fi[M[j] - k], p), o, p);
            total = addm(mulm(total, Q, p), term, p);
        }
        out[j] = mulm(total, F.f[M[j]], p);
        fact_free(&F);
    }
    free(V);
    return 0;
}

/* ---- brute force (brute_force.hpp:22-146) ---- */

typedef struct {
    int n, M, nc;
    const u64* mask;
    const unsigned long long* wt;
    int* idx;
    u64* items;
    int owner[64];
    char seen[64];
    unsigned long long fact[64];
    unsigned long long nbad;
} Brute;

static int kuhn_augment(Brute* B, int i) {
    u64 m = B->items[i];
    while (m) {
        const int s = __builtin_ctzll(m);
        m &= m - 1;
        if (B->seen[s]) continue;
        B->seen[s] = 1;
        if (B->owner[s] < 0 || kuhn_augment(B, B->owner[s])) {
            B->owner[s] = i;
            return 1;
        }
    }
    return 0;
}

static int has_full_matching(Brute* B) {
    for (int s = 0; s < B->M; ++s) B->owner[s] = -1;
    for (int i = 0; i < B->n; ++i) {
        memset(B->seen, 0, sizeof B->seen);
        if (!kuhn_augment(B, i)) return 0;
    }
    return 1;
}

static void brute_rec(Brute* B, int pos, int start) {
    if (pos == B->n) {
        if (has_full_matching(B)) return;
        /* multinomial count of orderings of this multiset times the choice weights */
        unsigned long long w = B->fact[B->n];
        for (int i = 0; i < B->n;) {
            int j = i;
            while (j < B->n && B->idx[j] == B->idx[i]) ++j;
            w /= B->fact[j - i];
            for (int t = i; t < j; ++t) w *= B->wt[B->idx[i]];
            i = j;
        }
        B->nbad += w;
        return;
    }
    for (int c = start; c < B->nc; ++c) {
        B->idx[pos] = c;
        B->items[pos] = B->mask[c];
        brute_rec(B, pos + 1, c);
    }
}

int or_brute_force(int n, int M, int d, int rep, unsigned long long* nbad, unsigned long long* total) {
    if (M < 1 || M > 20 || n < 1 || n > 20 || d < 1) return -1;
    size_t cap = (size_t)1 << M, nc = 0;
    u64* mask = (u64*)malloc(sizeof(u64) * cap);
    unsigned long long* wt = (unsigned long long*)malloc(sizeof(unsigned long long) * cap);
    if (!mask || !wt) return -1;
    for (u64 m = 1; m < ((u64)1 << M); ++m) {
        const int s = __builtin_popcountll(m);
        if (!rep) {
            if (s == d) mask[nc] = m, wt[nc++] = 1;
            continue;
        }
        if (s > d) continue;
        /* number of d-tuples whose set of values is exactly this s-set: surjections [d] -> [s] */
        long long sur = 0, c = 1;
        for (int i = 0; i <= s; ++i) {
            long long pw = 1;
            for (int t = 0; t < d; ++t) pw *= (s - i);
            sur += (i % 2 ? -1 : 1) * c * pw;
            c = c * (s - i) / (i + 1);
        }
        mask[nc] = m, wt[nc++] = (unsigned long long)sur;
    }
    Brute B;
    memset(&B, 0, sizeof B);
    B.n = n, B.M = M, B.nc = (int)nc, B.mask = mask, B.wt = wt;
    B.idx = (int*)malloc(sizeof(int) * (size_t)n);
    B.items = (u64*)malloc(sizeof(u64) * (size_t)n);
    B.fact[0] = 1;
    for (int i = 1; i <= n; ++i) B.fact[i] = B.fact[i - 1] * (unsigned long long)i;
    unsigned long long Q = 0, T = 1;
    for (size_t c = 0; c < nc; ++c) Q += wt[c];
    for (int i = 0; i < n; ++i) T *= Q;
    brute_rec(&B, 0, 0);
    *nbad = B.nbad, *total = T;
    free(B.idx), free(B.items), free(mask), free(wt);
    return 0;
}
