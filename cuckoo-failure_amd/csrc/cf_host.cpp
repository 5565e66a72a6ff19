// cf_host.cpp -- host-side logic: prime lists, need(), per-prime device constants, tile plan.
#include "cf_host.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace cf {

typedef unsigned __int128 u128;

uint64_t mulmod_u64(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)((u128)a * b % p); }

uint64_t powmod_u64(uint64_t a, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p;
    a %= p;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod_u64(r, a, p);
        a = mulmod_u64(a, a, p);
    }
    return r;
}

// Deterministic Miller-Rabin with the first twelve primes as bases, exact for n < 2^64
// (the same test as fast/cuckoo_fast.cu:403-421, so the prime list is the same).
bool is_prime_u64(uint64_t n) {
    static const uint64_t base[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t b : base)
        if (n % b == 0) return n == b;
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) d >>= 1, ++s;
    for (uint64_t b : base) {
        uint64_t x = powmod_u64(b, d, n);
        if (x == 1 || x == n - 1) continue;
        bool composite = true;
        for (int i = 1; i < s && composite; ++i) {
            x = mulmod_u64(x, x, n);
            composite = x != n - 1;
        }
        if (composite) return false;
    }
    return true;
}

// fast/cuckoo_fast.cu:423-432: walk c = (2^bits - 2) >> 20 downwards, p = c 2^20 + 1.
std::vector<uint64_t> prime_range(int prime_bits, long first, long count) {
    std::vector<uint64_t> out;
    long idx = 0;
    for (uint64_t c = ((uint64_t(1) << prime_bits) - 2) >> 20; c >= 1 && (long)out.size() < count; --c) {
        const uint64_t p = (c << 20) + 1;
        if (!is_prime_u64(p)) continue;
        if (idx++ >= first) out.push_back(p);
    }
    return out;
}

// fast/cuckoo_fast.cu:452-459. The rounding of this long double expression decides which residue
// lines exist, so it is kept exactly: log2 C(M, d) summed in long double and rounded to double,
// then n * that + 3 in long double over PRIME_BITS - 1.
double log2_q(int M, int d) {
    long double b = 0;
    for (int i = 0; i < d; ++i) b += std::log2((long double)(M - i)) - std::log2((long double)(i + 1));
    return (double)b;
}
double log2_q_replacement(int M, int d) { return (double)(d * std::log2((long double)M)); }

long need(int n, int M, int d, int prime_bits) {
    return (long)std::ceil(((long double)n * log2_q(M, d) + 3) / (prime_bits - 1)) + 1;
}

size_t transform_length(size_t nmax) {
    size_t L = 1;
    while (L < nmax) L <<= 1;
    return std::max<size_t>(16, L);
}

// include/cuckoo/modint.hpp:183-196: for k = 31 down to min_pow2, odd c descending, p = c 2^k + 1
// < 2^32 (2-adic valuation of p - 1 exactly k), then sorted largest first.
std::vector<uint32_t> ntt_primes32(int count, int min_pow2) {
    std::vector<uint32_t> out;
    for (int k = 31; k >= min_pow2 && (int)out.size() < count; --k) {
        int64_t c = (int64_t)(((uint64_t(1) << 32) - 2) >> k);
        if (c % 2 == 0) --c;
        for (; c >= 1 && (int)out.size() < count; c -= 2) {
            const uint64_t p = ((uint64_t)c << k) + 1;
            if (is_prime_u64(p)) out.push_back((uint32_t)p);
        }
    }
    if ((int)out.size() < count) throw std::runtime_error("not enough NTT-friendly primes");
    std::sort(out.rbegin(), out.rend());
    return out;
}

int two_adic_valuation(uint64_t x) {
    int v = 0;
    while (x && !(x & 1)) x >>= 1, ++v;
    return v;
}

int engine_log2_length(int nm, uint64_t p) {
    int lg = 10;
    while ((1 << lg) < nm) ++lg;
    if (lg == 15) lg = 16;  // tables of 16385 .. 32768 rows also run on the 2^16 transform
    if (lg > two_adic_valuation(p - 1)) return -1;
    return lg;
}

static double balanced(uint64_t v, uint64_t p) { return v > p / 2 ? (double)v - (double)p : (double)v; }

void task_constants(uint64_t p, int log2L, TaskDesc& d) {
    const uint64_t L = uint64_t(1) << log2L;
    // a quadratic non-residue x gives a root x^((p-1)/L) of order exactly L
    uint64_t x = 2;
    while (powmod_u64(x, (p - 1) / 2, p) != p - 1) ++x;
    const uint64_t w = powmod_u64(x, (p - 1) / L, p), wi = powmod_u64(w, L - 1, p);
    d.p = (double)p;
    d.pinv = 1.0 / (double)p;
    d.root = balanced(w, p);
    d.rooti = balanced(wi, p);
    d.linv = balanced(powmod_u64(L % p, p - 2, p), p);
    const uint64_t w16 = powmod_u64(w, L / 16, p), w16i = powmod_u64(wi, L / 16, p);
    for (int i = 0; i < 8; ++i) {
        d.W[i] = balanced(powmod_u64(w16, i, p), p);
        d.Wi[i] = balanced(powmod_u64(w16i, i, p), p);
    }
}

// A task spread over 16 workgroups: inner constants of length L/16 (from the same non-residue, so
// the inner root is the 16th power of the order-L root), 1/L of the whole transform.
void wide_constants(uint64_t p, int log2L, TaskDesc& d) {
    task_constants(p, log2L - 4, d);  // the inner transform's constants
    const uint64_t L = uint64_t(1) << log2L;
    uint64_t x = 2;
    while (powmod_u64(x, (p - 1) / 2, p) != p - 1) ++x;
    const uint64_t w = powmod_u64(x, (p - 1) / L, p), wi = powmod_u64(w, L - 1, p);
    d.wroot = balanced(w, p);
    d.wrooti = balanced(wi, p);
    d.linv = balanced(powmod_u64(L % p, p - 2, p), p);
}

SlotPlan plan_slots(int nm) {
    SlotPlan sp;
    sp.nm = nm;
    sp.NB = (nm + 1 + 15) / 16;
    sp.NS = 16 * sp.NB;
    sp.NSEG = (sp.NB + 15) / 16;
    const int NB = sp.NB, NSEG = sp.NSEG;
    sp.segs.assign((size_t)NB * NSEG, -1);
    std::vector<int> free_segs;
    int nseg = 0;
    const int nflush = nm >> 4;  // column block X is flushed at the end of step 16X + 15 <= nm - 1
    for (int X = 0; X < nflush; ++X) {
        // row block X was read during steps 16X .. 16X + 15: its segments are dead
        for (int s = 0; 16 * s < X; ++s) free_segs.push_back(sp.segs[(size_t)X * NSEG + s]);
        if (X % 16) continue;
        for (int Y = X + 1; Y < NB; ++Y) {
            int g;
            if (!free_segs.empty()) g = free_segs.back(), free_segs.pop_back();
            else g = nseg++;
            sp.segs[(size_t)Y * NSEG + X / 16] = g;
        }
    }
    sp.pool = 16 * std::max(nseg, 1);
    return sp;
}

size_t task_bytes(const SlotPlan& sp, int F1, int q_cnt, int log2L) {
    // a 2^16 task also holds its exchange rows, outer twiddles and per-workgroup sums
    const int64_t extra = log2L >= 16 ? wide_extra_doubles(q_cnt, log2L) : 0;
    return 8 * (size_t)(arena_doubles(sp.pool, sp.NS, F1, q_cnt, 1 << log2L) + extra);
}

}  // namespace cf
