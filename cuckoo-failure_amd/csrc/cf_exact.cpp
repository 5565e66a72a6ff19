// cf_exact.cpp -- exact N_bad from residues and the exact p_fail columns of fast/post_gmp.py.
//
// post_gmp.py (fast/post_gmp.py:51-60, 79-91) merges residues pairwise with GMP, forms
// Q^n = C(M, d)^n, tests N_bad 2^lambda <= Q^n exactly, prints log2(N_bad / Q^n) from 320-bit
// floats and N_bad / Q^n truncated to 40 significant digits. Here, without GMP:
//   * CRT by Garner's mixed radix: N_bad = c_0 + c_1 p_0 + c_2 p_0 p_1 + ... with digits
//     c_j < p_j. For every later modulus p_i the running sums sum_{l<j} c_l P_l and P_j
//     (P_j = p_0 ... p_{j-1}) are kept mod p_i in Montgomery form, so each (j, i) pair costs
//     three 64-bit Montgomery products. The reconstructed integer is unique, so it equals GMP's.
//   * log2 from the binary128 mantissas of N_bad and Q^n (113 bits, far beyond the 15 printed
//     decimals), printed correctly rounded by quadmath.
//   * the 40 digits by exact division: floor(N_bad 10^s / Q^n) for the one s that gives 40 digits.
#include <quadmath.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cuckoo_amd.h"
#include "cf_bignum.hpp"

namespace {

using namespace cfbig;

// Montgomery arithmetic modulo an odd p < 2^63 with R = 2^64; mul(a, b) = a b R^-1 mod p for
// any a < 2^64, b < p.
struct Mont {
    u64 p = 0, ninv = 0, r2 = 0;
    Mont() = default;
    explicit Mont(u64 m) : p(m) {
        u64 inv = m;  // correct to 3 bits for odd m; each Newton step doubles that
        for (int i = 0; i < 5; ++i) inv *= 2 - m * inv;
        ninv = (u64)0 - inv;
        const u64 r = (u64)((((u128)1) << 64) % m);
        r2 = (u64)((u128)r * r % m);
    }
    u64 mul(u64 a, u64 b) const {
        const u128 t = (u128)a * b;
        const u64 q = (u64)t * ninv;
        const u64 r = (u64)((t + (u128)q * p) >> 64);
        return r >= p ? r - p : r;
    }
    u64 mulmod(u64 a, u64 b) const { return mul(mul(a, b), r2); }
};

// a^-1 mod m by the extended Euclidean algorithm; 0 when gcd(a, m) != 1.
u64 inverse(u64 a, u64 m) {
    __int128 t = 0, nt = 1;
    u64 r = m, nr = a % m;
    while (nr) {
        const u64 q = r / nr;
        const __int128 tt = t - (__int128)q * nt;
        t = nt, nt = tt;
        const u64 rr = r - q * nr;
        r = nr, nr = rr;
    }
    if (r != 1) return 0;
    if (t < 0) t += m;
    return (u64)t;
}

// One Garner step for a later modulus i: acc_i += c_j P_j, P_{j+1} = P_j p_j (mod p_i).
struct Later {
    const u64 *p, *ninv, *r2;
    u64 *acc, *Pm;
    void apply(size_t i, u64 cj, u64 pj) const {
        Mont I;
        I.p = p[i], I.ninv = ninv[i], I.r2 = r2[i];
        const u64 a = acc[i] + I.mul(cj, Pm[i]);
        acc[i] = a >= I.p ? a - I.p : a;
        Pm[i] = I.mul(Pm[i], I.mul(pj, I.r2));
    }
};

// The mixed-radix digits of the x in [0, prod p) with x = res[i] mod p[i]; false if the moduli
// are not usable (even, beyond 2^63, or not pairwise coprime). Digits are found in blocks of
// kBlock: inside a block one after another, then the block's updates of every later modulus run
// on several threads (each modulus is independent of the others).
bool garner(const uint64_t* p, const uint64_t* res, size_t k, std::vector<u64>& c, std::string& err) {
    std::vector<u64> pp(p, p + k), ninv(k), r2(k), acc(k, 0), Pm(k);
    for (size_t i = 0; i < k; ++i) {
        if (p[i] < 3 || !(p[i] & 1) || (p[i] >> 63)) {
            err = "modulus " + std::to_string(p[i]) + " is not odd in [3, 2^63)";
            return false;
        }
        const Mont m(p[i]);
        ninv[i] = m.ninv, r2[i] = m.r2;
        Pm[i] = m.mul(m.r2, 1);  // P_0 = 1, as R mod p_i
    }
    const Later L{pp.data(), ninv.data(), r2.data(), acc.data(), Pm.data()};
    constexpr size_t kBlock = 256;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nthreads = k < 4 * kBlock ? 1 : std::min<size_t>(hw, 16);
    c.assign(k, 0);
    for (size_t j0 = 0; j0 < k; j0 += kBlock) {
        const size_t j1 = std::min(k, j0 + kBlock);
        for (size_t j = j0; j < j1; ++j) {
            const Mont J(pp[j]);
            const u64 inv = inverse(J.mul(Pm[j], 1), pp[j]);
            if (!inv) {
                err = "modulus " + std::to_string(pp[j]) + " shares a factor with an earlier one";
                return false;
            }
            const u64 rj = res[j] % pp[j];
            const u64 diff = rj >= acc[j] ? rj - acc[j] : rj + pp[j] - acc[j];
            c[j] = J.mulmod(diff, inv);
            for (size_t i = j + 1; i < j1; ++i) L.apply(i, c[j], pp[j]);
        }
        auto later = [&](size_t a, size_t b) {
            for (size_t i = a; i < b; ++i)
                for (size_t j = j0; j < j1; ++j) L.apply(i, c[j], pp[j]);
        };
        const size_t rest = k - j1;
        if (nthreads == 1 || rest < 2 * kBlock) {
            later(j1, k);
            continue;
        }
        std::vector<std::thread> pool;
        const size_t chunk = (rest + nthreads - 1) / nthreads;
        for (size_t a = j1; a < k; a += chunk) pool.emplace_back(later, a, std::min(k, a + chunk));
        for (std::thread& t : pool) t.join();
    }
    return true;
}

// C(M, d) exactly (0 for M < d)
Nat binomial(int M, int d) {
    if (d < 0 || M < d) return Nat{};
    Nat q = from_u64(1);
    for (int i = 0; i < d; ++i) mul_add_small(q, (u64)(M - i), 0);
    for (int i = 2; i <= d; ++i) divmod_small(q, (u64)i);
    return q;
}

std::string digits40(const Nat& x, const Nat& Qn, double lg) {
    if (x.zero()) return "0";
    const int digits = 40;
    long e = (long)std::floor(lg * 0.30102999566398120);  // a start; the loop settles the exponent
    const Nat ten = from_u64(10);
    for (int guard = 0; guard < 64; ++guard) {
        const long s = digits - 1 - e;
        const Nat q = s >= 0 ? div_floor(mul(x, pow(ten, (u64)s)), Qn) : div_floor(x, mul(Qn, pow(ten, (u64)-s)));
        const std::string t = to_decimal(q);
        if ((int)t.size() == digits) return t.substr(0, 1) + "." + t.substr(1) + "e" + std::to_string(e);
        e += (int)t.size() > digits ? 1 : -1;
    }
    return "?";
}

// x = c_0 + p_0 (c_1 + p_1 (c_2 + ...)) and m = p_0 p_1 ... in one pass from the top digit
void from_digits(const uint64_t* p, const std::vector<u64>& c, Nat& x, Nat& m) {
    x = Nat{};
    m = from_u64(1);
    for (size_t j = c.size(); j-- > 0;) {
        mul_add_small(x, p[j], c[j]);
        mul_add_small(m, p[j], 0);
    }
}

}  // namespace

extern "C" {

long cf_post_need(int n, int M, int d, int pbits) {
    if (n < 0 || d < 0 || pbits < 1 || M < d) return -1;
    double b = 0;
    for (int i = 0; i < d; ++i) b += std::log2((double)(M - i)) - std::log2((double)(i + 1));
    return (long)std::ceil(((double)n * b + 3) / (double)pbits) + 1;
}

int cf_exact_combine(int d, int n, int M, const uint64_t* primes, const uint64_t* res, size_t k, cf_exact* out,
                     char* nbad_dec, size_t dec_len) {
    if (!out || d < 0 || n < 0 || M < 0 || (k && (!primes || !res))) return CF_E_ARGS;
    memset(out, 0, sizeof *out);
    std::vector<u64> c;
    std::string err;
    if (!garner(primes, res, k, c, err)) {
        snprintf(out->error, sizeof out->error, "%s", err.c_str());
        return CF_E_ARGS;
    }
    Nat x, m;
    from_digits(primes, c, x, m);
    const Nat Qn = pow(binomial(M, d), (u64)n);
    out->nbad_bits = (int64_t)bit_length(x);
    out->qn_bits = (int64_t)bit_length(Qn);
    // post_gmp.py:84 asserts m > 4 Q^n (enough primes) and x <= Q^n (a count of outcomes)
    if (cmp(m, shl(Qn, 2)) <= 0) {
        snprintf(out->error, sizeof out->error, "the primes' product does not exceed 4 Q^n");
        return CF_E_PRIMES;
    }
    if (cmp(x, Qn) > 0) {
        snprintf(out->error, sizeof out->error, "the residues combine to more than Q^n");
        return CF_E_ARGS;
    }
    double lgd = -INFINITY;
    if (x.zero()) {
        snprintf(out->log2_pfail, sizeof out->log2_pfail, "-inf");
    } else {
        long ex = 0, eq = 0;
        const __float128 mx = mantissa(x, ex), mq = mantissa(Qn, eq);
        const __float128 lg = (__float128)(ex - eq) + (log2q(mx) - log2q(mq));
        quadmath_snprintf(out->log2_pfail, sizeof out->log2_pfail, "%.15Qf", lg);
        lgd = (double)lg;
    }
    static const int lam[3] = {40, 64, 128};
    for (int t = 0; t < 3; ++t) out->le[t] = cmp(shl(x, (size_t)lam[t]), Qn) <= 0;
    snprintf(out->p_fail, sizeof out->p_fail, "%s", digits40(x, Qn, lgd).c_str());
    if (nbad_dec) {
        const std::string s = to_decimal(x);
        if (s.size() + 1 > dec_len) {
            snprintf(out->error, sizeof out->error, "decimal buffer too small (%zu digits)", s.size());
            return CF_E_ARGS;
        }
        memcpy(nbad_dec, s.c_str(), s.size() + 1);
    }
    return CF_OK;
}

}  // extern "C"
