// cf_bignum.hpp -- natural numbers for the exact combiner: little-endian 64-bit limbs, no
// leading zero limb (zero is the empty vector). Schoolbook arithmetic; the largest numbers here
// (N_bad and Q^n at n = 65536, about 3.2 million bits) take well under a second per product.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace cfbig {

using u64 = uint64_t;
using u128 = unsigned __int128;

struct Nat {
    std::vector<u64> w;
    bool zero() const { return w.empty(); }
    void trim() {
        while (!w.empty() && !w.back()) w.pop_back();
    }
};

inline Nat from_u64(u64 v) {
    Nat a;
    if (v) a.w.push_back(v);
    return a;
}

inline size_t bit_length(const Nat& a) {
    if (a.zero()) return 0;
    return 64 * (a.w.size() - 1) + (64 - __builtin_clzll(a.w.back()));
}

inline int cmp(const Nat& a, const Nat& b) {
    if (a.w.size() != b.w.size()) return a.w.size() < b.w.size() ? -1 : 1;
    for (size_t i = a.w.size(); i-- > 0;)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
}

// a = a * m + c
inline void mul_add_small(Nat& a, u64 m, u64 c) {
    u64 carry = c;
    for (u64& x : a.w) {
        const u128 t = (u128)x * m + carry;
        x = (u64)t;
        carry = (u64)(t >> 64);
    }
    if (carry) a.w.push_back(carry);
    a.trim();
}

inline Nat mul(const Nat& a, const Nat& b) {
    Nat r;
    if (a.zero() || b.zero()) return r;
    r.w.assign(a.w.size() + b.w.size(), 0);
    for (size_t i = 0; i < a.w.size(); ++i) {
        const u64 ai = a.w[i];
        u64 carry = 0;
        u64* out = r.w.data() + i;
        for (size_t j = 0; j < b.w.size(); ++j) {
            const u128 t = (u128)ai * b.w[j] + out[j] + carry;
            out[j] = (u64)t;
            carry = (u64)(t >> 64);
        }
        out[b.w.size()] = carry;
    }
    r.trim();
    return r;
}

inline Nat pow(Nat base, u64 e) {
    Nat r = from_u64(1);
    for (; e; e >>= 1) {
        if (e & 1) r = mul(r, base);
        if (e > 1) base = mul(base, base);
    }
    return r;
}

inline Nat shl(const Nat& a, size_t s) {
    Nat r;
    if (a.zero()) return r;
    const size_t q = s / 64, b = s % 64;
    r.w.assign(a.w.size() + q + 1, 0);
    for (size_t i = 0; i < a.w.size(); ++i) {
        r.w[i + q] |= a.w[i] << b;
        if (b) r.w[i + q + 1] |= a.w[i] >> (64 - b);
    }
    r.trim();
    return r;
}

// a -= b, for a >= b
inline void sub_in(Nat& a, const Nat& b) {
    u64 borrow = 0;
    for (size_t i = 0; i < a.w.size(); ++i) {
        const u64 bi = i < b.w.size() ? b.w[i] : 0;
        const u64 t = a.w[i] - bi;
        const u64 b1 = a.w[i] < bi;
        a.w[i] = t - borrow;
        borrow = b1 | (t < borrow);
        if (i >= b.w.size() && !borrow) break;
    }
    a.trim();
}

// (hi:lo) / m for hi < m: the quotient fits one limb
inline u64 div_2by1(u64 hi, u64 lo, u64 m, u64& rem) {
    u64 q, r;
    __asm__("divq %4" : "=a"(q), "=d"(r) : "a"(lo), "d"(hi), "rm"(m));
    rem = r;
    return q;
}

// a = a / m; returns a mod m (m > 0)
inline u64 divmod_small(Nat& a, u64 m) {
    u64 r = 0;
    for (size_t i = a.w.size(); i-- > 0;) a.w[i] = div_2by1(r, a.w[i], m, r);
    a.trim();
    return r;
}

// floor(a / b) for b > 0 and a small quotient (binary long division: one compare and at most
// one subtraction of a shifted b per quotient bit). The quotient comes back as a Nat.
inline Nat div_floor(Nat a, const Nat& b) {
    Nat q;
    const size_t la = bit_length(a), lb = bit_length(b);
    if (la < lb) return q;
    const size_t top = la - lb;
    q.w.assign(top / 64 + 1, 0);
    for (size_t s = top + 1; s-- > 0;) {
        const Nat t = shl(b, s);
        if (cmp(t, a) <= 0) {
            sub_in(a, t);
            q.w[s / 64] |= (u64)1 << (s % 64);
        }
    }
    q.trim();
    return q;
}

inline std::string to_decimal(Nat a) {
    if (a.zero()) return "0";
    const u64 chunk = 10000000000000000000ull;  // 10^19
    std::vector<u64> parts;
    while (!a.zero()) parts.push_back(divmod_small(a, chunk));
    std::string s = std::to_string(parts.back());
    char buf[24];
    for (size_t i = parts.size() - 1; i-- > 0;) {
        snprintf(buf, sizeof buf, "%019llu", (unsigned long long)parts[i]);
        s += buf;
    }
    return s;
}

// a ~ mant * 2^ex with mant in [1, 2) rounded to binary128 from the top 128 bits of a (a > 0).
// The limbs below those bits change mant by less than 2^-63 of its last place.
inline __float128 mantissa(const Nat& a, long& ex) {
    const size_t n = a.w.size();
    u128 top = (u128)a.w[n - 1] << 64;
    if (n >= 2) top |= a.w[n - 2];
    const int lz = __builtin_clzll(a.w[n - 1]);
    top <<= lz;  // leading 1 at bit 127
    if (lz && n >= 3) top |= a.w[n - 3] >> (64 - lz);
    ex = (long)bit_length(a) - 1;
    return (__float128)top / (__float128)((u128)1 << 127);
}

}  // namespace cfbig
