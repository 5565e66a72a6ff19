// cf_crt.hpp -- exact reconstruction of N_bad from its residues modulo 32-bit primes, for the
// cuckoo_fail driver (the reference's include/cuckoo/crt.hpp:15-87 does the same job); the decimal
// N_bad is built on the exact combiner's bignum (cf_bignum.hpp).
//
// mixed_radix() gives the unique digits x with N = x0 + x1 p0 + x2 p0 p1 + ...; any exact method
// yields the same digits. log2_from_digits() must print the same decimal digits as the reference,
// so it follows the reference's floating-point steps exactly (crt.hpp:32-45): a long double
// mantissa folded from the top digit down, renormalised with frexp after every step.
#pragma once
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "cf_bignum.hpp"

namespace cfcli {

inline uint64_t powmod32(uint64_t a, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    a %= m;
    for (; e; e >>= 1) {
        if (e & 1) r = r * a % m;
        a = a * a % m;
    }
    return r;
}

// Digits x_i in [0, p_i): for each i, the value of the lower digits modulo p_i is accumulated
// by Horner's rule from the top, and x_i solves x_i * (p_0 ... p_{i-1}) = r_i - lower (mod p_i).
inline std::vector<uint32_t> mixed_radix(const std::vector<uint32_t>& p, const std::vector<uint32_t>& r) {
    const size_t K = p.size();
    std::vector<uint32_t> x(K);
    for (size_t i = 0; i < K; ++i) {
        const uint64_t pi = p[i];
        uint64_t low = 0, radix = 1;  // value of digits 0..i-1 and p_0...p_{i-1}, both mod p_i
        for (size_t j = i; j-- > 0;) low = (low * (p[j] % pi) + x[j]) % pi;
        for (size_t j = 0; j < i; ++j) radix = radix * (p[j] % pi) % pi;
        const uint64_t diff = (r[i] % pi + pi - low) % pi;
        x[i] = (uint32_t)(diff * powmod32(radix, pi - 2, pi) % pi);
    }
    return x;
}

inline long double log2_from_digits(const std::vector<uint32_t>& p, const std::vector<uint32_t>& x) {
    long double mant = 0;  // N = mant 2^ex
    long ex = 0;
    for (size_t i = x.size(); i-- > 0;) {
        mant = mant * p[i] + std::ldexp(static_cast<long double>(x[i]), static_cast<int>(-ex));
        if (mant != 0) {
            int de;
            mant = std::frexp(mant, &de);
            ex += de;
        }
    }
    if (mant == 0) return -INFINITY;
    return static_cast<long double>(ex) + std::log2(mant);
}

// N itself from its mixed-radix digits, as a natural number of the exact combiner's bignum:
// Horner's rule from the top digit, N <- N p_i + x_i.
inline cfbig::Nat from_mixed_radix(const std::vector<uint32_t>& p, const std::vector<uint32_t>& x) {
    cfbig::Nat v;
    for (size_t i = x.size(); i-- > 0;) cfbig::mul_add_small(v, p[i], x[i]);
    return v;
}

}  // namespace cfcli
