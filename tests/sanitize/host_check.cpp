// host_check.cpp -- the engine's host code under AddressSanitizer and UndefinedBehaviorSanitizer
// (tests only). Reads one command per line on stdin and prints its result; tests/test_sanitize.py
// compares every answer with Python's exact integers.
//   crt32 k x        residues of x modulo the first k 32-bit NTT primes -> mixed radix -> x in decimal
//   exact d n M k x  cf_exact_combine of x modulo 50-bit primes 0..k-1 -> "bits log2 p_fail le0 le1 le2"
//   pow a e          a^e (bignum) in decimal
//   div a b          floor(a / b) (bignum) in decimal
//   primes bits first count   the prime list
//   need n M d bits  the reference's need()
//   slots nm         the tile-slot plan's pool size and segment count
//   consts p lg      per-prime device constants (p^-1 and the root's order checked here)
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cf_bignum.hpp"
#include "cf_crt.hpp"
#include "cf_host.hpp"
#include "cuckoo_amd.h"

using namespace cfbig;

static Nat parse(const std::string& s) {
    Nat a;
    for (char c : s) mul_add_small(a, 10, (u64)(c - '0'));
    return a;
}

static u64 mod_small(Nat a, u64 m) { return divmod_small(a, m); }

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string cmd;
        in >> cmd;
        if (cmd == "crt32") {
            int k;
            std::string x;
            in >> k >> x;
            const std::vector<uint32_t> p = cf::ntt_primes32(k, 1);
            std::vector<uint32_t> r(k);
            for (int i = 0; i < k; ++i) r[i] = (uint32_t)mod_small(parse(x), p[i]);
            const std::vector<uint32_t> dg = cfcli::mixed_radix(p, r);
            std::printf("%s %.6Lf\n", to_decimal(cfcli::from_mixed_radix(p, dg)).c_str(), cfcli::log2_from_digits(p, dg));
        } else if (cmd == "exact") {
            int d, n, M, k;
            std::string x;
            in >> d >> n >> M >> k >> x;
            const std::vector<uint64_t> p = cf::prime_range(50, 0, k);
            std::vector<uint64_t> r(k);
            for (int i = 0; i < k; ++i) r[i] = mod_small(parse(x), p[i]);
            cf_exact out;
            std::vector<char> dec(20 * (size_t)k + 2);
            const int rc = cf_exact_combine(d, n, M, p.data(), r.data(), (size_t)k, &out, dec.data(), dec.size());
            if (rc) std::printf("rc %d %s\n", rc, out.error);
            else
                std::printf("%s %lld %s %s %d %d %d\n", dec.data(), (long long)out.nbad_bits, out.log2_pfail, out.p_fail,
                            out.le[0], out.le[1], out.le[2]);
        } else if (cmd == "pow") {
            std::string a;
            u64 e;
            in >> a >> e;
            std::printf("%s\n", to_decimal(pow(parse(a), e)).c_str());
        } else if (cmd == "div") {
            std::string a, b;
            in >> a >> b;
            std::printf("%s\n", to_decimal(div_floor(parse(a), parse(b))).c_str());
        } else if (cmd == "primes") {
            int bits;
            long first, count;
            in >> bits >> first >> count;
            for (uint64_t p : cf::prime_range(bits, first, count)) std::printf("%llu ", (unsigned long long)p);
            std::printf("\n");
        } else if (cmd == "need") {
            int n, M, d, bits;
            in >> n >> M >> d >> bits;
            std::printf("%ld\n", cf::need(n, M, d, bits));
        } else if (cmd == "slots") {
            int nm;
            in >> nm;
            const cf::SlotPlan sp = cf::plan_slots(nm);
            std::printf("%d %d %d %zu\n", sp.pool, sp.NSEG, sp.NB, sp.segs.size());
        } else if (cmd == "consts") {
            unsigned long long p;
            int lg;
            in >> p >> lg;
            cf::TaskDesc t{};
            if (lg >= 16) cf::wide_constants(p, lg, t);
            else cf::task_constants(p, lg, t);
            // a root of order exactly 2^k has root^(2^(k-1)) = -1: the transform's root (2^lg; the
            // inner 2^(lg-4) and the outer 2^lg of a 16-workgroup task)
            auto order_is = [&](double root, int k) {
                uint64_t r = (uint64_t)(root < 0 ? root + (double)p : root);
                for (int i = 0; i < k - 1; ++i) r = cf::mulmod_u64(r, r, p);
                return r == p - 1;
            };
            const bool ok = lg >= 16 ? order_is(t.root, lg - 4) && order_is(t.wroot, lg) : order_is(t.root, lg);
            std::printf("%d\n", ok ? 1 : 0);
        } else if (!cmd.empty()) {
            std::printf("? %s\n", cmd.c_str());
        }
        std::fflush(stdout);
    }
    return 0;
}
