// cuckoo_fail -- the 32-bit command line of the reference (src/cuckoo_fail.cpp), computed on the
// MI355X engine:
//
//   cuckoo_fail single   -n N -M M -d D [--replacement] [--algo forward|cubic|ntt] [--exact] [--mod P]
//   cuckoo_fail separate -n N -m m1,m2,...,md [--algo forward|suffix] [--exact] [--mod P]
//
// Kept as the reference has them, because scripts read them: the usage text and error lines, the
// printed lines and their number formats, the prime set (32-bit NTT-friendly primes, largest
// first, as many as the reference's count rule takes), and the long-double steps that produce the
// printed log2 values. Only the time fields differ. Every residue is computed on the GPU: --algo
// ntt (default) by the suffix-table engine, cubic and forward by the reference's O(n^3) and
// O(n^4) algorithms as device kernels; `separate` by the separate-tables kernels. The host picks
// the primes, then rebuilds N_bad from the residues (mixed radix, cf_crt.hpp) and prints.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cf_bignum.hpp"
#include "cf_cli.hpp"
#include "cf_crt.hpp"
#include "cuckoo_amd.h"

namespace {

const char* const kUsage =
    "usage:\n"
    "  cuckoo_fail single   -n N -M M -d D [--replacement] [--algo forward|cubic|ntt] [--exact] [--mod P]\n"
    "  cuckoo_fail separate -n N -m m1,m2,...,md       [--algo forward|suffix]     [--exact] [--mod P]\n"
    "\n"
    "Computes N_bad, the number of candidate-set outcomes that admit no placement,\n"
    "and p_fail = N_bad / Q^n. The algorithms work modulo 32-bit primes; enough\n"
    "primes are used to reconstruct N_bad exactly by Chinese remaindering.\n"
    "  --algo     single: forward O(n^4) | cubic O(n^3) | ntt O(n^2 log n) (default)\n"
    "             separate: forward O(n^{2d+2}) | suffix O(n^{d+2}) (default)\n"
    "  --exact    also print N_bad and Q^n in decimal\n"
    "  --mod P    report N_bad mod P only (single run, no reconstruction)\n"
    "\n"
    "This build runs every algorithm of both modes on the GPU.";

// What was asked for. `separate` is true for the separate-tables model; `lengths` are its tables.
struct Request {
    bool separate = false;
    std::string algo;
    int n = 0, M = 0, d = 0;
    std::vector<int> lengths;
    bool replacement = false, exact = false;
    long long mod = 0;
};

// Numbers are read the way atoi/atoll read them: the leading integer, 0 when there is none.
long long leading_int(const char* s) { return std::strtoll(s, nullptr, 10); }

// "6,7,8" -> {6, 7, 8}; every comma ends a piece, so "6,,8" and "6," have a piece that reads 0.
std::vector<int> split_lengths(const char* s) {
    std::vector<int> out;
    std::string piece;
    for (const char* c = s;; ++c) {
        if (*c == ',' || *c == '\0') {
            out.push_back((int)leading_int(piece.c_str()));
            piece.clear();
            if (*c == '\0') break;
        } else {
            piece += *c;
        }
    }
    return out;
}

bool prime_below_2_32(long long x) {
    if (x < 3 || x >= (1ll << 32) || x % 2 == 0) return false;
    for (long long q = 3; q * q <= x; q += 2)
        if (x % q == 0) return false;
    return true;
}

// Reads the command line into r; on a problem returns the reason (printed as "error: ..."), which
// is the reference's text for the same problem.
std::string read_request(int argc, char** argv, Request& r) {
    if (argc < 2) return "missing mode";
    const std::string mode = argv[1];
    if (mode != "single" && mode != "separate") return "mode must be single or separate";
    r.separate = mode == "separate";
    for (int i = 2; i < argc; ++i) {
        const std::string opt = argv[i];
        if (opt == "--replacement") {
            r.replacement = true;
            continue;
        }
        if (opt == "--exact") {
            r.exact = true;
            continue;
        }
        const bool valued = opt == "-n" || opt == "-M" || opt == "-d" || opt == "-m" || opt == "--algo" || opt == "--mod";
        if (!valued) return "unknown option " + opt;
        if (i + 1 >= argc) return "missing value for " + opt;
        const char* val = argv[++i];
        if (opt == "-n") r.n = (int)leading_int(val);
        if (opt == "-M") r.M = (int)leading_int(val);
        if (opt == "-d") r.d = (int)leading_int(val);
        if (opt == "-m") r.lengths = split_lengths(val);
        if (opt == "--algo") r.algo = val;
        if (opt == "--mod") r.mod = leading_int(val);
    }
    if (!r.separate) {
        if (r.algo.empty()) r.algo = "ntt";
        if (r.algo != "ntt" && r.algo != "cubic" && r.algo != "forward") return "bad --algo";
        // include/cuckoo/single_table.hpp:31-33
        if (r.d < 1 || r.n < 1 || r.M <= r.n) return "one table: need d >= 1 and 1 <= n < M";
    } else {
        if (r.algo.empty()) r.algo = "suffix";
        if (r.algo != "suffix" && r.algo != "forward") return "bad --algo";
        // include/cuckoo/separate_tables.hpp:17-25
        if (r.lengths.empty()) return "separate tables: need d >= 1 tables";
        long long total = 0;
        for (int m : r.lengths) {
            if (m < 1) return "separate tables: lengths must be positive";
            total += m;
        }
        if (r.n < 1 || total <= r.n) return "separate tables: need 1 <= n < M";
    }
    if (r.mod && !prime_below_2_32(r.mod)) return "--mod must be a prime below 2^32";
    return "";
}

// log2 of Q, the number of candidate choices of one item, summed in long double in the
// reference's order (src/cuckoo_fail.cpp:101-111): the printed 2^bits and p_fail digits depend
// on these exact floating-point steps.
long double log2_choices(const Request& r) {
    long double acc = 0;
    if (r.separate) {
        for (int m : r.lengths) acc += std::log2((long double)m);
        return acc;
    }
    if (r.replacement) return r.d * std::log2((long double)r.M);
    for (int i = 0; i < r.d; ++i) acc += std::log2((long double)(r.M - i)) - std::log2((long double)(i + 1));
    return acc;
}

// Q^n exactly: C(M, d)^n, M^(d n), or prod m_t^n.
cfbig::Nat outcomes(const Request& r) {
    cfbig::Nat q = cfbig::from_u64(1);
    if (r.separate) {
        for (int m : r.lengths) cfbig::mul_add_small(q, (uint64_t)m, 0);
    } else if (r.replacement) {
        for (int i = 0; i < r.d; ++i) cfbig::mul_add_small(q, (uint64_t)r.M, 0);
    } else {
        for (int i = 0; i < r.d; ++i) {  // C(M, i+1) = C(M, i) (M - i) / (i + 1), exact at every step
            cfbig::mul_add_small(q, (uint64_t)(r.M - i), 0);
            cfbig::divmod_small(q, (uint64_t)(i + 1));
        }
    }
    return cfbig::pow(q, (uint64_t)r.n);
}

// The 32-bit primes the reference uses for a product of 2^bits: the NTT-friendly list for
// transforms of max(2, next_pow2(n)) (include/cuckoo/modint.hpp:183-196), taken in its order. The
// count follows the reference's rule (src/cuckoo_fail.cpp:158-169), since "primes: K" is printed:
// start with floor(bits/31) + 2, and while the log2 of their product is at most bits + 2, grow by
// floor(shortfall/31) + 1.
struct PrimeSet {
    std::vector<uint32_t> p;
    long double log2_product = 0;
};

bool choose_primes(const Request& r, long double bits, PrimeSet& out) {
    if (r.mod) {
        out.p.assign(1, (uint32_t)r.mod);
        return true;
    }
    int min_pow2 = 1;
    while ((1L << min_pow2) < r.n) ++min_pow2;
    std::vector<uint32_t> list;
    std::vector<long double> sums{0};  // sums[k]: log2 of the product of the first k primes
    auto ensure = [&](int k) {
        if ((int)list.size() >= k) return true;
        list.assign((size_t)k, 0);
        if (cf_ntt_primes32(k, min_pow2, list.data()) != k) return false;
        sums.assign(1, 0);
        for (uint32_t p : list) sums.push_back(sums.back() + std::log2((long double)p));
        return true;
    };
    int k = (int)(bits / 31) + 2;
    for (;;) {
        if (!ensure(k)) return false;
        if (sums[(size_t)k] > bits + 2) break;
        k += (int)((bits + 2 - sums[(size_t)k]) / 31) + 1;
    }
    out.p.assign(list.begin(), list.begin() + k);
    out.log2_product = sums[(size_t)k];
    return true;
}

// The residues N_bad mod p for every prime, on the device.
int device_residues(const Request& r, const PrimeSet& ps, std::vector<uint32_t>& res, double& secs) {
    cf_engine* e = cfcli::open_engine();
    if (!e) return CF_E_DEVICE;
    const cfcli::Clock clock;
    std::vector<uint64_t> p64(ps.p.begin(), ps.p.end()), out(ps.p.size());
    int rc;
    if (r.separate) {
        const std::vector<int32_t> m(r.lengths.begin(), r.lengths.end());
        rc = cf_run_separate(e, r.algo == "suffix" ? CF_ALGO_SUFFIX : CF_ALGO_FORWARD, r.n, m.data(), (int)m.size(),
                             p64.data(), p64.size(), out.data());
    } else if (r.algo == "ntt") {
        const int32_t n = r.n, M = r.M;
        rc = cf_run_fused(e, r.d, r.replacement, &n, &M, 1, p64.data(), p64.size(), nullptr, out.data());
    } else {
        rc = cf_run_algo(e, r.algo == "cubic" ? CF_ALGO_CUBIC : CF_ALGO_FORWARD, r.d, r.replacement, r.n, r.M, p64.data(),
                         p64.size(), out.data());
    }
    secs = clock.secs();
    if (rc != CF_OK) std::fprintf(stderr, "error: %s: %s\n", cf_strerror(rc), cf_last_error(e));
    cf_close(e);
    res.assign(out.begin(), out.end());
    return rc;
}

void print_header(const Request& r, long double bits) {
    if (r.separate) {
        std::string m;
        for (size_t t = 0; t < r.lengths.size(); ++t) m += (t ? "," : "") + std::to_string(r.lengths[t]);
        std::printf("separate tables: n=%d items, m=(%s), d=%zu tables\n", r.n, m.c_str(), r.lengths.size());
    } else {
        std::printf("one table: n=%d items, M=%d slots, d=%d choices, %s\n", r.n, r.M, r.d,
                    r.replacement ? "d independent draws with replacement" : "uniform d-subset of candidate slots");
    }
    std::printf("algorithm: %s;  outcomes Q^n = 2^%.3f\n", r.algo.c_str(), (double)bits);
}

void print_result(const Request& r, long double bits, const PrimeSet& ps, const std::vector<uint32_t>& res, double secs) {
    if (r.mod) {
        std::printf("N_bad mod %u = %u  (%.3f s)\n", ps.p[0], res[0], secs);
        return;
    }
    std::printf("primes: %zu (product 2^%.1f)  time: %.3f s\n", ps.p.size(), (double)ps.log2_product, secs);
    const std::vector<uint32_t> digits = cfcli::mixed_radix(ps.p, res);
    const long double log2_nbad = cfcli::log2_from_digits(ps.p, digits);
    if (std::isinf((double)log2_nbad)) {
        std::puts("N_bad = 0,  p_fail = 0");
        return;
    }
    const long double log2_p = log2_nbad - bits;
    std::printf("p_fail = %.15e\n", (double)std::exp2(log2_p));
    std::printf("log2(p_fail) = %.6f,  log10(p_fail) = %.6f\n", (double)log2_p, (double)(log2_p * std::log10(2.0L)));
    if (r.exact) {
        std::printf("N_bad = %s\n", cfbig::to_decimal(cfcli::from_mixed_radix(ps.p, digits)).c_str());
        std::printf("Q^n   = %s\n", cfbig::to_decimal(outcomes(r)).c_str());
    }
}

}  // namespace

int main(int argc, char** argv) {
    Request r;
    const std::string problem = read_request(argc, argv, r);
    if (!problem.empty()) {
        std::fprintf(stderr, "error: %s\n\n", problem.c_str());
        std::puts(kUsage);
        return 2;
    }
    const long double bits = r.n * log2_choices(r);  // log2 of Q^n
    print_header(r, bits);
    PrimeSet ps;
    if (!choose_primes(r, bits, ps)) {
        std::fprintf(stderr, "error: not enough NTT-friendly primes\n");
        return 1;
    }
    std::vector<uint32_t> res;
    double secs = 0;
    const int rc = device_residues(r, ps, res, secs);
    if (rc != CF_OK) return cfcli::exit_code(rc);
    print_result(r, bits, ps, res, secs);
    return 0;
}
