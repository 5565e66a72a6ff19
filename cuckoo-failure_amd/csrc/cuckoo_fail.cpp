// cuckoo_fail -- the reference's 32-bit command line (src/cuckoo_fail.cpp), computed on the
// MI355X engine:
//
//   cuckoo_fail single   -n N -M M -d D [--replacement] [--algo forward|cubic|ntt] [--exact] [--mod P]
//   cuckoo_fail separate -n N -m m1,m2,...,md [--algo forward|suffix] [--exact] [--mod P]
//
// The prime set (32-bit NTT-friendly primes, largest first, enough for Q^n), the printed lines
// and their formats are those of the reference, so the output is identical apart from the time
// fields. Every residue is computed on the GPU; the host only chooses the primes and reconstructs
// N_bad. --algo ntt (default) runs the suffix-table engine, --algo cubic and --algo forward the
// reference's O(n^3) and O(n^4) algorithms as device kernels (cf_run_algo); `separate` runs the
// separate-tables model's suffix or forward algorithm as device kernels (cf_run_separate).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "cf_cli.hpp"
#include "cf_crt.hpp"
#include "cuckoo_amd.h"

namespace {

void usage() {
    std::puts(
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
        "This build runs every algorithm of both modes on the GPU.");
}

struct Options {
    std::string mode, algo;
    int n = 0, M = 0, d = 0;
    std::vector<int> m;
    bool replacement = false, exact = false;
    long long mod = 0;
};

// "a,b,c" -> {a, b, c}, each piece read as atoi reads it (src/cuckoo_fail.cpp:41-51)
std::vector<int> parse_list(const std::string& s) {
    std::vector<int> out;
    for (size_t pos = 0; pos <= s.size();) {
        size_t comma = s.find(',', pos);
        if (comma == std::string::npos) comma = s.size();
        out.push_back(std::atoi(s.substr(pos, comma - pos).c_str()));
        pos = comma + 1;
    }
    return out;
}

bool is_prime32(long long x) {
    if (x < 2) return false;
    for (long long q = 2; q * q <= x; ++q)
        if (x % q == 0) return false;
    return true;
}

// Argument handling of src/cuckoo_fail.cpp:54-87.
Options parse(int argc, char** argv) {
    if (argc < 2) throw std::invalid_argument("missing mode");
    Options o;
    o.mode = argv[1];
    if (o.mode != "single" && o.mode != "separate") throw std::invalid_argument("mode must be single or separate");
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) throw std::invalid_argument("missing value for " + a);
            return argv[++i];
        };
        if (a == "-n") o.n = std::atoi(next().c_str());
        else if (a == "-M") o.M = std::atoi(next().c_str());
        else if (a == "-d") o.d = std::atoi(next().c_str());
        else if (a == "-m") o.m = parse_list(next());
        else if (a == "--algo") o.algo = next();
        else if (a == "--mod") o.mod = std::atoll(next().c_str());
        else if (a == "--replacement") o.replacement = true;
        else if (a == "--exact") o.exact = true;
        else throw std::invalid_argument("unknown option " + a);
    }
    if (o.mode == "single") {
        if (o.algo.empty()) o.algo = "ntt";
        if (o.algo != "forward" && o.algo != "cubic" && o.algo != "ntt") throw std::invalid_argument("bad --algo");
        if (o.d < 1 || o.n < 1 || o.M <= o.n) throw std::invalid_argument("one table: need d >= 1 and 1 <= n < M");
    } else {
        if (o.algo.empty()) o.algo = "suffix";
        if (o.algo != "forward" && o.algo != "suffix") throw std::invalid_argument("bad --algo");
        // separate_tables.hpp:17-25
        if (o.m.empty()) throw std::invalid_argument("separate tables: need d >= 1 tables");
        long long M = 0;
        for (int x : o.m) {
            if (x < 1) throw std::invalid_argument("separate tables: lengths must be positive");
            M += x;
        }
        if (o.n < 1 || M <= o.n) throw std::invalid_argument("separate tables: need 1 <= n < M");
    }
    if (o.mod && (o.mod < 3 || o.mod >= (1ll << 32) || !is_prime32(o.mod))) throw std::invalid_argument("--mod must be a prime below 2^32");
    return o;
}

// log2 Q as src/cuckoo_fail.cpp:101-111 (long double).
long double log2_Q(const Options& o) {
    long double r = 0;
    if (o.mode == "separate") {
        for (int x : o.m) r += std::log2(static_cast<long double>(x));
    } else if (o.replacement) r = o.d * std::log2(static_cast<long double>(o.M));
    else
        for (int i = 0; i < o.d; ++i) r += std::log2(static_cast<long double>(o.M - i)) - std::log2(static_cast<long double>(i + 1));
    return r;
}

// Q^n in decimal: C(M, d)^n = prod_i (M - i)^n / (d!)^n, every partial quotient an integer.
cfcli::Decimal exact_Qn(const Options& o) {
    cfcli::Decimal b;
    b.d = {1};
    if (o.mode == "separate") {
        for (int x : o.m)
            for (int t = 0; t < o.n; ++t) b.mul_add(x, 0);
    } else if (o.replacement) {
        for (int i = 0; i < o.d * o.n; ++i) b.mul_add(o.M, 0);
    } else {
        for (int i = 0; i < o.d; ++i)
            for (int t = 0; t < o.n; ++t) b.mul_add(o.M - i, 0);
        for (int j = 2; j <= o.d; ++j)
            for (int t = 0; t < o.n; ++t) b.div_exact(j);
    }
    return b;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    try {
        o = parse(argc, argv);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "error: %s\n\n", ex.what());
        usage();
        return 2;
    }
    if (o.mode == "single") {
        std::printf("one table: n=%d items, M=%d slots, d=%d choices, %s\n", o.n, o.M, o.d,
                    o.replacement ? "d independent draws with replacement" : "uniform d-subset of candidate slots");
    } else {
        std::printf("separate tables: n=%d items, m=(", o.n);
        for (size_t j = 0; j < o.m.size(); ++j) std::printf("%s%d", j ? "," : "", o.m[j]);
        std::printf("), d=%zu tables\n", o.m.size());
    }
    const long double bits = o.n * log2_Q(o);
    std::printf("algorithm: %s;  outcomes Q^n = 2^%.3f\n", o.algo.c_str(), static_cast<double>(bits));

    // ---- primes (src/cuckoo_fail.cpp:154-170) ----
    std::vector<uint32_t> primes;
    long double have = 0;
    if (o.mod) {
        primes.push_back(static_cast<uint32_t>(o.mod));
    } else {
        size_t L = 1;
        while (L < (size_t)o.n) L <<= 1;
        if (L < 2) L = 2;
        int min_pow2 = 0;
        while ((size_t(1) << min_pow2) < L) ++min_pow2;
        int count = static_cast<int>(bits / 31) + 2;
        while (true) {
            primes.assign((size_t)count, 0);
            if (cf_ntt_primes32(count, min_pow2, primes.data()) != count) {
                std::fprintf(stderr, "error: not enough NTT-friendly primes\n");
                return 1;
            }
            have = 0;
            for (uint32_t p : primes) have += std::log2(static_cast<long double>(p));
            if (have > bits + 2) break;
            count += static_cast<int>((bits + 2 - have) / 31) + 1;
        }
    }

    // ---- residues on the GPU ----
    cf_engine* e = cfcli::open_engine();
    if (!e) return 1;
    const cfcli::Clock clock;
    std::vector<uint64_t> p64(primes.begin(), primes.end()), res(primes.size());
    const int32_t qn = o.n, qM = o.M;
    const int model = o.replacement ? 1 : 0;
    const std::vector<int32_t> m32(o.m.begin(), o.m.end());
    const int rc = o.mode == "separate"
                       ? cf_run_separate(e, o.algo == "suffix" ? CF_ALGO_SUFFIX : CF_ALGO_FORWARD, o.n, m32.data(),
                                         (int)m32.size(), p64.data(), p64.size(), res.data())
                   : o.algo == "ntt"
                       ? cf_run_fused(e, o.d, model, &qn, &qM, 1, p64.data(), p64.size(), nullptr, res.data())
                       : cf_run_algo(e, o.algo == "cubic" ? CF_ALGO_CUBIC : CF_ALGO_FORWARD, o.d, model, o.n, o.M,
                                     p64.data(), p64.size(), res.data());
    const double secs = clock.secs();
    if (rc != CF_OK) {
        std::fprintf(stderr, "error: %s: %s\n", cf_strerror(rc), cf_last_error(e));
        cf_close(e);
        return cfcli::exit_code(rc);
    }
    cf_close(e);
    std::vector<uint32_t> residues(res.begin(), res.end());

    if (o.mod) {
        std::printf("N_bad mod %u = %u  (%.3f s)\n", primes[0], residues[0], secs);
        return 0;
    }
    std::printf("primes: %zu (product 2^%.1f)  time: %.3f s\n", primes.size(), static_cast<double>(have), secs);
    const std::vector<uint32_t> x = cfcli::mixed_radix(primes, residues);
    const long double log2N = cfcli::log2_from_digits(primes, x);
    if (std::isinf(static_cast<double>(log2N))) {
        std::puts("N_bad = 0,  p_fail = 0");
        return 0;
    }
    const long double log2p = log2N - bits;
    std::printf("p_fail = %.15e\n", static_cast<double>(std::exp2(log2p)));
    std::printf("log2(p_fail) = %.6f,  log10(p_fail) = %.6f\n", static_cast<double>(log2p), static_cast<double>(log2p * std::log10(2.0L)));
    if (o.exact) {
        std::printf("N_bad = %s\n", cfcli::from_digits(primes, x).str().c_str());
        std::printf("Q^n   = %s\n", exact_Qn(o).str().c_str());
    }
    return 0;
}
