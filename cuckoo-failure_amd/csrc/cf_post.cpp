// cf_post -- exact N_bad by CRT from residue files and the exact p_fail table, in place of the
// reference's fast/post_gmp.py (which needs gmpy2):
//
//   cf_post D out_file... [--pbits 49] [--dump DIR] > table.tsv
//
// Same arguments (--pbits default 31, 49 for cuckoo_fast's 50-bit primes), same input rule
// (lines of exactly 5 integers "i p n M residue"; a later line for the same (n, M, i) wins), same
// need(), same TSV columns and rows in (n, M) order, INCOMPLETE rows, --dump files
// N_<n>_<M>.txt, and exit status 1 where the script stops with an exception
// (fast/post_gmp.py:21-93). The combining is cf_exact_combine in the engine library.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "cuckoo_amd.h"

namespace {

[[noreturn]] void die(const std::string& msg) {
    std::fflush(stdout);
    std::fprintf(stderr, "cf_post: %s\n", msg.c_str());
    std::exit(1);
}

// A decimal integer as Python's int() reads one (optional sign, surrounding blanks already split off).
bool parse_int(const std::string& s, long long& v) {
    if (s.empty()) return false;
    errno = 0;
    char* end = nullptr;
    v = std::strtoll(s.c_str(), &end, 10);
    return errno == 0 && end == s.c_str() + s.size();
}

bool parse_u64(const std::string& s, unsigned long long& v) {
    if (s.empty() || s[0] == '-') return false;
    errno = 0;
    char* end = nullptr;
    v = std::strtoull(s.c_str(), &end, 10);
    return errno == 0 && end == s.c_str() + s.size();
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> args(argv + 1, argv + argc);
    std::string dump;
    bool have_dump = false;
    int pbits = 31;  // every prime in the files is above 2^pbits
    for (size_t k = 0; k < args.size(); ++k)
        if (args[k] == "--pbits") {
            long long v;
            if (k + 1 >= args.size() || !parse_int(args[k + 1], v)) die("--pbits needs an integer");
            pbits = (int)v;
            args.erase(args.begin() + k, args.begin() + k + 2);
            break;
        }
    for (size_t k = 0; k < args.size(); ++k)
        if (args[k] == "--dump") {
            if (k + 1 >= args.size()) die("--dump needs a directory");
            dump = args[k + 1], have_dump = true;
            args.erase(args.begin() + k, args.begin() + k + 2);
            break;
        }
    long long dll;
    if (args.empty() || !parse_int(args[0], dll))
        die("usage: cf_post D out_file... [--pbits 49] [--dump DIR] > table.tsv");
    const int d = (int)dll;

    // (n, M) -> prime index -> (p, residue)
    std::map<std::pair<long long, long long>, std::map<long long, std::pair<uint64_t, uint64_t>>> res;
    for (size_t a = 1; a < args.size(); ++a) {
        std::ifstream f(args[a]);
        if (!f) die("cannot open " + args[a]);
        std::string line;
        while (std::getline(f, line)) {
            std::istringstream ss(line);
            std::vector<std::string> t;
            for (std::string w; ss >> w;) t.push_back(w);
            if (t.size() != 5) continue;
            long long i, n, M;
            unsigned long long p, r;
            if (!parse_int(t[0], i) || !parse_int(t[2], n) || !parse_int(t[3], M) || !parse_u64(t[1], p) ||
                !parse_u64(t[4], r))
                die("bad residue line in " + args[a] + ": " + line);
            res[{n, M}][i] = {p, r};
        }
    }

    std::printf("n\tM\tratio\tlog2_pfail\tle_2^-40\tle_2^-64\tle_2^-128\tp_fail\tprimes\n");
    std::vector<char> dec;
    for (const auto& kv : res) {
        const long long n = kv.first.first, M = kv.first.second;
        const auto& got = kv.second;
        if (n == 0) die("division by zero (n = 0)");
        const long k = cf_post_need((int)n, (int)M, d, pbits);
        if (k < 0) die("need() undefined for M = " + std::to_string(M) + " < d");
        char ratio[64];
        std::snprintf(ratio, sizeof ratio, "%.6f", (double)M / (double)n);
        std::vector<uint64_t> ps, rs;
        bool full = true;
        for (long i = 0; i < k && full; ++i) {
            auto it = got.find(i);
            if (it == got.end()) full = false;
            else ps.push_back(it->second.first), rs.push_back(it->second.second);
        }
        if (!full) {
            std::printf("%lld\t%lld\t%s\tINCOMPLETE %zu/%ld\n", n, M, ratio, got.size(), k);
            continue;
        }
        cf_exact ex;
        if (have_dump) dec.assign(20 * (size_t)k + 2, 0);
        const int rc = cf_exact_combine(d, (int)n, (int)M, ps.data(), rs.data(), ps.size(), &ex,
                                        have_dump ? dec.data() : nullptr, dec.size());
        if (rc != CF_OK) die("AssertionError (" + std::to_string(n) + ", " + std::to_string(M) + "): " + ex.error);
        std::printf("%lld\t%lld\t%s\t%s\t%s\t%s\t%s\t%s\t%ld\n", n, M, ratio, ex.log2_pfail, ex.le[0] ? "yes" : "no",
                    ex.le[1] ? "yes" : "no", ex.le[2] ? "yes" : "no", ex.p_fail, k);
        std::fflush(stdout);
        if (have_dump) {
            const std::string fn = dump + "/N_" + std::to_string(n) + "_" + std::to_string(M) + ".txt";
            FILE* o = std::fopen(fn.c_str(), "w");
            if (!o) die("cannot write " + fn);
            std::fprintf(o, "N_bad=%s\nQ^n=C(%lld,%d)^%lld\n", dec.data(), M, d, n);
            std::fclose(o);
        }
    }
    return 0;
}
