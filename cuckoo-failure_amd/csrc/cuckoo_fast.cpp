// cuckoo_fast -- the reference's batch command line (fast/cuckoo_fast.cu:10-16, 653-904) on the
// MI355X engine:
//
//   cuckoo_fast D QUERIES FIRST COUNT [--batch B] [--out FILE]
//
// Same arguments, prime list, need(), resume rule, output lines "i p n M residue" (in the same
// order: primes ascending, queries in file order) and exit codes (2 usage/input, 3 device memory,
// 4 prime list, 1 device error), so an --out file can be continued by either program and
// fast/post_gmp.py combines it unchanged. stderr carries a banner and per-batch progress.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "cf_cli.hpp"
#include "cuckoo_amd.h"

namespace {

constexpr int kPrimeBits = 50;
constexpr int kTwoAdic = 20;

struct Query {
    int n, M;
    long need;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: cuckoo_fast D QUERIES FIRST COUNT [--batch B] [--out FILE]\n");
        return 2;
    }
    const int d = std::atoi(argv[1]);
    const long first = std::atol(argv[3]), count = std::atol(argv[4]);
    long batch = 0;
    const char* out_path = nullptr;
    for (int i = 5; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--batch") && i + 1 < argc) batch = std::atol(argv[++i]);
        else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) out_path = argv[++i];
        else {
            std::fprintf(stderr, "unknown option %s\n", argv[i]);
            return 2;
        }
    }
    if (first < 0 || count < 0) {
        std::fprintf(stderr, "FIRST and COUNT must be non-negative\n");
        return 2;
    }

    // ---- queries (one "n M" pair per line) ----
    std::vector<Query> qs;
    int nmax = 0, Mmax = 0;
    long maxneed = 0;
    {
        FILE* f = std::fopen(argv[2], "r");
        if (!f) {
            std::fprintf(stderr, "cannot open %s\n", argv[2]);
            return 2;
        }
        int n, M;
        while (std::fscanf(f, "%d %d", &n, &M) == 2) {
            if (n < 1 || M <= n || d < 1) {
                std::fprintf(stderr, "bad query %d %d\n", n, M);
                std::fclose(f);
                return 2;
            }
            qs.push_back({n, M, cf_need(n, M, d, kPrimeBits)});
            nmax = std::max(nmax, n);
            Mmax = std::max(Mmax, M);
            maxneed = std::max(maxneed, qs.back().need);
        }
        std::fclose(f);
    }
    if (qs.empty() || nmax < 8 || nmax + 1 > (1 << kTwoAdic)) {
        std::fprintf(stderr, "need queries with 8 <= max n < 2^%d\n", kTwoAdic);
        return 2;
    }

    // ---- resume: prime i is done iff the file has exactly as many records for i as queries need it ----
    std::set<long> done;
    FILE* out = stdout;
    if (out_path) {
        if (FILE* f = std::fopen(out_path, "r")) {
            std::vector<long> cnt((size_t)std::max<long>(first + count, 1), 0);
            long i;
            unsigned long long p, r;
            int n, M;
            while (std::fscanf(f, "%ld %llu %d %d %llu", &i, &p, &n, &M, &r) == 5)
                if (i >= 0 && i < (long)cnt.size()) ++cnt[i];
            std::fclose(f);
            for (long i2 = first; i2 < first + count; ++i2) {
                long want = 0;
                for (const Query& q : qs) want += i2 < q.need;
                if (want > 0 && cnt[i2] == want) done.insert(i2);
            }
        }
        out = std::fopen(out_path, "a");
        if (!out) {
            std::fprintf(stderr, "cannot append to %s\n", out_path);
            return 2;
        }
    }
    std::vector<long> todo;
    for (long i = first; i < first + count && i < maxneed; ++i)
        if (!done.count(i)) todo.push_back(i);
    if (todo.empty()) {
        std::fprintf(stderr, "nothing to do (%zu primes already in the output)\n", done.size());
        return 0;
    }
    std::vector<uint64_t> primes((size_t)count);
    primes.resize((size_t)std::max<long>(cf_prime_range(kPrimeBits, first, count, primes.data()), 0));
    if (primes.empty() || primes.back() < (uint64_t(1) << (kPrimeBits - 1))) {
        std::fprintf(stderr, "prime list exhausted above 2^%d\n", kPrimeBits - 1);
        return 4;
    }

    // ---- engine and batch size ----
    cf_engine* e = cfcli::open_engine();
    if (!e) return 1;
    // batches of one shape (all but the last) replay one captured graph; CF_GRAPHS=0 turns it off
    const char* graphs = std::getenv("CF_GRAPHS");
    cf_set_graphs(e, !(graphs && graphs[0] == '0'));
    const size_t J = qs.size();
    // CF_AUTOTUNE=1: time the kernel variants of this run's transform length on its largest n
    const char* tune = std::getenv("CF_AUTOTUNE");
    if (tune && tune[0] == '1') {
        const int lg = cf_engine_log2_length(nmax, 1 + (uint64_t(1) << 49));
        cf_tune t;
        if (lg >= 10 && lg <= 14 && cf_autotune(e, lg, nmax > (1 << (lg - 1)) ? nmax : 0, &t) == CF_OK)
            std::fprintf(stderr, "autotune 2^%d on n=%d: variant %d of %d (%.1f ms)\n", lg, t.n, t.chosen, t.variants,
                         t.ms[t.chosen]);
    }
    const size_t per_prime = cf_mem_per_prime(nmax, Mmax, J);
    if (batch <= 0) batch = cf_max_batch(e, nmax, Mmax, J);
    if (batch < 1) {
        std::fprintf(stderr, "not enough device memory: %.2f GB per prime\n", per_prime / 1073741824.0);
        cf_close(e);
        return 3;
    }
    const size_t B = (size_t)std::min<long>(batch, (long)todo.size());
    char dev[128] = "";
    cf_device_name(e, dev, sizeof dev);
    std::fprintf(stderr, "[hip %s] exact FP64 residues, %d-bit primes, d=%d, nmax=%d, L=%zu, %.3f GB per prime, batch %zu, %zu primes to do\n",
                 dev, kPrimeBits, d, nmax, cf_transform_length((size_t)nmax), per_prime / 1073741824.0, B, todo.size());

    std::vector<int32_t> qn(J), qM(J);
    for (size_t j = 0; j < J; ++j) qn[j] = qs[j].n, qM[j] = qs[j].M;
    const cfcli::Clock clock;
    for (size_t at = 0; at < todo.size(); at += B) {
        const size_t nb = std::min(B, todo.size() - at);
        std::vector<uint64_t> bp(nb), res(nb * J);
        std::vector<uint8_t> mask(nb * J);
        int nm = 0;
        for (size_t b = 0; b < nb; ++b) {
            bp[b] = primes[(size_t)(todo[at + b] - first)];
            for (size_t j = 0; j < J; ++j) {
                mask[b * J + j] = todo[at + b] < qs[j].need;
                if (mask[b * J + j]) nm = std::max(nm, qs[j].n);
            }
        }
        const int rc = cf_run_fused(e, d, 0, qn.data(), qM.data(), J, bp.data(), nb, mask.data(), res.data());
        if (rc != CF_OK) {
            std::fprintf(stderr, "%s: %s\n", cf_strerror(rc), cf_last_error(e));
            cf_close(e);
            return cfcli::exit_code(rc);
        }
        std::string lines;
        char line[160];
        for (size_t b = 0; b < nb; ++b)
            for (size_t j = 0; j < J; ++j)
                if (mask[b * J + j]) {
                    std::snprintf(line, sizeof line, "%ld %llu %d %d %llu\n", todo[at + b], (unsigned long long)bp[b], qs[j].n,
                                  qs[j].M, (unsigned long long)res[b * J + j]);
                    lines += line;
                }
        std::fputs(lines.c_str(), out);
        std::fflush(out);
        const double t = clock.secs();
        std::fprintf(stderr, "primes %ld..%ld (n<=%d): build done at %.1f s, queries at %.1f s\n", todo[at], todo[at + nb - 1], nm, t, t);
    }
    if (out != stdout) std::fclose(out);
    cf_close(e);
    return 0;
}
