// cuckoo_fast -- the batch command line of the reference (fast/cuckoo_fast.cu:10-16, 653-904), run
// on the MI355X engine:
//
//   cuckoo_fast D QUERIES FIRST COUNT [--batch B] [--out FILE]
//
// What a user or a script sees is kept: the arguments, the prime list and need(), the resume rule
// for --out, the output lines "i p n M residue" (primes ascending, and for each prime the queries
// it is needed for in file order), the messages of the argument checks and the exit codes
// (2 usage or input, 3 device memory, 4 prime list, 1 device). So an --out file can be continued
// by either program, and fast/post_gmp.py combines it unchanged. stderr carries a banner and one
// progress line per batch.
//
// The program is organised as four steps, each a function below: read the command line, read the
// query file, decide from the ledger of an existing --out file which primes are still to do, then
// run the work in device-sized batches and append their lines.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cf_cli.hpp"
#include "cuckoo_amd.h"

namespace {

// residues modulo primes c 2^20 + 1 below 2^50, or below 2^62 for the build that replaces the
// reference's -DFP_RB=64 one (bin/cuckoo_fast62)
#ifndef CF_PRIME_BITS
#define CF_PRIME_BITS 50
#endif
constexpr int kPrimeBits = CF_PRIME_BITS;
static_assert(kPrimeBits == 50 || kPrimeBits == 62, "the reference's two prime lists");
constexpr int kMaxLog2Rows = 20;

// An early end of the program: the exit status and, unless empty, the line for stderr.
struct Stop {
    int status;
    std::string message;
};

struct CommandLine {
    int d = 0;
    long first = 0, count = 0;
    long batch = 0;  // 0: as many primes as fit in device memory
    const char* queries_path = nullptr;
    const char* out_path = nullptr;
};

// Positional arguments are read the way atoi/atol read them (leading number, 0 otherwise).
long leading_number(const char* s) { return std::strtol(s, nullptr, 10); }

bool read_command_line(int argc, char** argv, CommandLine& cl, Stop& stop) {
    if (argc < 5) {
        stop = {2, "usage: cuckoo_fast D QUERIES FIRST COUNT [--batch B] [--out FILE]"};
        return false;
    }
    cl.d = (int)leading_number(argv[1]);
    cl.queries_path = argv[2];
    cl.first = leading_number(argv[3]);
    cl.count = leading_number(argv[4]);
    int i = 5;
    while (i < argc) {
        const bool has_value = i + 1 < argc;
        if (has_value && std::strcmp(argv[i], "--batch") == 0) {
            cl.batch = leading_number(argv[i + 1]);
            i += 2;
        } else if (has_value && std::strcmp(argv[i], "--out") == 0) {
            cl.out_path = argv[i + 1];
            i += 2;
        } else {
            stop = {2, std::string("unknown option ") + argv[i]};
            return false;
        }
    }
    if (cl.first < 0 || cl.count < 0) {
        stop = {2, "FIRST and COUNT must be non-negative"};
        return false;
    }
    return true;
}

// The query file: whitespace-separated "n M" pairs up to the first token that is not an integer.
struct QuerySet {
    std::vector<int32_t> n, M;
    std::vector<long> need;  // primes 0 .. need-1 carry this query
    int max_n = 0, max_M = 0;
    long max_need = 0;
    size_t size() const { return n.size(); }
    // how many of the queries prime i is needed for
    long records_for(long i) const {
        return (long)std::count_if(need.begin(), need.end(), [i](long k) { return i < k; });
    }
};

bool read_queries(const CommandLine& cl, QuerySet& qs, Stop& stop) {
    FILE* f = std::fopen(cl.queries_path, "r");
    if (!f) {
        stop = {2, std::string("cannot open ") + cl.queries_path};
        return false;
    }
    int n = 0, M = 0;
    bool ok = true;
    while (ok && std::fscanf(f, "%d %d", &n, &M) == 2) {
        if (cl.d < 1 || n < 1 || M <= n) {
            char buf[64];
            std::snprintf(buf, sizeof buf, "bad query %d %d", n, M);
            stop = {2, buf};
            ok = false;
            break;
        }
        qs.n.push_back(n);
        qs.M.push_back(M);
        qs.need.push_back(cf_need(n, M, cl.d, kPrimeBits));
        qs.max_n = std::max(qs.max_n, n);
        qs.max_M = std::max(qs.max_M, M);
        qs.max_need = std::max(qs.max_need, qs.need.back());
    }
    std::fclose(f);
    if (!ok) return false;
    if (qs.size() == 0 || qs.max_n < 8 || qs.max_n >= (1 << kMaxLog2Rows)) {
        stop = {2, "need queries with 8 <= max n < 2^" + std::to_string(kMaxLog2Rows)};
        return false;
    }
    return true;
}

// Record counts per prime index of an existing output file. A prime counts as finished when the
// file holds exactly as many records for its index as there are queries it is needed for
// (fast/cuckoo_fast.cu:697-715); records are read as long as they parse as five numbers, so a
// truncated last line ends the ledger, and records of other (n, M) still count.
class Ledger {
  public:
    void load(FILE* f, long limit) {
        seen_.assign((size_t)std::max(limit, 1L), 0);
        long i;
        unsigned long long p, r;
        int n, M;
        while (std::fscanf(f, "%ld %llu %d %d %llu", &i, &p, &n, &M, &r) == 5)
            if (i >= 0 && i < limit) ++seen_[(size_t)i];
    }
    bool finished(long i, long wanted) const {
        return wanted > 0 && i >= 0 && (size_t)i < seen_.size() && seen_[(size_t)i] == wanted;
    }
    size_t count_finished(const QuerySet& qs, long from, long to) const {
        size_t k = 0;
        for (long i = from; i < to; ++i) k += finished(i, qs.records_for(i));
        return k;
    }

  private:
    std::vector<long> seen_;
};

struct Plan {
    std::vector<long> todo;        // prime indices to compute, ascending
    std::vector<uint64_t> primes;  // primes[i - first] for i in [first, first + count)
    FILE* out = stdout;
};

bool plan_work(const CommandLine& cl, const QuerySet& qs, Plan& plan, Stop& stop) {
    const long end = cl.first + cl.count;
    Ledger ledger;
    bool have_ledger = false;
    if (cl.out_path) {
        if (FILE* f = std::fopen(cl.out_path, "r")) {
            ledger.load(f, end);
            std::fclose(f);
            have_ledger = true;
        }
        plan.out = std::fopen(cl.out_path, "a");
        if (!plan.out) {
            stop = {2, std::string("cannot append to ") + cl.out_path};
            return false;
        }
    }
    for (long i = cl.first; i < std::min(end, qs.max_need); ++i)
        if (!(have_ledger && ledger.finished(i, qs.records_for(i)))) plan.todo.push_back(i);
    if (plan.todo.empty()) {
        const size_t fin = have_ledger ? ledger.count_finished(qs, cl.first, end) : 0;
        stop = {0, "nothing to do (" + std::to_string(fin) + " primes already in the output)"};
        return false;
    }
    plan.primes.resize((size_t)cl.count);
    const long got = cf_prime_range(kPrimeBits, cl.first, cl.count, plan.primes.data());
    plan.primes.resize((size_t)std::max(got, 0L));
    if (plan.primes.empty() || plan.primes.back() < (uint64_t(1) << (kPrimeBits - 1))) {
        stop = {4, "prime list exhausted above 2^" + std::to_string(kPrimeBits - 1)};
        return false;
    }
    return true;
}

// CF_AUTOTUNE=1: time the kernel variants of the run's transform-length class on its largest n.
void autotune_if_asked(cf_engine* e, const QuerySet& qs) {
    const char* tune = std::getenv("CF_AUTOTUNE");
    if (!tune || tune[0] != '1') return;
    const int lg = cf_engine_log2_length(qs.max_n, 1 + (uint64_t(1) << 49));
    cf_tune t;
    if (lg < 10 || lg > 14) return;
    const int n_cal = qs.max_n > (1 << (lg - 1)) ? qs.max_n : 0;
    if (cf_autotune(e, lg, n_cal, &t) == CF_OK)
        std::fprintf(stderr, "autotune 2^%d on n=%d: variant %d of %d (%.1f ms)\n", lg, t.n, t.chosen, t.variants,
                     t.ms[t.chosen]);
}

int run_batches(const CommandLine& cl, const QuerySet& qs, Plan& plan) {
    cf_engine* e = cfcli::open_engine();
    if (!e) return 1;
    // batches of one shape (all but the last) replay one captured graph; CF_GRAPHS=0 turns it off
    const char* graphs = std::getenv("CF_GRAPHS");
    cf_set_graphs(e, !(graphs && graphs[0] == '0'));
    if (kPrimeBits == 50) autotune_if_asked(e, qs);
    const size_t J = qs.size();
    const bool wide = kPrimeBits > 50;
    const double gb = (wide ? cf_mem_per_prime62(qs.max_n, qs.max_M, J) : cf_mem_per_prime(qs.max_n, qs.max_M, J)) / 1073741824.0;
    const long fit = cl.batch > 0 ? cl.batch
                                  : (wide ? cf_max_batch62(e, qs.max_n, qs.max_M, J) : cf_max_batch(e, qs.max_n, qs.max_M, J));
    if (fit < 1) {
        std::fprintf(stderr, "not enough device memory: %.2f GB per prime\n", gb);
        cf_close(e);
        return 3;
    }
    const size_t per_batch = (size_t)std::min<long>(fit, (long)plan.todo.size());
    char dev[128] = "";
    cf_device_name(e, dev, sizeof dev);
    std::fprintf(stderr, "[hip %s] %s residues, %d-bit primes, d=%d, nmax=%d, L=%zu, %.3f GB per prime, batch %zu, %zu primes to do\n",
                 dev, wide ? "64-bit Montgomery" : "exact FP64", kPrimeBits, cl.d, qs.max_n, cf_transform_length((size_t)qs.max_n), gb, per_batch, plan.todo.size());

    const cfcli::Clock clock;
    std::vector<uint64_t> bp, res;
    std::vector<uint8_t> mask;
    for (size_t at = 0; at < plan.todo.size(); at += per_batch) {
        const size_t nb = std::min(per_batch, plan.todo.size() - at);
        const long* idx = plan.todo.data() + at;
        bp.resize(nb);
        res.assign(nb * J, 0);
        mask.resize(nb * J);
        int rows = 0;
        for (size_t b = 0; b < nb; ++b) {
            bp[b] = plan.primes[(size_t)(idx[b] - cl.first)];
            for (size_t j = 0; j < J; ++j) {
                mask[b * J + j] = idx[b] < qs.need[j];
                if (mask[b * J + j]) rows = std::max(rows, (int)qs.n[j]);
            }
        }
        const int rc = cf_run_fused(e, cl.d, 0, qs.n.data(), qs.M.data(), J, bp.data(), nb, mask.data(), res.data());
        if (rc != CF_OK) {
            std::fprintf(stderr, "%s: %s\n", cf_strerror(rc), cf_last_error(e));
            cf_close(e);
            return cfcli::exit_code(rc);
        }
        std::string text;
        char line[128];
        for (size_t b = 0; b < nb; ++b)
            for (size_t j = 0; j < J; ++j) {
                if (!mask[b * J + j]) continue;
                std::snprintf(line, sizeof line, "%ld %llu %d %d %llu\n", idx[b], (unsigned long long)bp[b], (int)qs.n[j],
                              (int)qs.M[j], (unsigned long long)res[b * J + j]);
                text += line;
            }
        std::fputs(text.c_str(), plan.out);
        std::fflush(plan.out);
        const double t = clock.secs();
        std::fprintf(stderr, "primes %ld..%ld (n<=%d): build done at %.1f s, queries at %.1f s\n", idx[0], idx[nb - 1],
                     rows, t, t);
    }
    cf_close(e);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    CommandLine cl;
    QuerySet qs;
    Plan plan;
    Stop stop{0, ""};
    const bool go = read_command_line(argc, argv, cl, stop) && read_queries(cl, qs, stop) && plan_work(cl, qs, plan, stop);
    int status = stop.status;
    if (!go) {
        if (!stop.message.empty()) std::fprintf(stderr, "%s\n", stop.message.c_str());
    } else {
        status = run_batches(cl, qs, plan);
    }
    if (plan.out && plan.out != stdout) std::fclose(plan.out);
    return status;
}
