// cf_cli.hpp -- small helpers shared by the two command-line drivers.
#pragma once
#include <chrono>
#include <cstdio>

#include "cuckoo_amd.h"

namespace cfcli {

// Opens the engine on device 0 (HIP_VISIBLE_DEVICES selects the GPU), or reports why not.
inline cf_engine* open_engine() {
    cf_engine* e = nullptr;
    const int rc = cf_open(0, &e);
    if (rc != CF_OK) {
        std::fprintf(stderr, "HIP error: cannot open device 0 (%s%s%s)\n", cf_strerror(rc), e ? ": " : "", e ? cf_last_error(e) : "");
        return nullptr;
    }
    return e;
}

// Exit status for an engine error, matching the reference's codes (1 device, 2 arguments, 3 memory).
inline int exit_code(int rc) {
    switch (rc) {
        case CF_E_ARGS: return 2;
        case CF_E_NOMEM: return 3;
        case CF_E_PRIMES: return 4;
        default: return 1;
    }
}

struct Clock {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double secs() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

}  // namespace cfcli
