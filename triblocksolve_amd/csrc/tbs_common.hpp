// tbs_common.hpp -- internal helpers shared by the HIP sources of
// libtriblock_hip.so: context struct, error reporting, launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "triblock_hip.h"

struct tbs_ctx {
    int device = 0;
    int nvar = 9;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // reduction scratch for tbs_norm2 (grown on demand)
    double *red_buf = nullptr;
    size_t red_cap = 0;
    double *red_host = nullptr;   // pinned, 1 value
};

namespace tbs {

void set_error(const char *fmt, ...);

// Record a HIP failure and return its positive status.
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define TBS_HIP(call)                                                        \
    do {                                                                     \
        hipError_t _e = (call);                                              \
        if (_e != hipSuccess) return tbs::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

// Check the launch that was just enqueued.
#define TBS_LAUNCHED(name)                                                   \
    do {                                                                     \
        hipError_t _e = hipGetLastError();                                   \
        if (_e != hipSuccess) return tbs::hip_fail(_e, name, __FILE__, __LINE__); \
    } while (0)

#define TBS_REQUIRE(cond, code, ...)                                         \
    do {                                                                     \
        if (!(cond)) { tbs::set_error(__VA_ARGS__); return (code); }         \
    } while (0)

// Propagate a non-zero status.
#define TBS_TRY(x) do { int _rc = (x); if (_rc) return _rc; } while (0)

// Make the context's device current (every entry point does this first).
int use_ctx(tbs_ctx *ctx);

inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace tbs
