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
    // Optional device region that plan allocations are carved from (bump,
    // 256-byte aligned) instead of hipMalloc; set by the serving engine around
    // one request's admission.  Region memory is never freed piecewise.
    char *region = nullptr;
    size_t region_size = 0, region_used = 0;
    // Second stream (default priority, created on first use) for work that
    // runs beside the main stream inside one call and joins back before the
    // call's later work on the main stream (the factorisation's block packing).
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // Pinned host staging for table uploads of small plans: a ring of
    // buffers (grown on demand, reused; stage_ev[i] marks the end of the last
    // upload read from buffer i), so building one plan's tables does not wait
    // for the previous plan's upload, which is queued behind earlier work.
    static constexpr int kStages = 3;
    char *stage[kStages] = {};
    size_t stage_cap[kStages] = {};
    hipEvent_t stage_ev[kStages] = {};
    bool stage_busy[kStages] = {};
    int stage_cur = 0;
    // Device buffer for small per-call tables (grown on demand; reused in
    // stream order on the context's stream).
    char *dstage = nullptr;
    size_t dstage_cap = 0;
};

namespace tbs {

void set_error(const char *fmt, ...);

// The context's second stream and its fork/join events, created on first use.
int aux_stream(tbs_ctx *ctx, hipStream_t *out);

// The context's pinned staging buffer, at least `bytes` long, once the last
// upload from it has completed; the caller queues its upload on the
// context's stream and then calls stage_release.
int stage_acquire(tbs_ctx *ctx, size_t bytes, char **out);
int stage_release(tbs_ctx *ctx);

// The context's device table buffer, at least `bytes` long.
int dev_stage(tbs_ctx *ctx, size_t bytes);

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

// Device allocation for library-owned arrays: from the context's region when
// one is set (*from_region = true), else hipMalloc.  TBS_ERR_NOMEM when
// neither can hold `bytes`.
int dev_alloc(tbs_ctx *ctx, size_t bytes, void **ptr, bool *from_region);
constexpr size_t kRegionAlign = 256;

// Internal access to a plan for the serving engine (same library).
// tbs_plan_create, with the mesh consistency checks skipped when !validate
// (for meshes the library built itself).
// structured_nk > 0: m is tbs_structured_mesh's grid of m->nline lines of
// structured_nk elements given by its sizes and line arrays only (epoint, ef,
// fc, lines, lineface may be null); the plan generates those on the device.
int plan_create(tbs_ctx *ctx, const tbs_mesh_host *m, int precision, tbs_plan **out, bool validate,
                int structured_nk = 0);
// tbs_plan_footprint, with the same choices (structured: such a grid).
int plan_footprint(const tbs_mesh_host *mesh, int precision, size_t *bytes, bool validate,
                   bool structured = false);
// Scatter the plan's solution into U (device, element order) on stream st.
int plan_scatter_U(tbs_plan *p, hipStream_t st, double *U);
// Device residual history [kHistCap] of a plan.
const double *plan_hist(const tbs_plan *p);
// Bytes the plan's factorisation draws from a context region.
size_t plan_factor_workspace(int nvar, int nelem);
// An empty batch (block size nvar, precision prec) whose device tables and
// pinned staging already hold nsys systems, nlines lines and ngroups
// residual groups, so that tbs_batch_set_plans within those sizes allocates
// nothing (the serving engine creates its batch this way).
int batch_create_reserved(tbs_ctx *ctx, int nvar, int prec, size_t nsys, size_t nlines, size_t ngroups,
                          tbs_batch **out);

// Internal-order vectors of a factored plan (n elements x nvar values each,
// line-major order) for solvers built on the plan's kernels.
struct PlanVecs {
    tbs_ctx *ctx;      // the plan's context (device, stream)
    double *rhs, *U, *R;
    long n;          // elements
    int nv;
    const int *lines;  // internal element m holds element lines[m]
};
int plan_vecs(tbs_plan *p, PlanVecs *out);
// Mark the plan active so its sweep kernels run (state reset, no history).
int plan_activate(tbs_plan *p, hipStream_t st);
// R = rhs + J U over the plan's internal vectors; also g = D'^-1(-R) for the
// next Thomas pass.  R is stored only when want_R; u_zero: U is known to be
// zero, so R = rhs and the J U pass is skipped.
int plan_residual(tbs_plan *p, hipStream_t st, int want_R, bool u_zero = false);
// U += d with M d = -R (block-Thomas on the g of the last residual pass).
int plan_thomas(tbs_plan *p, hipStream_t st);
// Element order <-> internal order for one vector.
int plan_gather(tbs_plan *p, hipStream_t st, const double *src, double *dst_internal);
int plan_scatter(tbs_plan *p, hipStream_t st, const double *src_internal, double *dst);

inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace tbs
