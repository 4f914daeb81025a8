/*
 * triblock_hip.h -- C ABI of the MI355X (gfx950) line-implicit Jacobi (LIJ)
 * block solver library, libtriblock_hip.so.
 *
 * Conventions
 *   - Every entry point returns int: 0 (TBS_OK) on success, a negative
 *     TBS_ERR_* code for an argument / state error, or a positive hipError_t
 *     value when a HIP call failed.  tbs_last_error() gives a message for the
 *     calling thread.  Nothing in the library calls exit().
 *   - "device" pointers are HIP device allocations owned by the caller; the
 *     library never frees them.  "host" pointers are ordinary host memory.
 *   - Block arrays use the reference layout: entry (i,j) of block b at
 *     i + nvar*j + nvar*nvar*b (column-major blocks, OKL @dim(NVAR,NVAR,n)).
 *     Vectors: entry i of element e at i + nvar*e.  Indices are 0-based int32
 *     (after the reference reader's 1 -> 0 shift); boundary faces are < 0.
 *   - Launches are asynchronous on the context's stream, like OCCA's default
 *     stream; tbs_ctx_synchronize() waits for them.
 *   - A context is thread-compatible (one host thread at a time).
 */
#ifndef TRIBLOCK_HIP_H
#define TRIBLOCK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TBS_OK              0
#define TBS_ERR_ARG        -1   /* bad argument (null pointer, size, nvar)   */
#define TBS_ERR_STATE      -2   /* call out of order (e.g. solve before factor) */
#define TBS_ERR_UNSUPPORTED -3  /* nvar / mode not built into this library    */
#define TBS_ERR_NOMEM      -4   /* admission / allocation refused              */
#define TBS_ERR_IO         -5   /* file missing, truncated or inconsistent     */
#define TBS_ERR_NUMERIC    -6   /* non-finite residual or divergence guard     */

#define TBS_FP64 0              /* factors and Thomas sweeps in FP64           */
#define TBS_FP32 1              /* factors and Thomas sweeps in FP32; residual,
                                   update and norm stay FP64                   */

typedef struct tbs_ctx tbs_ctx;
typedef struct tbs_plan tbs_plan;
typedef struct tbs_batch tbs_batch;

const char *tbs_last_error(void);
const char *tbs_version(void);
/* Block sizes (nvar) the drop-in kernels / the fused plan path are built for:
 * writes up to cap values, returns the count. */
int tbs_supported_nvar(int *out, int cap);
int tbs_plan_supported_nvar(int *out, int cap);

/* ---------------------------------------------------------------------- */
/* Context: replaces Platform (include/Platform.hxx:25-92,                */
/* src/Platform.cxx:11-78) -- device selection and the launch stream.     */
/* ---------------------------------------------------------------------- */
int tbs_ctx_create(int device_id, int nvar, tbs_ctx **out);
int tbs_ctx_destroy(tbs_ctx *ctx);
/* Use an existing HIP stream (hipStream_t) instead of the context's own. */
int tbs_ctx_set_stream(tbs_ctx *ctx, void *hip_stream);
void *tbs_ctx_stream(tbs_ctx *ctx);
int tbs_ctx_synchronize(tbs_ctx *ctx);
/* Wait for all work on the context's device (every stream). */
int tbs_device_synchronize(tbs_ctx *ctx);
int tbs_ctx_nvar(tbs_ctx *ctx);
/* Device memory helpers for hosts without another allocator (the reference
 * allocates through occa::device::malloc, src/Mesh.cxx:75-84,
 * src/Jacobian.cxx:240-259). */
int tbs_malloc(tbs_ctx *ctx, size_t bytes, void **dptr);
int tbs_free(tbs_ctx *ctx, void *dptr);
int tbs_memcpy_h2d(tbs_ctx *ctx, void *dst, const void *src, size_t bytes);
int tbs_memcpy_d2h(tbs_ctx *ctx, void *dst, const void *src, size_t bytes);
int tbs_memset(tbs_ctx *ctx, void *dst, int value, size_t bytes);
/* Device-to-device copy on the context's stream, finished on return: the
 * copy behind the reference's Platform::malloc<T>(entries, occa::memory src)
 * (include/Platform.hxx:70-75). */
int tbs_memcpy_d2d(tbs_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Pinned (page-locked) host memory: the reference's Platform::hostMalloc
 * allocates with the OCCA property "host": true (include/Platform.hxx:82-88).
 * bytes == 0 gives *hptr = NULL. */
int tbs_host_malloc(tbs_ctx *ctx, size_t bytes, void **hptr);
int tbs_host_free(tbs_ctx *ctx, void *hptr);
/* Timing events on the context's stream (hipEvent_t behind an opaque handle;
 * the reference times with OCCA stream tags, src/main.cxx:181-228).  They are
 * created without the system-scope fence a default event record performs (a
 * cache writeback and invalidate that would slow the kernels being timed), so
 * they order and time work but do not publish device writes to the host:
 * synchronise the context for that. */
int tbs_event_create(tbs_ctx *ctx, void **ev);
int tbs_event_destroy(tbs_ctx *ctx, void *ev);
int tbs_event_record(tbs_ctx *ctx, void *ev);
int tbs_event_elapsed_ms(void *ev_start, void *ev_stop, float *ms);
/* Achievable HBM bandwidth of the context's device: a read-only stream and a
 * copy (16-byte loads and stores) over `bytes` bytes of device memory the
 * call allocates and frees, each timed over `reps` launches after one
 * warm-up in 32 launch shapes (grid-stride or contiguous 256 KiB runs per
 * workgroup, 4 or 8 loads in flight per thread, 4 to 32 workgroups per CU,
 * streaming or default cache policy); the fastest shape counts.  gbps[0] = read GB/s, gbps[1] = copy GB/s (bytes read + bytes
 * written).  Synchronous.  No counterpart in the
 * reference: it is the ceiling the sweep kernels' GB/s (the reference prints
 * GB/s at src/main.cxx:243-252) are compared with. */
int tbs_stream_ceiling(tbs_ctx *ctx, size_t bytes, int reps, double *gbps);

/* ---------------------------------------------------------------------- */
/* Drop-in kernels: one entry per active reference OKL kernel, same       */
/* argument order, all pointers device pointers.                          */
/* ---------------------------------------------------------------------- */
/* B[n] = A[n], n < nelem*nvar*nvar.      src/okl/linesmoothLU_v5.okl:85-93  */
int tbs_copyAtoBjac(tbs_ctx *ctx, int nelem, const double *A, double *B);
/* B[n] = A[n], n < nelem*nvar.           src/okl/linesmoothLU_v5.okl:75-83  */
int tbs_copyAtoB(tbs_ctx *ctx, int nelem, const double *A, double *B);
/* B[n] += A[n], n < nelem*nvar.          src/okl/linesmoothLU_v5.okl:95-103 */
int tbs_addAtoB(tbs_ctx *ctx, int nelem, const double *A, double *B);

/* Packed in-line lower blocks by element id (host loop in the reference,
 * here a device gather):  A[e] = (e==fc[2f]) ? jacO2[f] : jacO1[f] with
 * f = lineface[m], for every line position k >= 1.
 * Replaces Jacobian::assembleTriBlocks, src/Jacobian.cxx:76-117.          */
int tbs_assembleTriBlocks(tbs_ctx *ctx, int nline, const int *linesize,
                          const int *linepoint, const int *lines,
                          const int *lineface, const int *fc,
                          const double *jacO1, const double *jacO2,
                          double *A);

/* Block-Thomas factorisation, in place: Dia holds D on entry and the LU
 * factors D' on exit; DinvC[elast] = D'^-1 C for every non-last element.
 * Replaces lineLU, src/okl/lineLU_v2.okl:58-69.                            */
int tbs_lineLU(tbs_ctx *ctx, int nelem, int nintfaces, int nlines,
               const int *fc, const int *linesize, const int *linepoint,
               const int *lines, const int *lineface, double *Dia,
               const double *Of1, const double *Of2, double *DinvC);

/* Thomas forward elimination + back substitution: dU = -M^-1 R per line.
 * Replaces triblock_solveDU, src/okl/linesmoothLU_TriBlock_v3.okl:92-104. */
int tbs_triblock_solveDU(tbs_ctx *ctx, int nelem, int nintfaces, int eftot,
                         int nlines, int linelemtot, const int *linesize,
                         const int *linepoint, const int *lines,
                         const double *Dia, const double *DinvC,
                         const double *A, double *dU, const double *R);

/* Residual accumulation R(e) += D(e)U(e) + sum_faces Off(f)U(nb); the
 * caller presets R = rhs (as src/main.cxx:217 does).
 * Replaces triblock_lineRes, src/okl/linesmoothLU_TriBlock_v3.okl:228-243. */
int tbs_triblock_lineRes(tbs_ctx *ctx, int nelem, int nintfaces, int eftot,
                         int nlines, int linelemtot, const int *epoint,
                         const int *ef, const int *fc, const int *linesize,
                         const int *linepoint, const int *lines,
                         const double *Dia, const double *Of1,
                         const double *Of2, const double *U, double *R);

/* b(:,e) = L_e U_e x(:,e) for every element e < nelem, with L_e \ U_e the
 * packed LU factors of block e (unit-lower L strictly below the diagonal, U
 * on and above it, as tbs_lineLU leaves Dia).  The reference's dgemvLU
 * helper, src/okl/linesmoothLU_TriBlock_v3.okl:48-69, applied per block. */
int tbs_dgemvLU(tbs_ctx *ctx, int nelem, const double *LU, const double *x, double *b);

/* ||v||_2 of n doubles (device), deterministic fixed-order reduction; the
 * reference has no norm (build-defined convergence measure).              */
int tbs_norm2(tbs_ctx *ctx, long n, const double *v, double *out_host);
/* tbs_norm2 with no device allocation: its partials live in a static device
 * array of the library (512 KiB per GPU, held until the process ends), so the
 * call makes no hipMalloc/hipFree.  No counterpart in the reference; it is
 * for callers that cannot allocate (a serving loop, a harness that forbids
 * hipMalloc).  Limit: n <= 2^29 values (TBS_ERR_ARG beyond).  The result is
 * bitwise tbs_norm2's.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (two threads, two contexts) race. */
int tbs_norm2_noalloc(tbs_ctx *ctx, long n, const double *v, double *out_host);

/* ---------------------------------------------------------------------- */
/* Fused LIJ path ("plan"): the outer loop of src/main.cxx:180-229 as     */
/* graph-captured HIP kernels over a device-resident, re-laid-out copy of  */
/* the system.  Mesh topology is given as HOST arrays (the Mesh class's    */
/* host vectors, include/Mesh.hxx:41-48); Jacobian values as DEVICE        */
/* arrays in the reference layout (Jacobian::o_jacD/o_jacO1/o_jacO2).      */
/* ---------------------------------------------------------------------- */
typedef struct tbs_mesh_host {
    int nvar, nelem, nintface, eftot, nline, nlineelem;
    const int *epoint;     /* [nelem+1]      */
    const int *ef;         /* [eftot]        */
    const int *fc;         /* [2*nintface]   */
    const int *linesize;   /* [nline]        */
    const int *linepoint;  /* [nline+1]      */
    const int *lines;      /* [nlineelem]    */
    const int *lineface;   /* [nlineelem]    */
} tbs_mesh_host;

typedef struct tbs_plan_info {
    int nvar, nelem, nline, max_faces, precision;
    size_t device_bytes;          /* HBM held by the plan                      */
    double alg_bytes_per_sweep;   /* algorithmic bytes of one sweep (model)    */
    double ref_bytes_per_sweep;   /* the reference's own byte model,
                                     src/main.cxx:146-175, per sweep          */
    double thomas_bytes;          /* algorithmic bytes of one Thomas launch   */
    double residual_bytes;        /* algorithmic bytes of one residual launch */
} tbs_plan_info;

/* Footprint the plan would need, without allocating (admission control). */
int tbs_plan_footprint(const tbs_mesh_host *mesh, int precision, size_t *bytes);
/* TBS_ERR_UNSUPPORTED if one solution vector would exceed 4 GiB (about
 * 59 million elements at nvar 9): split such systems over several plans. */
int tbs_plan_create(tbs_ctx *ctx, const tbs_mesh_host *mesh, int precision,
                    tbs_plan **out);
int tbs_plan_destroy(tbs_plan *plan);
int tbs_plan_info_get(tbs_plan *plan, tbs_plan_info *info);
/* Factorise: the block-Thomas recurrence of lineLU (src/okl/lineLU_v2.okl:
 * 58-182) carried out in inverse form by one fused kernel that writes the
 * sweep-form factors W = D'^-1, E = D'^-1 A and Gamma = D'^-1 C directly.
 * Asynchronous on the context's stream: the Jacobian arrays must stay valid
 * until the stream has run it.  (The residual blocks are packed on a second
 * stream of the context beside the factor kernel; the context's stream waits
 * for them, so work queued after the call sees the whole factorisation.) */
int tbs_plan_factor(tbs_plan *plan, const double *jacD, const double *jacO1,
                    const double *jacO2);
/* Factorise n plans of one context (same block size and precision) in one
 * launch of the factor kernel over all their lines, longest first, with one
 * packing launch per plan beside it: jacD[i], jacO1[i], jacO2[i] are plan i's
 * device
 * arrays.  The same factors, bitwise, as n calls of tbs_plan_factor, in the
 * time of the longest line rather than the sum over plans.  No counterpart
 * in the reference (one system per run); for batches of systems. */
int tbs_plans_factor(tbs_ctx *ctx, tbs_plan *const *plans, int n, const double *const *jacD,
                     const double *const *jacO1, const double *const *jacO2);
/* Solve J U = -rhs by LIJ sweeps from the initial U (device, reference
 * layout, overwritten with the result).  Stops after max_sweeps or when
 * ||R_k|| <= rtol * ||R_0||.  hist (host, >= max_sweeps+1) gets ||R_k||_2,
 * k = 0..nsweeps.  R_final (device, may be NULL) gets rhs + J U.         */
int tbs_plan_solve(tbs_plan *plan, const double *rhs, double *U,
                   int max_sweeps, double rtol, double *hist, int *nsweeps,
                   double *R_final);

/* Autotuner (no counterpart in the reference, whose OKL kernels take their
 * tiling from JIT defines): times every compiled variant of the sweep's
 * kernels on this plan's own data (reps launches each, HIP events) and keeps
 * the fastest of each: the Thomas variant (factor windows prefetched per
 * wave, windows staged per LDS group, lines per workgroup) and the residual
 * workgroup shape (64-element groups per workgroup).  Every variant gives
 * bitwise the same results.  Resets the iterate: call before tbs_plan_begin /
 * tbs_plan_solve.  chosen = the Thomas variant; us[v] = mean microseconds of
 * Thomas variant v (the whole record: tbs_plan_tuning).  */
int tbs_plan_autotune(tbs_plan *plan, int reps, int *chosen, double *us, int cap);
int tbs_plan_set_variant(tbs_plan *plan, int variant);
/* Number of Thomas variants; pf[v] = 100 * (windows prefetched) + (ring
 * cycles per loop iteration) + 10000 * (lines per workgroup - 1). */
int tbs_plan_variants(tbs_plan *plan, int *pf, int cap);

/* The tuple the autotuner chose (or set by hand), with its timings. */
#define TBS_TUNE_THOMAS_MAX 8
#define TBS_TUNE_RESIDUAL_MAX 4
typedef struct {
    int thomas_variant;        /* index of the Thomas variant */
    int thomas_prefetch;       /* factor windows each wave keeps in flight */
    int thomas_lds_windows;    /* windows whose vectors a wave stages in LDS per group */
    int thomas_lines_per_wg;   /* lines per Thomas workgroup (two waves each) */
    int residual_variant;      /* index of the residual shape */
    int residual_waves_per_wg; /* 64-element groups (waves) per residual workgroup */
    int two_stream;            /* batches: eager sweeps as two halves on two streams */
    int n_thomas, n_residual;  /* variants compiled for this block size and precision */
    int residual_shapes[TBS_TUNE_RESIDUAL_MAX];  /* waves per workgroup of each residual variant */
    double thomas_us[TBS_TUNE_THOMAS_MAX];     /* last autotune: mean launch time per variant */
    double residual_us[TBS_TUNE_RESIDUAL_MAX]; /* (0 = not timed) */
    double pass_us[2];         /* batches: Thomas + residual pass, one stream / two streams */
} tbs_tuning;
int tbs_plan_tuning(tbs_plan *plan, tbs_tuning *out);
/* Set thomas_variant, residual_variant (two_stream is ignored for a plan). */
int tbs_plan_set_tuning(tbs_plan *plan, const tbs_tuning *t);

/* Low-level stepping, for timing and graph capture (all asynchronous). */
int tbs_plan_begin(tbs_plan *plan, const double *rhs, const double *U0);
int tbs_plan_launch_thomas(tbs_plan *plan);
int tbs_plan_launch_residual(tbs_plan *plan);   /* residual + norm of sweep */
/* One kernel of a sweep, for timing it alone: TBS_STEP_THOMAS (Thomas pass
 * and U update), TBS_STEP_RESIDUAL (R, per-group partial norms, next Thomas
 * right-hand side) or TBS_STEP_NORM (norm, history, convergence over the
 * partials).  A sweep is the three in this order.  With two events
 * (tbs_event_create, both or neither), the kernel is launched with
 * hipExtLaunchKernelGGL, which records them at the kernel's own start and
 * end, so tbs_event_elapsed_ms gives its execution time alone. */
enum { TBS_STEP_THOMAS = 0, TBS_STEP_RESIDUAL = 1, TBS_STEP_NORM = 2 };
int tbs_plan_launch_kernel(tbs_plan *plan, int step, void *ev_start, void *ev_stop);
int tbs_plan_launch_sweeps(tbs_plan *plan, int nsweeps, int use_graph);
int tbs_plan_end(tbs_plan *plan, double *U, double *R_final);
int tbs_plan_history(tbs_plan *plan, double *hist, int cap, int *count);
/* Diagnostics: device address and byte size of an internal array
 * (0 g, 1 d', 2 U, 3 rhs, 4 R, 5 W, 6 E windows, 7 Gamma windows,
 * 8 residual blocks); vectors are in line-major internal order. */
int tbs_plan_internal(tbs_plan *plan, int which, void **dptr, size_t *bytes);

/* ---------------------------------------------------------------------- */
/* Batched solves: many independent plans (uneven line lengths) advanced   */
/* by one launch per kernel per sweep, each with its own convergence test. */
/* No counterpart in the reference (it solves one system per run); for     */
/* callers that solve many small systems at once.                          */
/* ---------------------------------------------------------------------- */
int tbs_batch_create(tbs_ctx *ctx, tbs_plan *const *plans, int nplans,
                     tbs_batch **out);
int tbs_batch_destroy(tbs_batch *batch);
/* Replace the batch's systems (0 allowed: sweeps become no-ops).  The
 * descriptor tables are rewritten in place on the batch's stream, ordered
 * after the sweeps already enqueued; device tables grow when needed. */
int tbs_batch_set_plans(tbs_batch *batch, tbs_plan *const *plans, int nplans);
/* The plan autotuner's search over all systems of the batch at once, plus
 * the two-stream split of eager sweeps (timed both ways with the chosen
 * kernels). */
int tbs_batch_autotune(tbs_batch *batch, int reps, int *chosen, double *us, int cap);
int tbs_batch_tuning(tbs_batch *batch, tbs_tuning *out);
/* Set thomas_variant, residual_variant and two_stream. */
int tbs_batch_set_tuning(tbs_batch *batch, const tbs_tuning *t);
/* Enqueue nsweeps sweeps of every still-active system (asynchronous); the
 * plans must have been started with tbs_plan_begin.                        */
int tbs_batch_launch_sweeps(tbs_batch *batch, int nsweeps, int max_sweeps, double rtol);
/* Wait for the batch's stream, then copy every system's state:
 * states[4*s + 0] = residual norms recorded (sweeps + 1), [1] = active,
 * [2] = status (1 converged, 2 non-finite, 3 sweep cap, 4 diverging).    */
int tbs_batch_states(tbs_batch *batch, int *states);
/* rhs[i], U[i]: device pointers of plan i; hist: host [nplans][max_sweeps+1];
 * nsweeps: host [nplans].                                                  */
int tbs_batch_solve(tbs_batch *batch, const double *const *rhs,
                    double *const *U, int max_sweeps, double rtol,
                    int sweeps_per_check, double *hist, int *nsweeps);

/* ---------------------------------------------------------------------- */
/* Line-preconditioned GMRES (the paper's Krylov solver around the LIJ      */
/* smoother, kirby-mavriplis-2025 Table 6; the reference repository ships   */
/* no GMRES code).  Flexible right-preconditioned GMRES(restart) on a       */
/* factored plan solves J U = -rhs; the preconditioner is `sweeps` LIJ      */
/* sweeps from zero (1 = the block-tridiagonal line solve).  The Krylov     */
/* basis (2*restart+1 vectors of nelem*nvar doubles) is allocated at create. */
/* ---------------------------------------------------------------------- */
typedef struct tbs_gmres tbs_gmres;
int tbs_gmres_create(tbs_plan *plan, int restart, tbs_gmres **out);
int tbs_gmres_destroy(tbs_gmres *gmres);
/* rhs, U: device, element order; U is the initial guess and the result.
 * hist (host, hist_cap entries): ||R_0|| then the GMRES residual estimate
 * after each iteration; iters: iterations run.                             */
int tbs_gmres_solve(tbs_gmres *gmres, const double *rhs, double *U, int max_iters, double rtol,
                    int sweeps, double *hist, int hist_cap, int *iters);
/* [0] operator applications, [1] preconditioner sweeps, [2] restart cycles,
 * [3] true residual norm ||rhs + J U|| at the end of the last solve.       */
int tbs_gmres_stats(tbs_gmres *gmres, double *stats, int n);

/* ---------------------------------------------------------------------- */
/* Serving: continuous batching of solve requests on one GPU (no          */
/* counterpart in the reference; for request-driven callers).  Requests    */
/* wait in a FIFO; each tbs_server_step runs K sweeps of the active batch, */
/* admits queued requests whose HBM footprint fits the budget (generation, */
/* factorisation and first residual on a second stream, overlapping the   */
/* batch), retires converged ones (U copied to host) and lets admitted ones */
/* join.  A request's result does not depend on what it is batched with.  */
/* ---------------------------------------------------------------------- */
typedef struct tbs_server tbs_server;
typedef struct tbs_server_request {
    int NL, NK;          /* seeded structured system: NL lines of NK elements */
    uint64_t seed;
    double shift;        /* diagonal shift of the generator                  */
} tbs_server_request;
/* hbm_budget 0 = 80% of the device's free memory at creation.  The server's
 * HBM arena is allocated at the full budget here (in smaller pieces later
 * only if the device refuses that), so admissions allocate nothing; so are
 * its pinned result buffers, staging and batch tables, and one small
 * request runs through it before the call returns.  The server starts one
 * host thread of its own (admission set-up); the tbs_server_* calls on a
 * server must come from one thread at a time. */
int tbs_server_create(tbs_ctx *ctx, size_t hbm_budget, int precision, double rtol,
                      int max_sweeps, int sweeps_per_step, tbs_server **out);
int tbs_server_destroy(tbs_server *srv);
/* HBM bytes a request needs when admitted: its plan, rhs and U (held until
 * retirement) and the generated Jacobian (returned once the request's
 * set-up has completed).  Admission keeps the sum of what admitted
 * requests hold within the budget. */
int tbs_server_footprint(tbs_server *srv, const tbs_server_request *req, size_t *bytes);
/* TBS_ERR_NOMEM if the request alone exceeds the budget. */
int tbs_server_submit(tbs_server *srv, const tbs_server_request *req, long *id);
/* A caller's system: the mesh as the host arrays of the Mesh class
 * (src/Mesh.cxx:11-73, checked as tbs_plan_create checks it; block size =
 * the server context's), the Jacobian blocks jacD/jacO1/jacO2 and rhs in the
 * reference layout (Jacobian::jacD/jacO1/jacO2/rhs, src/Jacobian.cxx:11-74).
 * where = TBS_HOST: the values are copied during the call.  where =
 * TBS_DEVICE: device pointers, read on the device when the request is
 * admitted; they must stay valid until tbs_server_poll reports the request.
 * The result U is in element order, as for tbs_server_submit's requests. */
#define TBS_HOST   0
#define TBS_DEVICE 1
int tbs_server_submit_system(tbs_server *srv, const tbs_mesh_host *mesh, const double *jacD,
                             const double *jacO1, const double *jacO2, const double *rhs, int where,
                             long *id);
/* HBM bytes such a request holds when admitted (as tbs_server_footprint). */
int tbs_server_footprint_system(tbs_server *srv, const tbs_mesh_host *mesh, size_t *bytes);
/* One scheduling step: collect the running batch's states, let finished
 * systems leave and set-up-complete admissions join, launch the next
 * sweeps, start copying finished results out (a request is reported done
 * by a later step, once its copies have completed; at once when nothing
 * else is in flight), admit from the queue.  nactive counts running and
 * retiring requests. */
int tbs_server_step(tbs_server *srv, int *nactive, int *nqueued, int *nadmitting);
/* ids of requests finished since the last poll */
int tbs_server_poll(tbs_server *srv, long *ids, int cap, int *n);
/* U (host, element order), residual history, sweeps, status (1 converged,
 * 2 non-finite, 3 sweep cap, 4 diverging, -1 admission failed: the call
 * still returns TBS_OK and tbs_last_error() then gives the admission's
 * message) and times[3] = submit, admit, done (seconds on the server's
 * clock).                                                                   */
int tbs_server_result(tbs_server *srv, long id, double *U, size_t ucap, double *hist,
                      int hcap, int *nsweeps, int *status, double *times);
int tbs_server_release(tbs_server *srv, long id);
/* Where a finished request's time went (seconds on the server's clock; up to
 * n values): t[0] submitted, [1] admitted (its set-up enqueued), [2] set-up
 * seen complete by a step (it joins the batch), [3] convergence seen by a
 * step, [4] result copies enqueued, [5] copies seen complete (the result is
 * ready; times[2] of tbs_server_result). */
int tbs_server_timeline(tbs_server *srv, long id, double *t, int n);
int tbs_server_info(tbs_server *srv, size_t *budget, size_t *used, double *now);
/* Host time per engine phase, cumulative seconds: [0] waiting for the batch's
 * sweeps, [1] membership update, [2] sweep launch, [3] retirement, [4]
 * admission on the serving thread (slab and hand-off), [5] plan building and
 * [6] whole set-ups on the admission thread (which runs beside the steps);
 * counts: [0] steps, [1] admitted, [2] retired, [3] batch membership
 * updates. */
int tbs_server_stats(tbs_server *srv, double *seconds, int nseconds, long *counts, int ncounts);

/* ---------------------------------------------------------------------- */
/* Synthetic inputs (the reference reads files only; these generate the    */
/* seeded structured-grid workloads).  Counter-based SplitMix64: value =   */
/* a + (b-a)*u, u = (mix64(key + (i+1)*G) >> 11) * 2^-53.                  */
/* ---------------------------------------------------------------------- */
/* Structured NL x NK mesh (lines along the fast index) into caller arrays
 * sized from tbs_structured_sizes; 0-based like the reader's output.      */
int tbs_structured_sizes(int NL, int NK, int *nelem, int *nintface,
                         int *eftot, int *nline, int *nlineelem);
int tbs_structured_mesh(int NL, int NK, int *epoint, int *ef, int *fc,
                        int *linesize, int *linepoint, int *lines,
                        int *lineface);
/* Lines of uneven length (lens[nline]) stacked like the structured grid:
 * element (l,k) couples in-line to (l,k+-1) and across to (l+-1,k) where
 * that position exists.  Equal lengths reproduce tbs_structured_mesh.    */
int tbs_ragged_sizes(int nline, const int *lens, int *nelem, int *nintface,
                     int *eftot);
int tbs_ragged_mesh(int nline, const int *lens, int *epoint, int *ef, int *fc,
                    int *linesize, int *linepoint, int *lines, int *lineface);
/* Seeded shapes of a batch of nsys systems: nlines[s] ~ U{nlmin..nlmax},
 * lengths ~ U{lmin..lmax}, scaled to sum to `total` (if > 0), clamped to
 * [lmin, lmax], remainder settled from the last line backwards.  lens gets
 * all lengths system by system (capacity cap); *nlens their count.       */
int tbs_batch_shapes(uint64_t seed, int nsys, long total, int lmin, int lmax,
                     int nlmin, int nlmax, int *nlines, int *lens, int cap,
                     int *nlens);
/* Fill a Jacobian system: D = U[-1,1] + shift*I, Of1/Of2 = U[-0.1,0.1],
 * rhs = U[-1,1].  Host or device variant (device: on the ctx stream).    */
int tbs_generate_host(uint64_t seed, int nvar, int nelem, int nintface,
                      double shift, double *jacD, double *jacO1,
                      double *jacO2, double *rhs);
int tbs_generate_device(tbs_ctx *ctx, uint64_t seed, int nvar, int nelem,
                        int nintface, double shift, double *jacD,
                        double *jacO1, double *jacO2, double *rhs);

/* Binary files of the reference (src/Mesh.cxx:11-73, read_mesh_data.F90;
 * src/Jacobian.cxx:11-74, read_jacobian_data.F90).  Readers validate the
 * header against the file size.  sizes[6] = nvar, nelem, nintface, eftot,
 * nline, nlineelem.                                                       */
int tbs_read_mesh_header(const char *path, int sizes[6]);
int tbs_read_mesh(const char *path, int *epoint, int *ef, int *fc,
                  int *linesize, int *linepoint, int *lines, int *lineface);
int tbs_write_mesh(const char *path, const int sizes[6], const int *epoint,
                   const int *ef, const int *fc, const int *linesize,
                   const int *linepoint, const int *lines,
                   const int *lineface);
/* Host-side check of every index the drop-in kernels follow (epoint monotone
 * from 0 to eftot, ef below nintface, fc and lines inside [0, nelem),
 * linepoint/linesize contiguous up to nlineelem <= nelem, lineface of every
 * non-first line element an interior face joining it to its predecessor):
 * TBS_ERR_IO naming the first inconsistency.  The reference trusts its files
 * (src/Mesh.cxx:11-73); Mesh::fromFile calls this after reading. */
int tbs_mesh_validate(const tbs_mesh_host *mesh);
int tbs_read_jacobian_header(const char *path, int sizes[3]);
int tbs_read_jacobian(const char *path, double *jacD, double *jacO1,
                      double *jacO2, double *rhs, double *U0);
int tbs_write_jacobian(const char *path, int nvar, int nelem, int nintface,
                       const double *jacD, const double *jacO1,
                       const double *jacO2, const double *rhs,
                       const double *U0);

#ifdef __cplusplus
}
#endif
#endif /* TRIBLOCK_HIP_H */
