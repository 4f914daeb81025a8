// ctx.hip -- device context (stream, device id, block size) and the shared
// error / memory helpers.  Plays the role of the reference's Platform class
// (include/Platform.hxx:25-92, src/Platform.cxx:11-78): it selects the device
// and owns the launch queue; there is no JIT and no MPI.
#include "tbs_common.hpp"

#include <algorithm>
#include <cstring>

namespace tbs {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    return (int)e;
}

int use_ctx(tbs_ctx *ctx) {
    TBS_REQUIRE(ctx != nullptr, TBS_ERR_ARG, "null context");
    TBS_HIP(hipSetDevice(ctx->device));
    return TBS_OK;
}

int dev_alloc(tbs_ctx *ctx, size_t bytes, void **ptr, bool *from_region) {
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    if (ctx->region) {
        const size_t off = (ctx->region_used + kRegionAlign - 1) / kRegionAlign * kRegionAlign;
        TBS_REQUIRE(off + bytes <= ctx->region_size, TBS_ERR_NOMEM,
                    "region exhausted: %zu + %zu bytes > %zu", off, bytes, ctx->region_size);
        *ptr = ctx->region + off;
        ctx->region_used = off + bytes;
        if (from_region) *from_region = true;
        return TBS_OK;
    }
    if (from_region) *from_region = false;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e == hipErrorOutOfMemory) {
        hipGetLastError();
        *ptr = nullptr;
        set_error("device allocation of %zu bytes refused", bytes);
        return TBS_ERR_NOMEM;
    }
    TBS_HIP(e);
    return TBS_OK;
}

int aux_stream(tbs_ctx *ctx, hipStream_t *out) {
    if (!ctx->aux) {
        TBS_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
        TBS_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
        TBS_HIP(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    }
    *out = ctx->aux;
    return TBS_OK;
}

int stage_acquire(tbs_ctx *ctx, size_t bytes, char **out) {
    const int i = ctx->stage_cur;
    if (ctx->stage_busy[i]) TBS_HIP(hipEventSynchronize(ctx->stage_ev[i]));
    ctx->stage_busy[i] = false;
    if (bytes > ctx->stage_cap[i]) {
        const size_t cap = std::max(bytes, ctx->stage_cap[i] + ctx->stage_cap[i] / 2);
        if (ctx->stage[i]) TBS_HIP(hipHostFree(ctx->stage[i]));
        ctx->stage[i] = nullptr;
        ctx->stage_cap[i] = 0;
        TBS_HIP(hipHostMalloc((void **)&ctx->stage[i], cap, hipHostMallocDefault));
        ctx->stage_cap[i] = cap;
    }
    if (!ctx->stage_ev[i]) TBS_HIP(hipEventCreateWithFlags(&ctx->stage_ev[i], hipEventDisableTiming));
    *out = ctx->stage[i];
    return TBS_OK;
}

int dev_stage(tbs_ctx *ctx, size_t bytes) {
    if (bytes <= ctx->dstage_cap) return TBS_OK;
    if (ctx->dstage) {
        TBS_HIP(hipStreamSynchronize(ctx->stream));   // earlier users of the old buffer
        TBS_HIP(hipFree(ctx->dstage));
    }
    ctx->dstage = nullptr;
    ctx->dstage_cap = 0;
    const size_t cap = std::max<size_t>(bytes, 1 << 16);
    TBS_HIP(hipMalloc((void **)&ctx->dstage, cap));
    ctx->dstage_cap = cap;
    return TBS_OK;
}

int stage_release(tbs_ctx *ctx) {
    const int i = ctx->stage_cur;
    TBS_HIP(hipEventRecord(ctx->stage_ev[i], ctx->stream));
    ctx->stage_busy[i] = true;
    ctx->stage_cur = (i + 1) % tbs_ctx::kStages;
    return TBS_OK;
}

}  // namespace tbs

extern "C" {

const char *tbs_last_error(void) { return tbs::g_last_error.c_str(); }
const char *tbs_version(void) { return "triblocksolve-mi355x 0.1 (gfx950)"; }

int tbs_ctx_create(int device_id, int nvar, tbs_ctx **out) {
    TBS_REQUIRE(out != nullptr, TBS_ERR_ARG, "null output pointer");
    *out = nullptr;
    int supported[64];
    int ns = tbs_supported_nvar(supported, 64);
    bool ok = false;
    for (int i = 0; i < ns; ++i) ok |= (supported[i] == nvar);
    TBS_REQUIRE(ok, TBS_ERR_UNSUPPORTED, "block size nvar=%d is not built into this library", nvar);
    int ndev = 0;
    TBS_HIP(hipGetDeviceCount(&ndev));
    TBS_REQUIRE(device_id >= 0 && device_id < ndev, TBS_ERR_ARG,
                "device_id %d out of range (%d devices)", device_id, ndev);
    TBS_HIP(hipSetDevice(device_id));
    tbs_ctx *c = new tbs_ctx();
    c->device = device_id;
    c->nvar = nvar;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return tbs::hip_fail(e, "hipStreamCreate", __FILE__, __LINE__); }
    c->own_stream = true;
    e = hipHostMalloc((void **)&c->red_host, sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return tbs::hip_fail(e, "hipHostMalloc", __FILE__, __LINE__);
    }
    *out = c;
    return TBS_OK;
}

int tbs_ctx_destroy(tbs_ctx *ctx) {
    if (!ctx) return TBS_OK;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->aux) {
        hipStreamSynchronize(ctx->aux);
        hipStreamDestroy(ctx->aux);
    }
    if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
    for (int i = 0; i < tbs_ctx::kStages; ++i) {
        if (ctx->stage_ev[i]) hipEventDestroy(ctx->stage_ev[i]);
        if (ctx->stage[i]) hipHostFree(ctx->stage[i]);
    }
    if (ctx->dstage) hipFree(ctx->dstage);
    if (ctx->red_buf) hipFree(ctx->red_buf);
    if (ctx->red_host) hipHostFree(ctx->red_host);
    if (ctx->own_stream && ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return TBS_OK;
}

int tbs_ctx_set_stream(tbs_ctx *ctx, void *hip_stream) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    if (ctx->own_stream && ctx->stream) {
        TBS_HIP(hipStreamSynchronize(ctx->stream));
        TBS_HIP(hipStreamDestroy(ctx->stream));
    }
    ctx->stream = (hipStream_t)hip_stream;
    ctx->own_stream = false;
    return TBS_OK;
}

void *tbs_ctx_stream(tbs_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int tbs_ctx_synchronize(tbs_ctx *ctx) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_HIP(hipStreamSynchronize(ctx->stream));
    return TBS_OK;
}

int tbs_device_synchronize(tbs_ctx *ctx) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_HIP(hipDeviceSynchronize());
    return TBS_OK;
}

int tbs_ctx_nvar(tbs_ctx *ctx) { return ctx ? ctx->nvar : TBS_ERR_ARG; }

int tbs_malloc(tbs_ctx *ctx, size_t bytes, void **dptr) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(dptr != nullptr, TBS_ERR_ARG, "null output pointer");
    *dptr = nullptr;
    if (bytes == 0) return TBS_OK;
    hipError_t e = hipMalloc(dptr, bytes);
    if (e == hipErrorOutOfMemory) {
        hipGetLastError();
        tbs::set_error("device allocation of %zu bytes refused", bytes);
        return TBS_ERR_NOMEM;
    }
    TBS_HIP(e);
    return TBS_OK;
}

int tbs_free(tbs_ctx *ctx, void *dptr) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    if (dptr) TBS_HIP(hipFree(dptr));
    return TBS_OK;
}

int tbs_memcpy_h2d(tbs_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    if (bytes == 0) return TBS_OK;
    TBS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    TBS_HIP(hipStreamSynchronize(ctx->stream));
    return TBS_OK;
}

int tbs_memcpy_d2h(tbs_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    if (bytes == 0) return TBS_OK;
    TBS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TBS_HIP(hipStreamSynchronize(ctx->stream));
    return TBS_OK;
}

int tbs_memcpy_d2d(tbs_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    if (bytes == 0) return TBS_OK;
    TBS_REQUIRE(dst && src, TBS_ERR_ARG, "null pointer");
    TBS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    TBS_HIP(hipStreamSynchronize(ctx->stream));
    return TBS_OK;
}

int tbs_host_malloc(tbs_ctx *ctx, size_t bytes, void **hptr) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(hptr != nullptr, TBS_ERR_ARG, "null output pointer");
    *hptr = nullptr;
    if (bytes == 0) return TBS_OK;
    hipError_t e = hipHostMalloc(hptr, bytes, hipHostMallocDefault);
    if (e == hipErrorOutOfMemory) {
        hipGetLastError();
        tbs::set_error("pinned host allocation of %zu bytes refused", bytes);
        return TBS_ERR_NOMEM;
    }
    TBS_HIP(e);
    return TBS_OK;
}

int tbs_host_free(tbs_ctx *ctx, void *hptr) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    if (hptr) TBS_HIP(hipHostFree(hptr));
    return TBS_OK;
}

int tbs_memset(tbs_ctx *ctx, void *dst, int value, size_t bytes) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    if (bytes == 0) return TBS_OK;
    TBS_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return TBS_OK;
}

int tbs_event_create(tbs_ctx *ctx, void **ev) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(ev != nullptr, TBS_ERR_ARG, "null output pointer");
    hipEvent_t e;
    TBS_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    *ev = (void *)e;
    return TBS_OK;
}

int tbs_event_destroy(tbs_ctx *ctx, void *ev) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    if (ev) TBS_HIP(hipEventDestroy((hipEvent_t)ev));
    return TBS_OK;
}

int tbs_event_record(tbs_ctx *ctx, void *ev) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(ev != nullptr, TBS_ERR_ARG, "null event");
    TBS_HIP(hipEventRecord((hipEvent_t)ev, ctx->stream));
    return TBS_OK;
}

int tbs_event_elapsed_ms(void *ev_start, void *ev_stop, float *ms) {
    TBS_REQUIRE(ev_start && ev_stop && ms, TBS_ERR_ARG, "null argument");
    TBS_HIP(hipEventSynchronize((hipEvent_t)ev_stop));
    TBS_HIP(hipEventElapsedTime(ms, (hipEvent_t)ev_start, (hipEvent_t)ev_stop));
    return TBS_OK;
}

}  // extern "C"
