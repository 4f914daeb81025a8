// Platform.cxx -- HIP device context behind the reference's Platform API
// (src/Platform.cxx:11-78 selects an OCCA mode and device; here the only
// device mode is HIP on gfx950).
#include "Platform.hxx"

#include <cstdio>
#include <cstring>

void tbs_check(int rc, const char *what) {
    if (rc != TBS_OK)
        throw std::runtime_error(std::string(what) + ": " + tbs_last_error() + " (status " +
                                 std::to_string(rc) + ")");
}

Platform::Platform(int thread_model, int dev, int nv) : mode(thread_model), device_id(dev), nvar(nv) {
    if (thread_model != HIP_MODE)
        throw std::runtime_error("compute_mode " + std::to_string(thread_model) +
                                 " is not built: this library runs HIP_MODE (1) on MI355X only");
    tbs_check(tbs_ctx_create(device_id, nvar, &ctx), "tbs_ctx_create");
}

Platform::~Platform() {
    if (ctx) tbs_ctx_destroy(ctx);
}

void Platform::free(void *p, size_t bytes) {
    if (p) tbs_free(ctx, p);
    allocated_ -= (bytes <= allocated_) ? bytes : allocated_;
}

void Platform::finish() { tbs_check(tbs_ctx_synchronize(ctx), "Platform::finish"); }

void *Platform::hostMalloc(size_t bytes, const void *src, HostMemory &h_mem) {
    h_mem.reset();
    void *p = nullptr;
    tbs_check(tbs_host_malloc(ctx, bytes, &p), "Platform::hostMalloc");
    h_mem.ctx_ = ctx;
    h_mem.ptr_ = p;
    h_mem.bytes_ = bytes;
    if (p && src) std::memcpy(p, src, bytes);
    return p;
}
