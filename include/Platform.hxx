/*
 * Platform.hxx -- device selection and memory for the HIP build.
 * Mirrors the reference's Platform (include/Platform.hxx:25-92): same mode
 * numbers and constructor shape, minus the MPI communicator (one process per
 * GPU, no collectives) and the OCCA JIT (kernels are compiled ahead of time
 * into libtriblock_hip.so).  Only HIP_MODE runs here; the Serial/OpenMP CPU
 * modes of the reference are not part of this product.
 */
#ifndef TBS_PLATFORM_HXX
#define TBS_PLATFORM_HXX

#include <cstddef>
#include <stdexcept>
#include <string>

#include "triblock_hip.h"

#define SERIAL_MODE 0
#define HIP_MODE    1
#define CUDA_MODE   2
#define OPENCL_MODE 3
#define OPENMP_MODE 4
#define DPCPP_MODE  5
#define METAL_MODE  6

/* Throws std::runtime_error carrying tbs_last_error() when rc != TBS_OK. */
void tbs_check(int rc, const char *what);

/* Pinned host memory returned by Platform::hostMalloc; frees it when it goes
 * out of scope (the reference hands back an occa::memory that owns it). */
class HostMemory {
  public:
    HostMemory() = default;
    ~HostMemory() { reset(); }
    HostMemory(const HostMemory &) = delete;
    HostMemory &operator=(const HostMemory &) = delete;
    HostMemory(HostMemory &&o) noexcept : ctx_(o.ctx_), ptr_(o.ptr_), bytes_(o.bytes_) { o.ptr_ = nullptr; }
    void *ptr() const { return ptr_; }
    size_t size() const { return bytes_; }
    void reset() {
        if (ptr_) tbs_host_free(ctx_, ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }

  private:
    friend class Platform;
    tbs_ctx *ctx_ = nullptr;
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
};

class Platform {
  public:
    int mode;
    int device_id;
    int nvar;
    tbs_ctx *ctx = nullptr;

    Platform(int thread_model, int device_id = 0, int nvar = 9);
    ~Platform();
    Platform(const Platform &) = delete;
    Platform &operator=(const Platform &) = delete;

    /* a device buffer as the source of an allocation (the reference passes
     * an occa::memory) */
    struct Device {
        const void *ptr;
    };

    /* device allocation of `entries` values of T, zero-filled, or copied from
     * host memory src (the reference's malloc<T>(entries, src),
     * include/Platform.hxx:62-68) */
    template <class T>
    T *malloc(size_t entries, const void *src = nullptr) {
        T *p = alloc<T>(entries);
        if (!entries) return p;
        if (src)
            tbs_check(tbs_memcpy_h2d(ctx, p, src, entries * sizeof(T)), "Platform::malloc");
        else
            tbs_check(tbs_memset(ctx, p, 0, entries * sizeof(T)), "Platform::malloc");
        return p;
    }
    /* device allocation copied from another device buffer (the reference's
     * malloc<T>(entries, occa::memory src), include/Platform.hxx:70-75) */
    template <class T>
    T *malloc(size_t entries, Device src) {
        T *p = alloc<T>(entries);
        if (entries) tbs_check(tbs_memcpy_d2d(ctx, p, src.ptr, entries * sizeof(T)), "Platform::malloc");
        return p;
    }
    /* pinned host allocation of `bytes`, copied from src when it is not null;
     * h_mem owns it (the reference's hostMalloc, include/Platform.hxx:82-88) */
    void *hostMalloc(size_t bytes, const void *src, HostMemory &h_mem);
    void free(void *p, size_t bytes = 0);
    void finish();                         /* occa::device::finish analogue */
    size_t memoryAllocated() const { return allocated_; }
    void *stream() const { return tbs_ctx_stream(ctx); }

  private:
    template <class T>
    T *alloc(size_t entries) {
        void *p = nullptr;
        tbs_check(tbs_malloc(ctx, entries * sizeof(T), &p), "Platform::malloc");
        allocated_ += entries * sizeof(T);
        return static_cast<T *>(p);
    }
    size_t allocated_ = 0;
};

#endif
