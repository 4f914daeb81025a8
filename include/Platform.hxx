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

    /* device allocation of `entries` values of T (zero-filled) */
    template <class T>
    T *malloc(size_t entries) {
        void *p = nullptr;
        tbs_check(tbs_malloc(ctx, entries * sizeof(T), &p), "Platform::malloc");
        if (entries) tbs_check(tbs_memset(ctx, p, 0, entries * sizeof(T)), "Platform::malloc");
        allocated_ += entries * sizeof(T);
        return static_cast<T *>(p);
    }
    void free(void *p, size_t bytes = 0);
    void finish();                         /* occa::device::finish analogue */
    size_t memoryAllocated() const { return allocated_; }
    void *stream() const { return tbs_ctx_stream(ctx); }

  private:
    size_t allocated_ = 0;
};

#endif
