// stream_ceiling.hip -- the achievable HBM bandwidth of the device, measured
// with the two simplest streams: a read-only pass and a copy.  The sweep
// kernels are HBM-bound, so their fraction of this ceiling (next to the
// fraction of the published peak) says how much of the gap to peak is the
// kernels' and how much the memory system's.
#include "tbs_common.hpp"

#include <algorithm>

namespace {

typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));
constexpr int kUnroll = 4;   // 16-byte loads in flight per thread

// XOR of every 16-byte word, one 4-byte result per thread (the results are
// 1/(64*kUnroll*iterations) of the bytes read, and keep the loads live).
__global__ __launch_bounds__(256) void stream_read_kernel(const u32x4v *__restrict__ src, size_t n16,
                                                          unsigned *__restrict__ out) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    u32x4v acc = {0u, 0u, 0u, 0u};
    size_t i = tid;
    for (; i + (kUnroll - 1) * nth < n16; i += kUnroll * nth) {
        u32x4v v[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) v[k] = __builtin_nontemporal_load(src + i + k * nth);
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) acc ^= v[k];
    }
    for (; i < n16; i += nth) acc ^= __builtin_nontemporal_load(src + i);
    out[tid] = acc.x ^ acc.y ^ acc.z ^ acc.w;
}

// dst = src in 16-byte words.
__global__ __launch_bounds__(256) void stream_copy_kernel(const u32x4v *__restrict__ src, u32x4v *__restrict__ dst,
                                                          size_t n16) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    size_t i = tid;
    for (; i + (kUnroll - 1) * nth < n16; i += kUnroll * nth) {
        u32x4v v[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) v[k] = __builtin_nontemporal_load(src + i + k * nth);
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) __builtin_nontemporal_store(v[k], dst + i + k * nth);
    }
    for (; i < n16; i += nth) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

__global__ void stream_fill_kernel(u32x4v *__restrict__ p, size_t n16) {
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += nth) {
        const unsigned s = (unsigned)i * 2654435761u;
        p[i] = u32x4v{s, s ^ 1u, s ^ 2u, s ^ 3u};
    }
}

}  // namespace

extern "C" {

int tbs_stream_ceiling(tbs_ctx *ctx, size_t bytes, int reps, double *gbps) {
    TBS_TRY(tbs::use_ctx(ctx));
    TBS_REQUIRE(gbps && reps >= 1 && bytes >= ((size_t)1 << 20), TBS_ERR_ARG,
                "need gbps[2], reps >= 1 and at least 1 MiB");
    const size_t n16 = bytes / 16;
    hipDeviceProp_t prop;
    TBS_HIP(hipGetDeviceProperties(&prop, ctx->device));
    // 8 workgroups of 256 threads per CU: every SIMD holds several waves
    const unsigned grid = (unsigned)std::max(1, prop.multiProcessorCount * 8);
    const size_t nth = (size_t)grid * 256;
    hipStream_t st = ctx->stream;
    void *a = nullptr, *b = nullptr, *o = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TBS_OK;
    auto fail = [&](hipError_t e, const char *what) { rc = tbs::hip_fail(e, what, __FILE__, __LINE__); };
    hipError_t e;
    if ((e = hipMalloc(&a, n16 * 16)) != hipSuccess) fail(e, "hipMalloc src");
    else if ((e = hipMalloc(&b, n16 * 16)) != hipSuccess) fail(e, "hipMalloc dst");
    else if ((e = hipMalloc(&o, nth * sizeof(unsigned))) != hipSuccess) fail(e, "hipMalloc out");
    else if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) fail(e, "event");
    if (rc == TBS_OK) {
        hipLaunchKernelGGL(stream_fill_kernel, dim3(grid), dim3(256), 0, st, (u32x4v *)a, n16);
        for (int kind = 0; kind < 2 && rc == TBS_OK; ++kind) {
            auto launch = [&]() {
                if (kind == 0)
                    hipLaunchKernelGGL(stream_read_kernel, dim3(grid), dim3(256), 0, st, (const u32x4v *)a, n16,
                                       (unsigned *)o);
                else
                    hipLaunchKernelGGL(stream_copy_kernel, dim3(grid), dim3(256), 0, st, (const u32x4v *)a,
                                       (u32x4v *)b, n16);
            };
            launch();                                   // warm-up
            hipEventRecord(e0, st);
            for (int r = 0; r < reps; ++r) launch();
            hipEventRecord(e1, st);
            if ((e = hipGetLastError()) != hipSuccess) { fail(e, "stream kernel launch"); break; }
            if ((e = hipEventSynchronize(e1)) != hipSuccess) { fail(e, "hipEventSynchronize"); break; }
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            const double moved = (double)n16 * 16.0 * (kind == 0 ? 1.0 : 2.0);
            gbps[kind] = moved * reps / (ms * 1e-3) / 1e9;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipStreamSynchronize(st);
    if (a) hipFree(a);
    if (b) hipFree(b);
    if (o) hipFree(o);
    return rc;
}

}  // extern "C"
