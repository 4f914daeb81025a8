// stream_ceiling.hip -- the achievable HBM bandwidth of the device, measured
// with the two simplest streams: a read-only pass and a copy.  The sweep
// kernels are HBM-bound, so their fraction of this ceiling (next to the
// fraction of the published peak) says how much of the gap to peak is the
// kernels' and how much the memory system's.  Each stream is timed in several
// launch shapes (16-byte loads in flight per thread, workgroups per CU,
// streaming or default cache policy) and the fastest is reported: the
// ceiling is what the best of these simple kernels reaches.
#include "tbs_common.hpp"

#include <algorithm>

namespace {

typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ u32x4v ld(const u32x4v *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ __forceinline__ void st(u32x4v v, u32x4v *p) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// XOR of every 16-byte word, one 4-byte result per thread (the results are a
// negligible fraction of the bytes read, and keep the loads live).
template <int U, bool NT>
__global__ __launch_bounds__(256) void stream_read_kernel(const u32x4v *__restrict__ src, size_t n16,
                                                          unsigned *__restrict__ out) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    u32x4v acc = {0u, 0u, 0u, 0u};
    size_t i = tid;
    for (; i + (U - 1) * nth < n16; i += U * nth) {
        u32x4v v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) v[k] = ld<NT>(src + i + k * nth);
#pragma unroll
        for (int k = 0; k < U; ++k) acc ^= v[k];
    }
    for (; i < n16; i += nth) acc ^= ld<NT>(src + i);
    out[tid] = acc.x ^ acc.y ^ acc.z ^ acc.w;
}

// dst = src in 16-byte words.
template <int U, bool NT>
__global__ __launch_bounds__(256) void stream_copy_kernel(const u32x4v *__restrict__ src, u32x4v *__restrict__ dst,
                                                          size_t n16) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    size_t i = tid;
    for (; i + (U - 1) * nth < n16; i += U * nth) {
        u32x4v v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) v[k] = ld<NT>(src + i + k * nth);
#pragma unroll
        for (int k = 0; k < U; ++k) st<NT>(v[k], dst + i + k * nth);
    }
    for (; i < n16; i += nth) st<NT>(ld<NT>(src + i), dst + i);
}

// The same two streams with each workgroup on contiguous runs of kRun
// 16-byte words (U consecutive 4 KiB pieces in flight per workgroup), the
// access pattern of the sweep kernels' block slots.
constexpr size_t kRun = 16384;   // 256 KiB per run
template <int U, bool NT>
__global__ __launch_bounds__(256) void stream_read_blocked_kernel(const u32x4v *__restrict__ src, size_t n16,
                                                                  unsigned *__restrict__ out) {
    u32x4v acc = {0u, 0u, 0u, 0u};
    for (size_t base = (size_t)blockIdx.x * kRun; base < n16; base += (size_t)gridDim.x * kRun) {
        const size_t end = base + kRun < n16 ? base + kRun : n16;
        for (size_t j = base + threadIdx.x; j < end; j += U * 256) {
            u32x4v v[U];
#pragma unroll
            for (int k = 0; k < U; ++k) v[k] = j + k * 256 < end ? ld<NT>(src + j + k * 256) : u32x4v{0u, 0u, 0u, 0u};
#pragma unroll
            for (int k = 0; k < U; ++k) acc ^= v[k];
        }
    }
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc.x ^ acc.y ^ acc.z ^ acc.w;
}

template <int U, bool NT>
__global__ __launch_bounds__(256) void stream_copy_blocked_kernel(const u32x4v *__restrict__ src,
                                                                  u32x4v *__restrict__ dst, size_t n16) {
    for (size_t base = (size_t)blockIdx.x * kRun; base < n16; base += (size_t)gridDim.x * kRun) {
        const size_t end = base + kRun < n16 ? base + kRun : n16;
        for (size_t j = base + threadIdx.x; j < end; j += U * 256) {
            u32x4v v[U];
#pragma unroll
            for (int k = 0; k < U; ++k) v[k] = j + k * 256 < end ? ld<NT>(src + j + k * 256) : u32x4v{0u, 0u, 0u, 0u};
#pragma unroll
            for (int k = 0; k < U; ++k)
                if (j + k * 256 < end) st<NT>(v[k], dst + j + k * 256);
        }
    }
}

__global__ void stream_fill_kernel(u32x4v *__restrict__ p, size_t n16) {
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += nth) {
        const unsigned s = (unsigned)i * 2654435761u;
        p[i] = u32x4v{s, s ^ 1u, s ^ 2u, s ^ 3u};
    }
}

template <int U, bool NT>
void launch_stream(int kind, unsigned grid, hipStream_t s, const void *a, void *b, void *o, size_t n16) {
    if (kind == 0)
        hipLaunchKernelGGL((stream_read_kernel<U, NT>), dim3(grid), dim3(256), 0, s, (const u32x4v *)a, n16,
                           (unsigned *)o);
    else
        hipLaunchKernelGGL((stream_copy_kernel<U, NT>), dim3(grid), dim3(256), 0, s, (const u32x4v *)a,
                           (u32x4v *)b, n16);
}

template <int U, bool NT>
void launch_blocked(int kind, unsigned grid, hipStream_t s, const void *a, void *b, void *o, size_t n16) {
    if (kind == 0)
        hipLaunchKernelGGL((stream_read_blocked_kernel<U, NT>), dim3(grid), dim3(256), 0, s, (const u32x4v *)a,
                           n16, (unsigned *)o);
    else
        hipLaunchKernelGGL((stream_copy_blocked_kernel<U, NT>), dim3(grid), dim3(256), 0, s, (const u32x4v *)a,
                           (u32x4v *)b, n16);
}

using LaunchFn = void (*)(int, unsigned, hipStream_t, const void *, void *, void *, size_t);
constexpr LaunchFn kShapes[] = {launch_stream<4, true>,  launch_stream<4, false>,  launch_stream<8, true>,
                                launch_stream<8, false>, launch_blocked<4, true>,  launch_blocked<4, false>,
                                launch_blocked<8, true>, launch_blocked<8, false>};
constexpr int kWgPerCu[] = {4, 8, 16, 32};
constexpr int kMaxWgPerCu = 32;

}  // namespace

extern "C" {

int tbs_stream_ceiling(tbs_ctx *ctx, size_t bytes, int reps, double *gbps) {
    TBS_TRY(tbs::use_ctx(ctx));
    TBS_REQUIRE(gbps && reps >= 1 && bytes >= ((size_t)1 << 20), TBS_ERR_ARG,
                "need gbps[2], reps >= 1 and at least 1 MiB");
    const size_t n16 = bytes / 16;
    hipDeviceProp_t prop;
    TBS_HIP(hipGetDeviceProperties(&prop, ctx->device));
    const int ncu = std::max(1, prop.multiProcessorCount);
    const size_t max_threads = (size_t)ncu * kMaxWgPerCu * 256;
    hipStream_t st = ctx->stream;
    void *a = nullptr, *b = nullptr, *o = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TBS_OK;
    auto fail = [&](hipError_t e, const char *what) { rc = tbs::hip_fail(e, what, __FILE__, __LINE__); };
    hipError_t e;
    if ((e = hipMalloc(&a, n16 * 16)) != hipSuccess) fail(e, "hipMalloc src");
    else if ((e = hipMalloc(&b, n16 * 16)) != hipSuccess) fail(e, "hipMalloc dst");
    else if ((e = hipMalloc(&o, max_threads * sizeof(unsigned))) != hipSuccess) fail(e, "hipMalloc out");
    else if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) fail(e, "event");
    gbps[0] = gbps[1] = 0.0;
    if (rc == TBS_OK) {
        hipLaunchKernelGGL(stream_fill_kernel, dim3((unsigned)ncu * 8), dim3(256), 0, st, (u32x4v *)a, n16);
        for (int kind = 0; kind < 2 && rc == TBS_OK; ++kind)
            for (LaunchFn fn : kShapes)
                for (int wg : kWgPerCu) {
                    if (rc != TBS_OK) break;
                    const unsigned grid = (unsigned)(ncu * wg);   // grid * 256 <= max_threads
                    fn(kind, grid, st, a, b, o, n16);            // warm-up
                    hipEventRecord(e0, st);
                    for (int r = 0; r < reps; ++r) fn(kind, grid, st, a, b, o, n16);
                    hipEventRecord(e1, st);
                    if ((e = hipGetLastError()) != hipSuccess) { fail(e, "stream kernel launch"); break; }
                    if ((e = hipEventSynchronize(e1)) != hipSuccess) { fail(e, "hipEventSynchronize"); break; }
                    float ms = 0.f;
                    hipEventElapsedTime(&ms, e0, e1);
                    const double moved = (double)n16 * 16.0 * (kind == 0 ? 1.0 : 2.0);
                    gbps[kind] = std::max(gbps[kind], moved * reps / (ms * 1e-3) / 1e9);
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
