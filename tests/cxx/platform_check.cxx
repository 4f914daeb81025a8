// platform_check.cxx -- exercises the Platform memory API (include/Platform.hxx)
// on the GPU: malloc<T>(entries, host src), malloc<T>(entries, Device{src}),
// zero-filled malloc<T>(entries) and hostMalloc(bytes, src, h_mem).  Prints
// "platform ok" and exits 0 when every copy round-trips bit for bit.
#include <cstdio>
#include <cstring>
#include <vector>

#include "Platform.hxx"

int main() {
    try {
        Platform gpu(HIP_MODE, 0, 9);
        const size_t n = 100003;
        std::vector<double> h(n);
        for (size_t k = 0; k < n; ++k) h[k] = 0.5 * (double)k - 17.25;
        double *d1 = gpu.malloc<double>(n, h.data());                  // from host memory
        double *d2 = gpu.malloc<double>(n, Platform::Device{d1});      // from device memory
        int *dz = gpu.malloc<int>(n);                                  // zero-filled
        HostMemory pin;
        double *p = static_cast<double *>(gpu.hostMalloc(n * sizeof(double), nullptr, pin));
        if (!p || pin.ptr() != p || pin.size() != n * sizeof(double)) {
            std::printf("hostMalloc returned no buffer\n");
            return 1;
        }
        tbs_check(tbs_memcpy_d2h(gpu.ctx, p, d2, n * sizeof(double)), "d2h");
        if (std::memcmp(p, h.data(), n * sizeof(double)) != 0) {
            std::printf("device-to-device malloc copy differs\n");
            return 1;
        }
        std::vector<int> z(n, 1);
        tbs_check(tbs_memcpy_d2h(gpu.ctx, z.data(), dz, n * sizeof(int)), "d2h");
        for (int v : z)
            if (v != 0) {
                std::printf("malloc without a source is not zero-filled\n");
                return 1;
            }
        HostMemory pin2;
        void *q = gpu.hostMalloc(n * sizeof(double), h.data(), pin2);   // copied from src
        if (std::memcmp(q, h.data(), n * sizeof(double)) != 0) {
            std::printf("hostMalloc did not copy its source\n");
            return 1;
        }
        if (gpu.memoryAllocated() != 2 * n * sizeof(double) + n * sizeof(int)) {
            std::printf("memoryAllocated %zu\n", gpu.memoryAllocated());
            return 1;
        }
        gpu.free(d1, n * sizeof(double));
        gpu.free(d2, n * sizeof(double));
        gpu.free(dz, n * sizeof(int));
        pin.reset();
        if (pin.ptr() != nullptr) return 1;
        std::printf("platform ok\n");
        return 0;
    } catch (const std::exception &e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
}
