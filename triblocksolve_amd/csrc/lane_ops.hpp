// lane_ops.hpp -- wave64 cross-lane helpers used by the line kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tbs {

// Value of v in lane `lane` (uniform result).  lane must be wave-uniform.
__device__ __forceinline__ double rdlane(double v, int lane) {
    const uint64_t b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

__device__ __forceinline__ float rdlane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Wave-wide sum (all 64 lanes), fixed butterfly order.
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace tbs
