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

// out[j] = v of lane j, for j < N.  N <= 16: a DPP row broadcast per j
// (row_newbcast: each row of 16 lanes receives its own lane j, a VGPR result,
// one instruction per double); N > 16: v_readlane pairs (SGPR results).  In
// both cases lanes 0..15 see lane j's value, bit for bit.
template <int N, int J = 0>
__device__ __forceinline__ void bcast_rows(double v, double (&out)[N]) {
    if constexpr (J < N) {
        if constexpr (N <= 16)
            out[J] = __builtin_amdgcn_mov_dpp(v, 0x150 + J, 0xF, 0xF, true);
        else
            out[J] = rdlane(v, J);
        bcast_rows<N, J + 1>(v, out);
    }
}

// Wave-wide sum (all 64 lanes), fixed butterfly order.
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace tbs

// ---------------------------------------------------------------------------
// Buffer-resource (V#) loads and stores.  The descriptor holds a wave-uniform
// base and a byte range; a lane whose offset falls outside the range reads 0
// and its store is dropped, which is how the sweep kernels handle ragged line
// ends, padded lanes and padded groups without branches.  Ranges are < 4 GiB.
namespace tbs {

using rsrc_t = __amdgpu_buffer_rsrc_t;

// Word 3 of the descriptor for gfx9-family raw buffers: DATA_FORMAT = 32.
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, kRsrcWord3);
}

// Offset used by lanes that must neither load nor store.
constexpr uint32_t kOOB = 0xFFFFFF00u;
// Base offset of a whole wave access that must read zeros: it stays out of
// range after a wave's lane offsets (a few KiB) are added to it, where kOOB
// would wrap around to small offsets.
constexpr uint32_t kOOBBase = 0x80000000u;

template <typename T>
__device__ __forceinline__ T bload(rsrc_t r, uint32_t voff, uint32_t soff);

template <>
__device__ __forceinline__ double bload<double>(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ float bload<float>(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ int bload<int>(rsrc_t r, uint32_t voff, uint32_t soff) {
    return (int)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}

// 16 bytes = 2 doubles or 4 floats.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 bload16(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
}
// Streaming (non-temporal) variants for data read once per sweep: the cache
// policy bit nt asks L2 not to keep the lines, so reused data (the solution
// vector's neighbour tiles) stays resident.
constexpr int kCacheNT = 2;
__device__ __forceinline__ u32x4 bload16_nt(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, kCacheNT);
}
template <typename T>
__device__ __forceinline__ T bload_nt(rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ double bload_nt<double>(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, kCacheNT));
}
template <>
__device__ __forceinline__ float bload_nt<float>(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, kCacheNT));
}

template <typename T>
__device__ __forceinline__ T lane_elem(const u32x4 &v, int k);
template <>
__device__ __forceinline__ double lane_elem<double>(const u32x4 &v, int k) {
    const unsigned long long b = (unsigned long long)v[2 * k] | ((unsigned long long)v[2 * k + 1] << 32);
    return __builtin_bit_cast(double, b);
}
template <>
__device__ __forceinline__ float lane_elem<float>(const u32x4 &v, int k) {
    // copy out first: bit-casting the element lvalue v[k] reads the vector's
    // first element whatever k is
    const unsigned int u = v[k];
    return __builtin_bit_cast(float, u);
}

template <typename T>
__device__ __forceinline__ void bstore(rsrc_t r, T v, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ void bstore<double>(rsrc_t r, double v, uint32_t voff, uint32_t soff) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)voff, (int)soff, 0);
}
template <>
__device__ __forceinline__ void bstore<float>(rsrc_t r, float v, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)voff, (int)soff, 0);
}

template <typename T>
__device__ __forceinline__ void bstore_nt(rsrc_t r, T v, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ void bstore_nt<double>(rsrc_t r, double v, uint32_t voff, uint32_t soff) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)voff, (int)soff, kCacheNT);
}
template <>
__device__ __forceinline__ void bstore_nt<float>(rsrc_t r, float v, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)voff, (int)soff, kCacheNT);
}

// Lane-interleaved block layout ("SoA-64"): a group of 64 elements stores each
// BS-entry block as NC chunks of VW = 16/sizeof(T) consecutive entries per
// lane ([chunk][lane][VW]), then the BS - NC*VW remaining entries
// ([entry][lane]); every chunk load of a wave is 1 KiB contiguous.
template <int BS, typename T>
struct Soa64 {
    static constexpr int VW = 16 / (int)sizeof(T);
    static constexpr int NC = BS / VW;
    static constexpr int REM = BS - NC * VW;
    static constexpr int ELEMS = BS * 64;            // T values per block slot
    // index (in T units, within one slot) of entry q of lane `lane`
    __host__ __device__ static constexpr long index(int q, int lane) {
        return q < NC * VW ? (long)(q / VW) * 64 * VW + (long)lane * VW + q % VW
                           : (long)NC * VW * 64 + (long)(q - NC * VW) * 64 + lane;
    }
};

// Factor-window layout of the Thomas kernel: a window holds, for each of LW
// lanes, one row of N = NV entries; entries are stored in 16-byte chunks per
// lane ([chunk][lane][VW]) followed by the remaining entries ([entry][lane]),
// so one wave load moves LW*16 bytes.
template <int N, int LW, typename T>
struct WinLayout {
    static constexpr int VW = 16 / (int)sizeof(T);
    static constexpr int NC = N / VW;
    static constexpr int REM = N - NC * VW;
    static constexpr int ELEMS = N * LW;              // T values of data per window
    // window stride in T values, padded to a multiple of 16 bytes so every
    // 16-byte chunk load is aligned
    static constexpr int STRIDE = (int)(((long)ELEMS * (long)sizeof(T) + 15) / 16 * 16 / (long)sizeof(T));
    __host__ __device__ static constexpr long index(int j, int lane) {
        return j < NC * VW ? (long)(j / VW) * LW * VW + (long)lane * VW + j % VW
                           : (long)NC * VW * LW + (long)(j - NC * VW) * LW + lane;
    }
};

}  // namespace tbs
