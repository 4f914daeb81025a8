// lij_fast.hip -- the fused LIJ sweep path ("plan" / "batch").
//
// One LIJ sweep of the reference (src/main.cxx:207-229) is
//     dU = -M^-1 R  (Thomas per line),  U += dU,  R = rhs + J U.
// Here it is two kernels plus a tiny norm/convergence kernel, captured in a
// hipGraph:
//   fres_kernel   residual R = rhs + D U + sum_faces Off U(nb) for a 64-element
//                 group per wave (thread per element, SoA-64 block layout, all
//                 block loads 512 B coalesced), the per-group sum of R^2, and
//                 the Thomas right-hand side g = D'^-1 (-R) (W = D'^-1 is a
//                 factor).  R itself is only written when asked for.
//   fthomas_kernel two waves per line, one from each end ("twisted" block
//                 Thomas: the factorisation runs top-down to step m-1 and
//                 bottom-up to step m+1 and closes at step m).  Forward:
//                 d_k = g_k - E_k d_{k-1} with E_k = D'_k^-1 A_k (bottom half:
//                 d_{k+1}, E_k = D'_k^-1 C_{k+1}); step m takes both
//                 neighbours; backward: d_k = d'_k - Gamma_k d_{k+1} with
//                 Gamma_k = D'_k^-1 C_{k+1} (the reference's DinvC; bottom
//                 half: d_{k-1}, D'_k^-1 A_k), and U_k += d_k.  Each wave's
//                 chain is half a line long.  A window of SPW = floor(64/NV) consecutive steps
//                 is spread over the wave, lane = step*NV + row; each step's
//                 9x9 matvec uses the previous step's vector as uniform SGPR
//                 operands (v_readlane), so the per-step dependency chain is one
//                 FMA column plus the broadcast.  Factor windows are stored
//                 [window][column j][lane], so each load is contiguous.
//   ffinal_kernel per system: fixed-order sum of the group partials -> ||R||,
//                 history, convergence / divergence flags.
// These are algebraically the reference's sweep; rounding differs from the
// reference's solveLU-per-step form by O(eps * cond(D')) because the factors
// are applied as D'^-1 products (computed from the reference's LU factors by
// its own solveLU).
//
// Many independent systems (a batch) run in one launch per kernel through
// descriptor tables: lines -> (system, m0, n, window offset), groups ->
// (system, group).  Every per-system quantity is computed from that system's
// data alone and reduced in a fixed order, so a system's result is bitwise
// the same alone and inside any batch.
#include "tbs_common.hpp"
#include "lane_ops.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <climits>
#include <cstring>
#include <functional>
#include <type_traits>
#include <map>
#include <memory>

namespace {

using tbs::rdlane;

constexpr int kGroup = 64;      // elements per residual group (one wave)
constexpr int kHistCap = 4096;  // sweeps of history kept per system
constexpr int kPad = 64;                  // zero elements before and after U, rhs, R
constexpr int kSlotNone = INT_MIN;        // no lane of the group has this face
constexpr int kSlotGather = INT_MIN + 1;  // irregular neighbours: per-lane gather

// Two-ended ("twisted") split of a line of n elements between the two waves
// of its workgroup: one runs steps 0..m-1 top-down (q whole windows of spw
// steps), the other steps n-1..m+1 bottom-up (nb windows, the first pad
// steps of its first window empty), and step m closes the line from both
// sides.  A line's factor windows: the q top windows, one window holding
// step m (at its first step), then the nb bottom windows in bottom-up order.
struct Split { int q, m, B, nb, pad; };
__host__ __device__ inline Split split_line(int n, int spw) {
    Split s;
    const int wall = (n - 1 + spw - 1) / spw;   // windows of the n - 1 steps off the middle
    s.q = wall / 2;
    s.m = s.q * spw;
    s.B = n - 1 - s.m;
    s.nb = (s.B + spw - 1) / spw;
    s.pad = s.nb * spw - s.B;
    return s;
}
__host__ __device__ inline int line_windows(int n, int spw) {
    const Split s = split_line(n, spw);
    return s.q + 1 + s.nb;
}

template <typename TF>
struct SysDev {
    const double *blk;   // [ngroups][F+1][BS][64] residual blocks (slot 0 = D)
    const int *nbr;      // [ngroups][F][64] neighbour internal index or -1
    const int *slotc;    // [ngroups][F] neighbour map of the slot: kSlotNone,
                         // kSlotGather, or the shift c with nbr = m + c for
                         // every lane that has a face in this slot
    const TF *W;         // [ngroups][BS][64] D'^-1
    const TF *EF;        // [windows][NV][SPW*NV] E_k rows, zero padded
    const TF *GB;        // [windows][NV][SPW*NV] Gamma_k rows, zero padded
    const double *rhs;   // [n][NV] internal order; kPad zero elements either side
    double *U;           // [n][NV]; kPad zero elements either side
    TF *g;               // [n][NV] Thomas right-hand side
    TF *dUp;             // [n][NV] forward-sweep intermediate
    double *R;           // [n][NV] residual (written on request); padded like U
    double *gpart;       // [ngroups] sum of R^2 per group
    double *hist;        // [kHistCap]
    int *state;          // [0] = sweeps recorded, [1] = active, [2] = status
    int n, ngroups, F, pad;
};

// ----------------------------------------------------------- residual --
// acc[i] += B(i,j) x[j] over one SoA-64 block slot, in entry order q = i + NV*j
// (ascending j for every row i, the reference's summation order).
template <int NV, typename T>
__device__ __forceinline__ void slot_matvec(tbs::rsrc_t r, uint32_t slot_soff, int lane,
                                            const double (&x)[NV], double (&acc)[NV]) {
    using L = tbs::Soa64<NV * NV, T>;
#pragma unroll
    for (int c = 0; c < L::NC; ++c) {
        const tbs::u32x4 v = tbs::bload16_nt(r, (uint32_t)lane * 16u, slot_soff + (uint32_t)c * 1024u);
#pragma unroll
        for (int k = 0; k < L::VW; ++k) {
            const int q = c * L::VW + k;
            acc[q % NV] += (double)tbs::lane_elem<T>(v, k) * x[q / NV];
        }
    }
#pragma unroll
    for (int k = 0; k < L::REM; ++k) {
        const int q = L::NC * L::VW + k;
        const T b = tbs::bload_nt<T>(r, (uint32_t)lane * sizeof(T),
                                     slot_soff + (uint32_t)(L::NC * L::VW * 64 + k * 64) * sizeof(T));
        acc[q % NV] += (double)b * x[q / NV];
    }
}

template <int NV>
__device__ __forceinline__ void load_vec(tbs::rsrc_t r, uint32_t byte_off, double (&v)[NV]) {
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = tbs::bload<double>(r, byte_off + 8u * j, 0);
}

// Copy the vector tile of elements [m_base, m_base + 64) (AoS, NV doubles per
// element, contiguous) into this wave's LDS tile with 16-byte loads; elements
// outside the vector read as 0.  Each lane then reads its own element.  NT:
// the vector is read once (streaming cache policy).
template <int NV, bool NT = false>
__device__ __forceinline__ void stage_tile(tbs::rsrc_t r, int m_base, int lane, double *tile) {
    constexpr int TB = kGroup * NV * 8;          // tile bytes
    constexpr int NC = TB / 1024;                // full 1 KiB wave loads
    const uint32_t base = (uint32_t)(m_base * NV * 8);
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const tbs::u32x4 v = NT ? tbs::bload16_nt(r, base + 16u * lane + 1024u * k, 0)
                                : tbs::bload16(r, base + 16u * lane + 1024u * k, 0);
        tile[2 * lane + 128 * k] = tbs::lane_elem<double>(v, 0);
        tile[2 * lane + 128 * k + 1] = tbs::lane_elem<double>(v, 1);
    }
    static_assert(TB - NC * 1024 == 0 || TB - NC * 1024 == 512, "tile = 512*NV bytes");
    if (TB - NC * 1024 == 512)   // odd NV: one more 8-byte load per lane
        tile[NC * 128 + lane] = NT ? tbs::bload_nt<double>(r, base + (uint32_t)(NC * 1024) + 8u * lane, 0)
                                   : tbs::bload<double>(r, base + (uint32_t)(NC * 1024) + 8u * lane, 0);
    __builtin_amdgcn_wave_barrier();
}

template <int NV>
__device__ __forceinline__ void read_tile(const double *tile, int lane, double (&v)[NV]) {
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = tile[lane * NV + j];
    __builtin_amdgcn_wave_barrier();
}

template <int NV, typename TF>
__global__ __launch_bounds__(256) void fres_kernel(const SysDev<TF> *__restrict__ sys,
                                                   const int2 *__restrict__ gtab, int ng,
                                                   int want_R, int force) {
    constexpr int BS = NV * NV;
    __shared__ double tiles[4][kGroup * NV];
    // XCD-aware order: hardware block b runs on XCD b % 8; give each XCD a
    // contiguous run of groups so neighbouring lines share an L2.
    const int nblk = (ng + 3) / 4;
    const int per = (nblk + 7) / 8;
    const int blk = (int)(blockIdx.x % 8u) * per + (int)(blockIdx.x / 8u);
    if (blk >= nblk) return;
    const int wave = threadIdx.x >> 6;
    const int wid = __builtin_amdgcn_readfirstlane(blk * 4 + wave);
    if (wid >= ng) return;
    double *tile = tiles[wave];
    const int lane = threadIdx.x & 63;
    const int2 sg = gtab[wid];
    const SysDev<TF> &S = sys[sg.x];
    // force bit 0: run even if the system is inactive; bit 1: U is zero
    // (the first residual of a solve from zero is R = rhs: skip J U)
    if (!(force & 1) && !S.state[1]) return;
    const bool u_zero = (force & 2) != 0;
    const int g = sg.y;
    const int F = S.F;
    const int m = g * kGroup + lane;
    const int m0 = g * kGroup;
    constexpr uint32_t kSlot = (uint32_t)BS * kGroup * sizeof(double);
    const tbs::rsrc_t rB = tbs::make_rsrc(S.blk + (size_t)g * (F + 1) * BS * kGroup, (F + 1) * kSlot);
    const tbs::rsrc_t rN = tbs::make_rsrc(S.nbr + (size_t)g * F * kGroup, (uint32_t)F * kGroup * 4u);
    const tbs::rsrc_t rW = tbs::make_rsrc(S.W + (size_t)g * BS * kGroup, (uint32_t)BS * kGroup * sizeof(TF));
    // U and rhs descriptors start kPad elements before element 0, so tiles
    // shifted by a neighbour offset never start at a negative offset while
    // they still touch real elements.
    const uint32_t pbytes = (uint32_t)(S.n + 2 * kPad) * NV * 8u;
    const tbs::rsrc_t rU = tbs::make_rsrc(S.U - kPad * NV, pbytes);
    const tbs::rsrc_t rRhs = tbs::make_rsrc(S.rhs - kPad * NV, pbytes);
    const uint32_t moff = (uint32_t)m * NV * 8u;

    double u[NV], acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    if (!u_zero) {
    stage_tile<NV>(rU, m0 + kPad, lane, tile);
    read_tile<NV>(tile, lane, u);
    slot_matvec<NV, double>(rB, 0u, lane, u, acc);                     // D U
    for (int f = 0; f < F; ++f) {
        const int c = S.slotc[g * F + f];
        if (c == kSlotNone) continue;
        const int nb = tbs::bload<int>(rN, (uint32_t)(f * kGroup + lane) * 4u, 0);
        double un[NV];
        if (c != kSlotGather) {                                           // wave-uniform
            stage_tile<NV>(rU, m0 + c + kPad, lane, tile);
            read_tile<NV>(tile, lane, un);
        }
        if (nb >= 0) {                                                    // interior face
            if (c == kSlotGather) load_vec<NV>(rU, (uint32_t)(nb + kPad) * NV * 8u, un);
            double t[NV];
#pragma unroll
            for (int i = 0; i < NV; ++i) t[i] = 0.0;
            slot_matvec<NV, double>(rB, (uint32_t)(1 + f) * kSlot, lane, un, t);
#pragma unroll
            for (int i = 0; i < NV; ++i) acc[i] += t[i];
        }
    }
    }
    double rh[NV], r[NV], nr[NV];
    stage_tile<NV, true>(rRhs, m0 + kPad, lane, tile);
    read_tile<NV>(tile, lane, rh);
    double rr = 0.0;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        r[i] = rh[i] + acc[i];                                            // 0 for padded lanes
        rr += r[i] * r[i];
        nr[i] = -r[i];
    }
    double gv[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) gv[i] = 0.0;
    slot_matvec<NV, TF>(rW, 0u, lane, nr, gv);                          // g = D'^-1 (-R)
    const tbs::rsrc_t rG = tbs::make_rsrc(S.g, (uint32_t)S.n * NV * sizeof(TF));
    const uint32_t goff = (uint32_t)m * NV * sizeof(TF);
#pragma unroll
    for (int i = 0; i < NV; ++i) tbs::bstore<TF>(rG, (TF)gv[i], goff + (uint32_t)i * sizeof(TF), 0);
    if (want_R) {
        const tbs::rsrc_t rR = tbs::make_rsrc(S.R, (uint32_t)S.n * NV * 8u);
#pragma unroll
        for (int i = 0; i < NV; ++i) tbs::bstore<double>(rR, r[i], moff + 8u * i, 0);
    }
    rr = tbs::wave_sum(rr);
    if (lane == 0) S.gpart[g] = rr;
}

// -------------------------------------------------------------- Thomas --
// Vector I/O of a group of windows: V values of T, contiguous in memory,
// moved with 16-byte wave loads/stores and staged in LDS ([flat] layout equal
// to the memory layout).  NCH 1 KiB chunks cover V values (+1 spare value so
// the idle lane of a window may read one past the end).
template <int V, typename T>
struct GroupIO {
    static constexpr int VW = 16 / (int)sizeof(T);
    static constexpr int NCH = (V + 1 + 64 * VW - 1) / (64 * VW);
    static constexpr int LDS = NCH * 64 * VW;            // padded LDS length
};

template <int V, typename T, bool NT = false>
__device__ __forceinline__ void bulk_load(tbs::rsrc_t r, uint32_t base, int lane,
                                          tbs::u32x4 (&reg)[GroupIO<V, T>::NCH]) {
#pragma unroll
    for (int c = 0; c < GroupIO<V, T>::NCH; ++c)
        reg[c] = NT ? tbs::bload16_nt(r, base + (uint32_t)(c * 1024 + lane * 16), 0)
                    : tbs::bload16(r, base + (uint32_t)(c * 1024 + lane * 16), 0);
}

template <int V, typename T>
__device__ __forceinline__ void bulk_commit(const tbs::u32x4 (&reg)[GroupIO<V, T>::NCH], int lane, T *lds) {
    using IO = GroupIO<V, T>;
#pragma unroll
    for (int c = 0; c < IO::NCH; ++c) {
#pragma unroll
        for (int k = 0; k < IO::VW; ++k) lds[c * 64 * IO::VW + lane * IO::VW + k] = tbs::lane_elem<T>(reg[c], k);
    }
    __builtin_amdgcn_wave_barrier();
}

// Store exactly V values from LDS (16-byte stores for whole chunks, per-value
// stores for the tail so nothing past the group is overwritten).
template <int V, typename T, bool NT = false>
__device__ __forceinline__ void bulk_store(tbs::rsrc_t r, uint32_t base, int lane, const T *lds) {
    using IO = GroupIO<V, T>;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int c = 0; c < IO::NCH; ++c) {
        const int first = c * 64 * IO::VW + lane * IO::VW;
        if ((c + 1) * 64 * IO::VW <= V) {                      // whole chunk inside the group
            tbs::u32x4 v;
            if constexpr (sizeof(T) == 8) {
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const unsigned long long bits = __builtin_bit_cast(unsigned long long, (double)lds[first + k]);
                    v[2 * k] = (unsigned)bits;
                    v[2 * k + 1] = (unsigned)(bits >> 32);
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) v[k] = __builtin_bit_cast(unsigned, (float)lds[first + k]);
            }
            __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)(base + (uint32_t)(c * 1024 + lane * 16)), 0,
                                                   NT ? tbs::kCacheNT : 0);
        } else {
#pragma unroll
            for (int k = 0; k < IO::VW; ++k) {
                const int q = first + k;
                const uint32_t off = q < V ? base + (uint32_t)q * sizeof(T) : tbs::kOOB;
                if (NT) tbs::bstore_nt<T>(r, lds[q], off, 0);
                else tbs::bstore<T>(r, lds[q], off, 0);
            }
        }
    }
}

// Two waves per line (split_line): wave 0 eliminates steps 0..m-1 top-down,
// wave 1 steps n-1..m+1 bottom-up, each over windows of SPW steps with PF
// factor windows prefetched in a register ring and its vectors staged per
// group of G = K*(PF+1) windows in LDS; wave 0 then closes step m from both
// sides and both waves substitute back outwards from it.  The bottom wave
// runs the same loops over its own windows; only the place of a step's
// values in memory differs (element m0+n-1-s for its step s, less pad).
template <int NV, typename TF, int PF, int K>
__global__ __launch_bounds__(128) void fthomas_kernel(const SysDev<TF> *__restrict__ sys,
                                                      const int4 *__restrict__ ltab, int nl) {
    constexpr int SPW = 64 / NV;      // steps per window
    constexpr int LW = SPW * NV;      // lanes carrying data
    constexpr int NB = PF + 1;        // register ring depth
    constexpr int G = K * NB;         // windows per group (loop iteration)
    constexpr int VG = G * LW;        // vector values per group
    using IOT = GroupIO<VG, TF>;
    using IOD = GroupIO<VG, double>;
    __shared__ TF s_vin[2][2][IOT::LDS];    // [wave] forward: g; backward: d'
    __shared__ TF s_out[2][IOT::LDS];       // [wave] forward: d' staging
    __shared__ double s_u[2][2][IOD::LDS];  // [wave] backward: U (updated in place)
    __shared__ TF s_x[3][NV];               // d' at step m-1, d' at step m+1, d at step m
    const int l = blockIdx.x;
    if (l >= nl) return;
    // {system, m0, n, window offset}, wave-uniform by construction: the
    // group loops below must be scalar loops (the chain broadcasts read
    // lanes that every step writes)
    int4 lt = ltab[l];
    lt.x = __builtin_amdgcn_readfirstlane(lt.x);
    lt.y = __builtin_amdgcn_readfirstlane(lt.y);
    lt.z = __builtin_amdgcn_readfirstlane(lt.z);
    lt.w = __builtin_amdgcn_readfirstlane(lt.w);
    const SysDev<TF> &S = sys[lt.x];
    if (!S.state[1]) return;                                            // the whole workgroup
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const bool bot = wave == 1;
    const int lane = threadIdx.x & 63;
    const int st_of_lane = lane / NV;
    const int row = lane - st_of_lane * NV;
    const int m0 = lt.y, n = lt.z;
    const Split sp = split_line(n, SPW);
    const int nw = bot ? sp.nb : sp.q;                                  // the wave's windows
    const int ngrp = (nw + G - 1) / G;
    // the bottom wave runs voff empty windows before its first one, so that
    // its groups of windows start at whole groups of elements from element
    // m+1 (every group's values start at an offset >= 0 in the wave's range)
    const int voff = bot ? ngrp * G - nw : 0;
    constexpr int WSTRIDE = tbs::WinLayout<NV, LW, TF>::STRIDE;
    const uint32_t wbytes = (uint32_t)WSTRIDE * sizeof(TF);           // one window of factors
    const long wfirst = (long)lt.w + (bot ? sp.q + 1 : 0);
    const tbs::rsrc_t rEF = tbs::make_rsrc(S.EF + wfirst * WSTRIDE, (uint32_t)nw * wbytes);
    const tbs::rsrc_t rGB = tbs::make_rsrc(S.GB + wfirst * WSTRIDE, (uint32_t)nw * wbytes);
    // the wave's elements: [m0, m0+m) top-down, [m0+m+1, m0+n) bottom-up
    const int e0 = bot ? m0 + sp.m + 1 : m0;
    const uint32_t vb = (uint32_t)(bot ? sp.B : sp.m) * NV;           // vector values of the wave
    const tbs::rsrc_t rg = tbs::make_rsrc(S.g + (size_t)NV * e0, vb * sizeof(TF));
    const tbs::rsrc_t rd = tbs::make_rsrc(S.dUp + (size_t)NV * e0, vb * sizeof(TF));
    const tbs::rsrc_t rU = tbs::make_rsrc(S.U + (size_t)NV * e0, vb * 8u);
    // Where the values of a group start in memory (first element, relative
    // to the wave's first element; always >= 0) and where a step of window bb
    // (lane = st*NV + row) sits in the group's LDS copy, which has the memory
    // layout.  Groups are numbered in the direction of the forward pass: the
    // forward pass visits groups 0, 1, ..., the backward pass the reverse;
    // group jg holds the wave's windows jg*G ... jg*G+G-1 (bottom wave:
    // counted with its voff empty windows).
    const int lofs = lane < LW ? (bot ? row - st_of_lane * NV : lane) : 0;
    auto gbase = [&](int jg) -> int { return (bot ? ngrp - 1 - jg : jg) * G * SPW; };
    auto fwd_slot = [&](int bb) -> int { return lane < LW ? (bot ? VG - NV - bb * LW + lofs : bb * LW + lofs) : 0; };
    auto bwd_slot = [&](int bb) -> int { return lane < LW ? (bot ? (bb + 1) * LW - NV + lofs : (G - 1 - bb) * LW + lofs) : 0; };
    auto vbytes = [&](int elem, int bytes_per) -> uint32_t { return (uint32_t)(elem * NV * bytes_per); };

    using WL = tbs::WinLayout<NV, LW, TF>;
    auto wbase = [&](int w) -> uint32_t { return w >= 0 ? (uint32_t)w * wbytes : tbs::kOOBBase; };
    auto load_row = [&](tbs::rsrc_t r, int w, TF (&e)[NV]) {
        const uint32_t b = wbase(w);
#pragma unroll
        for (int c = 0; c < WL::NC; ++c) {
            const tbs::u32x4 v = tbs::bload16_nt(r, b + (uint32_t)(c * LW * 16 + lane * 16), 0);
#pragma unroll
            for (int k = 0; k < WL::VW; ++k) e[c * WL::VW + k] = tbs::lane_elem<TF>(v, k);
        }
#pragma unroll
        for (int k = 0; k < WL::REM; ++k)
            e[WL::NC * WL::VW + k] = tbs::bload_nt<TF>(r, b + (uint32_t)((WL::NC * WL::VW * LW + k * LW + lane) * sizeof(TF)), 0);
    };

    if (lane < NV) s_x[wave][lane] = TF(0);                             // an empty side contributes 0
    TF d[NV];
    TF eb[NB][NV];
    // ---- forward: d_k = g_k - E_k d_{k-1} (bottom wave: d_{k+1})
#pragma unroll
    for (int j = 0; j < NV; ++j) d[j] = TF(0);
#pragma unroll
    for (int p = 0; p < PF; ++p) load_row(rEF, p - voff, eb[p]);
    {
        tbs::u32x4 vr[IOT::NCH];
        bulk_load<VG, TF, true>(rg, vbytes(gbase(0), sizeof(TF)), lane, vr);
        bulk_commit<VG, TF>(vr, lane, s_vin[wave][0]);
    }
    for (int jg = 0; jg < ngrp; ++jg) {
        const int cur = jg & 1;
        tbs::u32x4 vr[IOT::NCH];
        if (jg + 1 < ngrp)                                              // next group's g
            bulk_load<VG, TF, true>(rg, vbytes(gbase(jg + 1), sizeof(TF)), lane, vr);
#pragma unroll
        for (int bb = 0; bb < G; ++bb) {
            const int w = jg * G + bb - voff;                           // the wave's window
            const int b = bb % NB;
            const int bn = (b + PF) % NB;
            load_row(rEF, w + PF, eb[bn]);
            const int slot = fwd_slot(bb);
            const TF gv = s_vin[wave][cur][slot];
            TF out = TF(0);
#pragma unroll
            for (int st = 0; st < SPW; ++st) {
                TF x0 = TF(0), x1 = TF(0), x2 = TF(0);
#pragma unroll
                for (int j = 0; j < NV; j += 3) {
                    x0 += eb[b][j] * d[j];
                    if (j + 1 < NV) x1 += eb[b][j + 1] * d[j + 1];
                    if (j + 2 < NV) x2 += eb[b][j + 2] * d[j + 2];
                }
                const TF v = gv - ((x0 + x1) + x2);
#pragma unroll
                for (int j = 0; j < NV; ++j) d[j] = rdlane(v, st * NV + j);
                out = (st_of_lane == st) ? v : out;
            }
            if (lane < LW) s_out[wave][slot] = out;
            // the wave's last step (step m-1, or m+1 for the bottom wave)
            if (w == nw - 1 && st_of_lane == SPW - 1 && lane < LW) s_x[wave][row] = out;
        }
        bulk_store<VG, TF>(rd, vbytes(gbase(jg), sizeof(TF)), lane, s_out[wave]);
        if (jg + 1 < ngrp) bulk_commit<VG, TF>(vr, lane, s_vin[wave][cur ^ 1]);
    }
    __syncthreads();
    // ---- step m: d_m = g_m - E_m d'_{m-1} - F_m d'_{m+1}, U_m += d_m
    if (!bot && lane < NV) {
        const tbs::rsrc_t rME = tbs::make_rsrc(S.EF + ((long)lt.w + sp.q) * WSTRIDE, wbytes);
        const tbs::rsrc_t rMF = tbs::make_rsrc(S.GB + ((long)lt.w + sp.q) * WSTRIDE, wbytes);
        TF em[NV], fm[NV];
        load_row(rME, 0, em);
        load_row(rMF, 0, fm);
        const size_t at = (size_t)NV * (m0 + sp.m) + lane;
        TF x0 = TF(0), x1 = TF(0);
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            x0 += em[j] * s_x[0][j];
            x1 += fm[j] * s_x[1][j];
        }
        const TF v = S.g[at] - (x0 + x1);
        s_x[2][lane] = v;
        S.U[at] += (double)v;
    }
    __syncthreads();

    // ---- backward: d_k = d'_k - Gamma_k d_{k+1} (bottom wave: d_{k-1}); U_k += d_k
    // The groups in reverse, each from its last window down.  The top wave's
    // last group may end past its last window: those windows are skipped
    // (d must reach window nw-1 unchanged); the bottom wave's empty windows
    // come last and change nothing that is stored.
#pragma unroll
    for (int j = 0; j < NV; ++j) d[j] = s_x[2][j];
    const int wtop = ngrp * G - voff;                                   // one past the last window visited
#pragma unroll
    for (int p = 0; p < PF; ++p) load_row(rGB, wtop - 1 - p, eb[p]);
    {
        tbs::u32x4 vr[IOT::NCH];
        tbs::u32x4 ur[IOD::NCH];
        bulk_load<VG, TF, true>(rd, vbytes(gbase(ngrp - 1), sizeof(TF)), lane, vr);
        bulk_load<VG, double>(rU, vbytes(gbase(ngrp - 1), 8), lane, ur);
        bulk_commit<VG, TF>(vr, lane, s_vin[wave][0]);
        bulk_commit<VG, double>(ur, lane, s_u[wave][0]);
    }
    for (int t = 0; t < ngrp; ++t) {
        const int jg = ngrp - 1 - t;
        const int cur = t & 1;
        tbs::u32x4 vr[IOT::NCH];
        tbs::u32x4 ur[IOD::NCH];
        if (jg > 0) {                                                   // next group's d', U
            bulk_load<VG, TF, true>(rd, vbytes(gbase(jg - 1), sizeof(TF)), lane, vr);
            bulk_load<VG, double>(rU, vbytes(gbase(jg - 1), 8), lane, ur);
        }
#pragma unroll
        for (int bb = 0; bb < G; ++bb) {
            const int w = wtop - 1 - (t * G + bb);
            const int b = bb % NB;
            const int bn = (b + PF) % NB;
            load_row(rGB, w - PF, eb[bn]);
            if (w >= nw) continue;                                      // past the top wave's last window
            const int slot = bwd_slot(bb);
            const TF pv = s_vin[wave][cur][slot];
            TF out = TF(0);
#pragma unroll
            for (int st = SPW - 1; st >= 0; --st) {
                TF x0 = TF(0), x1 = TF(0), x2 = TF(0);
#pragma unroll
                for (int j = 0; j < NV; j += 3) {
                    x0 += eb[b][j] * d[j];
                    if (j + 1 < NV) x1 += eb[b][j + 1] * d[j + 1];
                    if (j + 2 < NV) x2 += eb[b][j + 2] * d[j + 2];
                }
                const TF v = pv - ((x0 + x1) + x2);
#pragma unroll
                for (int j = 0; j < NV; ++j) d[j] = rdlane(v, st * NV + j);
                out = (st_of_lane == st) ? v : out;
            }
            if (lane < LW) s_u[wave][cur][slot] += (double)out;
        }
        bulk_store<VG, double>(rU, vbytes(gbase(jg), 8), lane, s_u[wave][cur]);
        if (jg > 0) {
            bulk_commit<VG, TF>(vr, lane, s_vin[wave][cur ^ 1]);
            bulk_commit<VG, double>(ur, lane, s_u[wave][cur ^ 1]);
        }
    }
}

// ------------------------------------------------ norm / convergence --
// One block per system: fixed-order sum of the group partials (thread t
// adds partials t, t + 1024, ... in order, then a fixed LDS tree).  Each
// thread issues the loads of up to kFinalPer partials before it adds any, so
// the pass costs one memory round trip per kFinalPer * 1024 partials.
constexpr int kFinalThreads = 1024;
constexpr int kFinalPer = 32;
template <typename TF>
__global__ __launch_bounds__(kFinalThreads) void ffinal_kernel(const SysDev<TF> *__restrict__ sys, int nsys,
                                                               int max_sweeps, double rtol, int force) {
    __shared__ double s[kFinalThreads];
    const int sidx = blockIdx.x;
    if (sidx >= nsys) return;
    const SysDev<TF> &S = sys[sidx];
    if (!force && !S.state[1]) return;
    const int ng = S.ngroups;
    double acc = 0.0;
    for (int base = 0; base < ng; base += kFinalPer * kFinalThreads) {
        double v[kFinalPer];
#pragma unroll
        for (int q = 0; q < kFinalPer; ++q) {
            const int i = base + q * kFinalThreads + (int)threadIdx.x;
            v[q] = i < ng ? S.gpart[i] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kFinalPer; ++q)
            if (base + q * kFinalThreads + (int)threadIdx.x < ng) acc += v[q];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kFinalThreads / 2; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double h = sqrt(s[0]);
        const int k = S.state[0];
        if (k < kHistCap) S.hist[k] = h;
        S.state[0] = k + 1;
        int active = 1, status = 0;
        if (!(h == h) || h > 1.0e300) { active = 0; status = 2; }          // non-finite
        else if (k >= 1 && h <= rtol * S.hist[0]) { active = 0; status = 1; }  // converged
        else if (k >= max_sweeps || k + 1 >= kHistCap) { active = 0; status = 3; }
        else if (k >= 8 && h > 1.0e6 * S.hist[0]) { active = 0; status = 4; }  // diverging
        S.state[1] = active;
        S.state[2] = status;
    }
}

template <typename TF>
__global__ void reset_state_kernel(const SysDev<TF> *__restrict__ sys, int nsys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsys) {
        sys[i].state[0] = 0;
        sys[i].state[1] = 1;
        sys[i].state[2] = 0;
    }
}

template <typename TF>
__global__ void gather_states_kernel(const SysDev<TF> *__restrict__ sys, int nsys, int *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4 * nsys) out[i] = sys[i >> 2].state[i & 3];
}

// ------------------------------------------------------------ packing --
// Residual blocks into the SoA-64 layout (tbs::Soa64<BS, double>): src code
// per (internal element, slot): x = array (0 none, 1 jacD, 2 jacO1, 3 jacO2),
// y = block index.  The destination is enumerated in order (coalesced writes).
template <int NV>
__global__ void pack_blocks_kernel(long total, int S, const int2 *__restrict__ src,
                                   const double *__restrict__ D, const double *__restrict__ O1,
                                   const double *__restrict__ O2, double *__restrict__ blk) {
    constexpr int BS = NV * NV;
    using L = tbs::Soa64<BS, double>;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int idx = (int)(t % L::ELEMS);
        const long gs = t / L::ELEMS;
        const int slot = (int)(gs % S);
        const long g = gs / S;
        int q, lane;
        if (idx < L::NC * L::VW * 64) {
            const int c = idx / (64 * L::VW), rem = idx % (64 * L::VW);
            lane = rem / L::VW;
            q = c * L::VW + rem % L::VW;
        } else {
            const int r = idx - L::NC * L::VW * 64;
            q = L::NC * L::VW + r / 64;
            lane = r % 64;
        }
        const long m = g * kGroup + lane;
        const int2 c = src[m * S + slot];
        double v = 0.0;
        if (c.x == 1) v = D[(size_t)BS * c.y + q];
        else if (c.x == 2) v = O1[(size_t)BS * c.y + q];
        else if (c.x == 3) v = O2[(size_t)BS * c.y + q];
        blk[t] = v;
    }
}

// Same packing through LDS: one 256-thread workgroup per (group, slot) reads
// the group's 64 source blocks whole (consecutive threads read consecutive
// entries of a block, so every load is contiguous), then writes the slot in
// SoA-64 order (consecutive threads write consecutive addresses).
// One 128-thread workgroup per half of a (group, slot): the 32 source blocks
// of lanes h*32 .. h*32+31 come into LDS whole, in 16-byte pieces (consecutive
// threads read consecutive pieces of a block; a block is contiguous in the
// source array, 8-byte aligned), and leave as those lanes' part of the SoA-64
// slot: 512-byte runs of 16-byte pieces.  (Half groups keep the LDS of a
// workgroup at 21 KB, so seven fit a CU.)
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));
constexpr int kPackLanes = 32;
template <int NV>
__global__ __launch_bounds__(128) void pack_lds_kernel(int S, const int2 *__restrict__ src,
                                                       const double *__restrict__ D,
                                                       const double *__restrict__ O1,
                                                       const double *__restrict__ O2,
                                                       double *__restrict__ blk) {
    constexpr int BS = NV * NV;
    constexpr int PB = (BS + 1) / 2;           // 16-byte pieces per block (the last may be 8 bytes)
    constexpr int EL = kPackLanes;
    using L = tbs::Soa64<BS, double>;
    static_assert(L::VW == 2, "SoA-64 slots of doubles hold 2 entries per 16-byte chunk");
    __shared__ double s[EL * BS];              // [element lane - h*EL][entry]
    __shared__ const double *s_src[EL];
    const long w = blockIdx.x >> 1;            // (group, slot)
    const int h = (int)(blockIdx.x & 1u);      // half of the group's lanes
    const long g = w / S;
    const int slot = (int)(w - g * S);
    const int t = threadIdx.x;
    if (t < EL) {
        const int2 c = src[(g * kGroup + h * EL + t) * S + slot];
        const double *a = c.x == 1 ? D : (c.x == 2 ? O1 : (c.x == 3 ? O2 : nullptr));
        s_src[t] = a ? a + (size_t)BS * c.y : nullptr;
    }
    __syncthreads();
    for (int i = t; i < EL * PB; i += 128) {
        const int e = i / PB, k = i - e * PB;
        const double *p = s_src[e];
        double *d = s + e * BS + 2 * k;
        if (2 * k + 1 < BS) {
            const f64x2u v = p ? *reinterpret_cast<const f64x2u *>(p + 2 * k) : f64x2u{0.0, 0.0};
            d[0] = v[0];
            d[1] = v[1];
        } else {
            d[0] = p ? p[2 * k] : 0.0;
        }
    }
    __syncthreads();
    double *out = blk + w * L::ELEMS;
    for (int i = t; i < L::NC * EL; i += 128) {              // chunk c, lane h*EL + e: entries 2c, 2c+1
        const int c = i / EL, e = i - c * EL;
        const double *q = s + e * BS + 2 * c;
        *reinterpret_cast<f64x2u *>(out + c * 128 + (h * EL + e) * 2) = f64x2u{q[0], q[1]};
    }
    for (int i = t; i < L::REM * EL; i += 128)               // odd BS: the last entry, [lane]
        out[L::NC * 128 + h * EL + i] = s[i * BS + BS - 1];
}

// ------------------------------------------------------------- factor --
// 1/x from the hardware reciprocal estimate refined by two Newton steps
// (each doubles the correct bits), for the Gauss-Jordan pivots.
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Block-Thomas factorisation of every line straight into the sweep-form
// factors.  src/okl/lineLU_v2.okl:58-182 runs the same recurrence on LU
// factors; with A_k, C_k the blocks at face f_k = lineface[m0+k] between
// elements k-1 and k (A = Of2 and C = Of1 when element k is the face's
// first element fc[2f], else the reverse, lineLU_v2.okl:116-117):
//   S_0 = D_0,   S_k = D_k - A_k Gamma_{k-1}               (:147-162)
//   [Gamma_k | E_k | W_k] = S_k^{-1} [C_{k+1} | A_k | I]
// by Gauss-Jordan elimination without pivoting (the reference's LU does not
// pivot either).  Gamma_k is the reference's DinvC, W_k = D'_k^{-1} and
// E_k = D'_k^{-1} A_k are what the sweep kernels read.
//
// Two waves per line (split_line): wave 0 runs steps 0..m-1 as above, wave 1
// runs steps n-1..m+1 of the same recurrence bottom-up (S_k = D_k -
// C_{k+1} Gamma''_{k+1}, [Gamma''_k | E''_k | W_k] = S_k^{-1} [A_k | C_{k+1} |
// I]: the same step with the roles of A and C exchanged), and wave 0 then
// closes step m with both neighbours, S_m = D_m - A_m Gamma_{m-1} -
// C_{m+1} Gamma''_{m+1}, [F_m | E_m | W_m] = S_m^{-1} [C_{m+1} | A_m | I]
// (F_m is stored in Gamma's place).
// In each wave, one lane per column: lanes P..P+NV-1 hold S_k (the
// pivot columns), X..X+NV-1 hold C_{k+1} and end with Gamma_k, 2NV.. hold A_k
// then E_k, 3NV.. hold I then W_k.  P and X swap every step, so the lanes
// that end a step holding Gamma_k form S_{k+1} in place; each pivot step
// broadcasts one column with readlane.  A lone wave issues one instruction
// every four cycles whatever its kind, so the step is written as one
// instruction stream for every lane, with no branches on the lane's role:
// each lane reads its new column through a per-lane LDS pointer (D, C or A
// of the step, or a resident identity block), every lane applies the
// S_k update with its old column scaled by 1 (pivot lanes) or 0 (the rest),
// and the outputs go out through per-lane destination offsets
// u_i + s * v_i precomputed for the lane's role.  The D, A, C blocks of 4
// steps at a time (2 for blocks larger than 10x10) are prefetched into
// double-buffered LDS one chunk ahead;
// their sources come from the per-element table fidx = {e, A source, C
// source of the face before} (source = 2 * face + 1 for Of2, 2 * face for
// Of1, -1 none: the zero block).
// What the factor kernel needs of one plan: its line tables, its Jacobian
// and its factor arrays.  One plan's record is a kernel argument; many plans
// factored in one launch read theirs from device memory through a line table
// {record, line}.
struct FactorRec {
    const int *linepoint, *linesize, *woff;
    const int4 *fidx;
    const double *D, *O1, *O2, *zero;   // zero: a block of zeros (the rhs padding)
    void *W, *EF, *GB;
};

template <int NV, typename TF>
__global__ __launch_bounds__(128) void ffactor_kernel(FactorRec one, int nline, const FactorRec *__restrict__ recs,
                                                      const int2 *__restrict__ ltab) {
    constexpr int BS = NV * NV;
    constexpr int CH = NV <= 10 ? 4 : 2;        // steps per chunk (LDS: two waves, two buffers)
    constexpr int NB = 3 * CH;                  // blocks per chunk: D_k, A_k, C_{k+1} of each step
    constexpr int SB = 3 * BS;                  // doubles per step
    constexpr int NLD = (CH * SB + 63) / 64;    // loads per lane per chunk
    constexpr int SPW = 64 / NV, LW = SPW * NV;
    using WL = tbs::WinLayout<NV, LW, TF>;
    using WS = tbs::Soa64<BS, TF>;
    static_assert(4 * NV <= 64, "lane roles need 4*NV lanes");
    static_assert(BS <= 256, "entry index kept in 8 bits");
    __shared__ double sblk_all[2][2][CH * SB];  // [wave][buffer]
    __shared__ double sI[BS];                   // identity: the W lanes' starting column
    __shared__ const double *sptr_all[2][2][NB];  // [wave][buffer]: sources of a chunk's blocks

    if ((int)blockIdx.x >= nline) return;
    int l = blockIdx.x;
    FactorRec fr = one;
    if (ltab) {
        const int2 e = ltab[blockIdx.x];
        fr = recs[__builtin_amdgcn_readfirstlane(e.x)];
        l = __builtin_amdgcn_readfirstlane(e.y);
    }
    const int *__restrict__ linepoint = fr.linepoint;
    const int *__restrict__ linesize = fr.linesize;
    const int *__restrict__ woff = fr.woff;
    const int4 *__restrict__ fidx = fr.fidx;
    const double *__restrict__ jacD = fr.D;
    const double *__restrict__ O1 = fr.O1;
    const double *__restrict__ O2 = fr.O2;
    const double *__restrict__ zero_blk = fr.zero;
    TF *__restrict__ W = (TF *)fr.W;
    TF *__restrict__ EF = (TF *)fr.EF;
    TF *__restrict__ GB = (TF *)fr.GB;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const bool bot = wave == 1;
    const int lane = threadIdx.x & 63;
    auto sblk = sblk_all[wave];
    auto sptr = sptr_all[wave];
    // Gamma''_{m+1} for step m, from the bottom wave's last step, in its
    // block buffer (no longer read by then)
    double *sG = &sblk_all[1][0][0];
    // wave-uniform by construction: the chunk and step loops must be scalar
    // loops (the pivot broadcasts read lanes that every step writes)
    const int m0 = __builtin_amdgcn_readfirstlane(linepoint[l]);
    const int n = __builtin_amdgcn_readfirstlane(linesize[l]);
    const Split sp = split_line(n, SPW);
    const int nsteps = bot ? sp.B : sp.m + 1;   // the top wave's last step is step m
    // the wave's windows: top q + 1 (the last holds step m), bottom nb
    const long wbase = ((long)__builtin_amdgcn_readfirstlane(woff[l]) + (bot ? sp.q + 1 : 0)) * WL::STRIDE;
    // a block of 128 threads advances its two waves' LDS use apart: within a
    // wave, LDS written by some lanes is read by others in program order
    auto wsync = []() { __builtin_amdgcn_wave_barrier(); };

    // ---- per-lane constants
    const int grp = lane / NV;                  // role group 0..3 (4+: idle lanes)
    const int j = lane - grp * NV;              // the lane's column
    if (!bot)
        for (int t = lane; t < BS; t += 64) sI[t] = (t % NV == t / NV) ? 1.0 : 0.0;
    // chunk entry t = lane + 64 r: block (step * 3 + which) << 8 | entry
    int code[NLD];
#pragma unroll
    for (int r = 0; r < NLD; ++r) {
        const int t = lane + 64 * r;
        const int st = t / SB, rem = t - st * SB;
        const int which = rem / BS;
        code[r] = t < CH * SB ? ((st * 3 + which) << 8) | (rem - which * BS) : -1;
    }
    // outputs: lanes of group 0/1 (X, parity-dependent) and 2 write window
    // entries WL::index(j, st * NV + i) = u_i + st * v_i; group 3 writes W
    // entries WS::index(i + NV j, lane of the element) = u_i + elane * v_i
    int u[NV], v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        if (grp == 3) {
            u[i] = (int)WS::index(i + NV * j, 0);
            v[i] = (int)(WS::index(i + NV * j, 1) - u[i]);
        } else {
            u[i] = (int)WL::index(j < NV ? j : 0, i);
            v[i] = (int)(WL::index(j < NV ? j : 0, NV) - WL::index(j < NV ? j : 0, 0));
        }
    }
    const double fP[2] = {grp == 0 ? 1.0 : 0.0, grp == 1 ? 1.0 : 0.0};   // pivot lanes at parity 0 / 1

    // lane t < NB: block t of a chunk is (wave step t / 3, slot t % 3) and
    // its source is a component of fidx ({element, A source, C source of the
    // face before}): top wave, line step k = s: slots D_k, A_k, C_{k+1}
    // (fidx[k].x, fidx[k].y, fidx[k+1].z); bottom wave, k = n-1-s: slots D_k,
    // C_{k+1}, A_k (fidx[k].x, fidx[k+1].z, fidx[k].y); -1: none
    const int lst = lane / 3, lwhich = lane - 3 * lst;
    const int nextslot = bot ? 1 : 2;
    auto load_idx = [&](int c) -> int {
        const int s = c * CH + lst;
        const int k = (bot ? n - 1 - s : s) + (lwhich == nextslot ? 1 : 0);
        const int comp = lwhich == 0 ? 0 : (lwhich == nextslot ? 2 : 1);
        int code_src = -1;
        if (lane < NB && s < nsteps && k < n) code_src = reinterpret_cast<const int *>(fidx + m0 + k)[comp];
        return code_src;
    };
    // (pointers are chosen by arithmetic on their values: a select between
    // two captured pointers becomes a select of their addresses, which puts
    // them in scratch memory)
    const uintptr_t aD = (uintptr_t)jacD, aO1 = (uintptr_t)O1, aZ = (uintptr_t)zero_blk;
    const uintptr_t dO = (uintptr_t)O2 - aO1;
    auto set_ptrs = [&](int b, int cs) {
        if (lane < NB) {
            const uintptr_t face = aO1 + (uintptr_t)(cs & 1) * dO + (uintptr_t)((long)(cs >> 1) * BS * 8);
            const uintptr_t elem = aD + (uintptr_t)((long)cs * BS * 8);
            const uintptr_t p = lwhich == 0 ? elem : face;
            sptr[b][lane] = (const double *)(cs >= 0 ? p : aZ);
        }
    };
    // (the sources are global memory: loads through a global-space pointer,
    // so they count only against vmcnt and the step's LDS waits do not wait
    // for the prefetch)
    using gptr = const __attribute__((address_space(1))) double *;
    auto load_blocks = [&](int b, double (&R)[NLD]) {
#pragma unroll
        for (int r = 0; r < NLD; ++r) {
            const int cd = code[r] < 0 ? 0 : code[r];
            const gptr src = (gptr)sptr[b][cd >> 8];
            R[r] = src[cd & 255];
        }
    };
    auto store_blocks = [&](int b, const double (&R)[NLD]) {
#pragma unroll
        for (int r = 0; r < NLD; ++r) {
            const int t = lane + 64 * r;
            if (r < NLD - 1 || t < CH * SB) sblk[b][t] = R[r];
        }
    };

    double R[NLD];
    // prologue: chunk 0 in LDS, chunk 1 in flight, indices of chunk 2 in flight
    set_ptrs(0, load_idx(0));
    __syncthreads();                            // sI, and wave-local sptr
    load_blocks(0, R);
    store_blocks(0, R);
    set_ptrs(1, load_idx(1));
    wsync();
    load_blocks(1, R);
    int xnext = load_idx(2);

    double col[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) col[i] = 0.0;

    const uintptr_t aW = (uintptr_t)W, aEF = (uintptr_t)EF, dGB = (uintptr_t)GB - (uintptr_t)EF;
    auto step = [&](auto par, int b, int s, int sw) {
        constexpr int PAR = decltype(par)::value;
        constexpr int P = PAR ? NV : 0;
        const int k = bot ? n - 1 - sw : sw;    // line step of wave step sw
        const double *Dk = &sblk[b][s * SB];
        const double *Ak = Dk + BS;
        // the lane's new column: pivot lanes D_k, X lanes C_{k+1}, A lanes
        // A_k, W lanes the identity (idle lanes read the identity too)
        const int xg = PAR ? 0 : 1;
        const double *cp0 = grp == (PAR ? 1 : 0) ? Dk : (grp == xg ? Dk + 2 * BS : (grp == 2 ? Ak : sI));
        const double *bcol = cp0 + (grp < 4 ? NV * j : 0);
        double vold[NV];
#pragma unroll
        for (int m = 0; m < NV; ++m) vold[m] = col[m] * fP[PAR];
        // S_k = D_k - A_k Gamma_{k-1} on the pivot lanes (A_k uniform); the
        // other lanes subtract exact zeros
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double tot = 0.0;
#pragma unroll
            for (int m = 0; m < NV; ++m) tot += Ak[i + NV * m] * vold[m];
            col[i] = bcol[i] - tot;
        }
        // Gauss-Jordan on the pivot columns
#pragma unroll
        for (int p = 0; p < NV; ++p) {
            double cpv[NV];
#pragma unroll
            for (int i = 0; i < NV; ++i) cpv[i] = tbs::rdlane(col[i], P + p);
            const double rp = recip(cpv[p]);
            const double x = col[p] * rp;
#pragma unroll
            for (int i = 0; i < NV; ++i)
                if (i != p) col[i] -= cpv[i] * x;
            col[p] = x;
        }
        // outputs: Gamma_k (X lanes), E_k (A lanes), W_k, at the wave's
        // window position of the step (the bottom wave's windows start with
        // pad empty steps)
        const int pos = bot ? sw + sp.pad : sw;
        const int w = pos / SPW, st = pos - w * SPW;
        const int mm = m0 + k;
        const bool isW = grp == 3;
        // every window entry of a real step is written: E_0, E''_{n-1} and
        // Gamma_{n-1} (no neighbour) as zeros, so the factor arrays need no
        // zero fill
        const bool act = grp == xg || grp == 2 || isW;
        const bool zero = bot ? (grp == 2 && sw == 0) : ((grp == xg && k == n - 1) || (grp == 2 && k == 0));
        const uintptr_t wdst = aW + (uintptr_t)((size_t)(mm / kGroup) * BS * kGroup * sizeof(TF));
        const uintptr_t fdst = aEF + (uintptr_t)(grp == xg) * dGB +
                               (uintptr_t)((wbase + (long)w * WL::STRIDE) * (long)sizeof(TF));
        using gdst = __attribute__((address_space(1))) TF *;
        const gdst dst = (gdst)(TF *)(isW ? wdst : fdst);
        const int sc = isW ? mm % kGroup : st;
        if (act) {
#pragma unroll
            for (int i = 0; i < NV; ++i) dst[u[i] + sc * v[i]] = zero ? TF(0) : (TF)col[i];
        }
    };

    // one uniform loop over step pairs (CH is even); at the first step of
    // chunk c > 0: chunk c into LDS, chunk c+1's blocks and chunk c+2's
    // indices into flight
    // step m (top wave): wait for the bottom wave's last Gamma'' (the one
    // block-wide barrier of the two waves after the start), then D_m -=
    // C_{m+1} Gamma''_{m+1} in the chunk's copy of D_m, so that the plain
    // step gives S_m = D_m - C_{m+1} Gamma''_{m+1} - A_m Gamma_{m-1}
    auto middle = [&](int b, int s) {
        __syncthreads();
        double *Dm = &sblk[b][s * SB];
        const double *Cm = Dm + 2 * BS;
        for (int t = lane; t < BS; t += 64) {
            const int i = t % NV, jj = t / NV;
            double tot = 0.0;
#pragma unroll
            for (int mm = 0; mm < NV; ++mm) tot += Cm[i + NV * mm] * sG[NV * jj + mm];
            Dm[t] -= tot;
        }
        wsync();
    };
    for (int k = 0; k < nsteps; k += 2) {
        const int c = k / CH, s = k - c * CH;
        if (s == 0 && c > 0) {
            wsync();                                    // LDS of chunk c-1 read
            store_blocks(c & 1, R);                     // chunk c
            set_ptrs((c + 1) & 1, xnext);               // sources of chunk c+1
            wsync();
            load_blocks((c + 1) & 1, R);
            xnext = load_idx(c + 2);
        }
        if (!bot && k == sp.m) middle(c & 1, s);
        step(std::integral_constant<int, 0>(), c & 1, s, k);
        if (k + 1 < nsteps) {
            if (!bot && k + 1 == sp.m) middle(c & 1, s + 1);
            step(std::integral_constant<int, 1>(), c & 1, s + 1, k + 1);
        }
    }
    if (bot) {
        // Gamma''_{m+1}: the X lanes of the last step (group 1 after an even
        // step, 0 after an odd one); zeros when the bottom side is empty
        wsync();
        const int xl = sp.B == 0 ? -1 : (((sp.B - 1) & 1) ? 0 : 1);
        if (xl < 0) {
            for (int t = lane; t < BS; t += 64) sG[t] = 0.0;
        } else if (grp == xl) {
#pragma unroll
            for (int i = 0; i < NV; ++i) sG[NV * j + i] = col[i];
        }
        __syncthreads();
    }
    // empty steps of the wave's windows (top: window q after step m; bottom:
    // the first pad steps of window 0): E and Gamma zero (the Thomas sweep
    // runs whole windows and needs finite factors there); and each window's
    // alignment padding (never read as a factor; written so the arrays hold
    // no stale memory)
    using gdst = __attribute__((address_space(1))) TF *;
    const int e_lo = bot ? 0 : 1, e_hi = bot ? sp.pad : SPW;
    const int wlast = bot ? 0 : sp.q;
    for (int st = e_lo; st < e_hi; ++st) {
        if (grp == 2) {
            const uintptr_t base = (uintptr_t)((wbase + (long)wlast * WL::STRIDE) * (long)sizeof(TF));
            const gdst e = (gdst)(TF *)(aEF + base), g = (gdst)(TF *)(aEF + dGB + base);
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                e[u[i] + st * v[i]] = TF(0);
                g[u[i] + st * v[i]] = TF(0);
            }
        }
    }
    if (WL::STRIDE > WL::ELEMS && grp == 2 && j < WL::STRIDE - WL::ELEMS) {
        const int nwin = bot ? sp.nb : sp.q + 1;
        for (int w = 0; w < nwin; ++w) {
            const long at = wbase + (long)w * WL::STRIDE + WL::ELEMS + j;
            ((gdst)(TF *)aEF)[at] = TF(0);
            ((gdst)(TF *)(aEF + dGB))[at] = TF(0);
        }
    }
}

// ------------------------------------------------------- plan tables --
// The derived tables of a plan, built on the device from the mesh arrays
// (uploaded as they are): per line element the factor sources fidx and the
// line index; per group lane the residual block sources src and neighbour
// indices nbr; per group slot the neighbour map slotc.

// One block per line.  fidx = {e, A source, C source}: A_k = Of2, C_k = Of1
// at face f_k when element k is the face's first element, else the reverse
// (src/okl/lineLU_v2.okl:116-117); source = 2f + 1 for Of2, 2f for Of1.
__global__ void tab_lines_kernel(const int *__restrict__ linepoint, const int *__restrict__ linesize,
                                 const int *__restrict__ woff, const int *__restrict__ lines,
                                 const int *__restrict__ lineface, const int *__restrict__ fc,
                                 int *__restrict__ mline, int4 *__restrict__ fidx, int4 *__restrict__ ltab) {
    const int l = blockIdx.x;
    const int m0 = linepoint[l], n = linesize[l];
    if (threadIdx.x == 0) ltab[l] = make_int4(0, m0, n, woff[l]);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int mm = m0 + k, e = lines[mm];
        int a = -1, c = -1;
        if (k >= 1) {
            const int f = lineface[mm];
            const bool first = e == fc[2 * f];
            a = 2 * f + (first ? 1 : 0);
            c = 2 * f + (first ? 0 : 1);
        }
        mline[mm] = l;
        fidx[mm] = make_int4(e, a, c, 0);
    }
}

// inv[lines[i]] = i: internal position of every element.
__global__ void tab_inv_kernel(int nle, const int *__restrict__ lines, int *__restrict__ inv) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nle; i += gridDim.x * blockDim.x) inv[lines[i]] = i;
}

// One thread per group lane mm < G: src row [S] = {0 none, 1 D, 2 Of1, 3 Of2;
// index} (slot 0 = D, slot 1 + s = the element's s-th face in ef order: Of2
// when the element is the face's first element), nbr[g][s][lane] = internal
// index of the neighbour across that face or -1; lanes past the last
// element stay empty.
__global__ void tab_src_kernel(long G, int nle, int S, int Fm, const int *__restrict__ lines,
                               const int *__restrict__ epoint, const int *__restrict__ ef,
                               const int *__restrict__ fc, const int *__restrict__ inv, int2 *__restrict__ src,
                               int *__restrict__ nbr) {
    for (long mm = (long)blockIdx.x * blockDim.x + threadIdx.x; mm < G; mm += (long)gridDim.x * blockDim.x) {
        const long g = mm / kGroup;
        const int lane = (int)(mm % kGroup);
        int2 *row = src + mm * S;
        for (int q = 0; q < S; ++q) row[q] = make_int2(0, 0);
        for (int s2 = 0; s2 < Fm; ++s2) nbr[(g * Fm + s2) * kGroup + lane] = -1;
        if (mm >= nle) continue;
        const int e = lines[mm];
        row[0] = make_int2(1, e);
        int s2 = 0;
        for (int k2 = epoint[e]; k2 < epoint[e + 1]; ++k2, ++s2) {
            const int f = ef[k2];
            if (f < 0) continue;
            const int e1 = fc[2 * f], e2 = fc[2 * f + 1];
            row[1 + s2] = make_int2(e == e1 ? 3 : 2, f);
            nbr[(g * Fm + s2) * kGroup + lane] = inv[e == e1 ? e2 : e1];
        }
    }
}

// One thread per (group, slot): the slot's neighbour map is the shift c when
// every lane with a face there has nbr = m + c, kSlotGather when they differ,
// kSlotNone when no lane has the face.  Also the group table.
__global__ void tab_slot_kernel(int ngroups, int F, int Fm, int nle, const int *__restrict__ nbr,
                                int *__restrict__ slotc, int2 *__restrict__ gtab) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < (long)ngroups * Fm;
         t += (long)gridDim.x * blockDim.x) {
        const int g = (int)(t / Fm), f = (int)(t % Fm);
        int c = kSlotNone;
        for (int lane = 0; lane < kGroup && f < F; ++lane) {
            const int mm = g * kGroup + lane;
            const int nb = nbr[((long)g * Fm + f) * kGroup + lane];
            if (mm >= nle || nb < 0) continue;
            if (c == kSlotNone) c = nb - mm;
            else if (c != nb - mm) { c = kSlotGather; break; }
        }
        slotc[t] = c;
        if (f == 0) gtab[g] = make_int2(0, g);
    }
}

// Vector gathers between reference order (element id) and internal order.
__global__ void gather_vec_kernel(long total, int nv, const int *__restrict__ lines,
                                  const double *__restrict__ src, double *__restrict__ dst) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const long m = t / nv;
        const int i = (int)(t % nv);
        dst[t] = src ? src[(size_t)nv * lines[m] + i] : 0.0;
    }
}
__global__ void scatter_vec_kernel(long total, int nv, const int *__restrict__ lines,
                                   const double *__restrict__ src, double *__restrict__ dst) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const long m = t / nv;
        const int i = (int)(t % nv);
        dst[(size_t)nv * lines[m] + i] = src[t];
    }
}
__global__ void zero_kernel(long n, double *__restrict__ p) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        p[t] = 0.0;
}

unsigned gs_grid(long n) {
    long g = (n + 255) / 256;
    return (unsigned)std::max(1L, std::min(g, 65536L));
}

}  // namespace

// ======================================================================
// Host side
// ======================================================================

struct tbs_plan {
    tbs_ctx *ctx = nullptr;
    int nv = 9, prec = TBS_FP64;
    int nelem = 0, nintface = 0, nline = 0, nle = 0, F = 0, S = 1, ngroups = 0, nwin = 0;
    bool factored = false;
    std::vector<int> h_linesize, h_linepoint, h_woff;
    // device: topology
    int *d_lines = nullptr, *d_mline = nullptr, *d_linepoint = nullptr, *d_linesize = nullptr;
    int *d_woff = nullptr;
    int *d_fc = nullptr, *d_lineface = nullptr;  // for lineLU / assemble
    int2 *d_src = nullptr;                       // packing source codes
    int4 *d_fidx = nullptr;                      // factor block sources per element
    // device: data
    double *d_blk = nullptr;
    int *d_nbr = nullptr;
    int *d_slotc = nullptr;
    void *d_W = nullptr, *d_EF = nullptr, *d_GB = nullptr;
    double *d_rhs = nullptr, *d_U = nullptr, *d_R = nullptr, *d_gpart = nullptr, *d_hist = nullptr;
    double *a_rhs = nullptr, *a_U = nullptr, *a_R = nullptr;   // allocations (kPad before element 0)
    void *d_g = nullptr, *d_dUp = nullptr;
    int *d_state = nullptr;
    // the plan as a one-system batch
    void *d_sys = nullptr;
    int4 *d_ltab = nullptr;
    int2 *d_gtab = nullptr;
    std::map<int, hipGraphExec_t> graphs;
    int max_sweeps = kHistCap - 1;
    double rtol = 0.0;
    int solves = 0;           // tbs_plan_solve calls so far
    int variant = 0;          // Thomas prefetch variant (autotuned)
    size_t bytes = 0;
    int *h_state = nullptr;      // pinned unless region-backed
    bool in_region = false;      // device arrays carved from a context region
    std::vector<void *> allocs;  // hipMalloc'd arrays (not region-backed)
    std::unique_ptr<char[]> h_tables;  // host image of the packed tables (source of an async upload)
};

struct tbs_batch {
    tbs_ctx *ctx = nullptr;
    std::vector<tbs_plan *> plans;
    int nv = 9, prec = TBS_FP64;
    void *d_sys = nullptr;
    int4 *d_ltab = nullptr;
    int2 *d_gtab = nullptr;
    int nl = 0, ng = 0;
    // eager FP64 sweeps run in two halves on the context's two streams: systems
    // [0, nsA) with lines [0, nlA) and groups [0, ngA) of the tables, the
    // rest after them (nsA = 0: one stream)
    int nsA = 0, nlA = 0, ngA = 0;
    std::map<std::pair<int, std::pair<int, double>>, hipGraphExec_t> graphs;
    int *h_states = nullptr;    // pinned [cap_sys][4]
    int *d_states = nullptr;    // [cap_sys][4] gathered system states
    int variant = 0;
    // capacities of the device tables (entries) and of the pinned staging
    size_t cap_sys = 0, cap_l = 0, cap_g = 0, cap_stage = 0;
    char *h_stage = nullptr;    // pinned image of the tables being uploaded
    hipEvent_t staged = nullptr; // the last upload from h_stage has completed
};

namespace {

template <typename TF>
SysDev<TF> make_sysdev(const tbs_plan *p) {
    SysDev<TF> s{};
    s.blk = p->d_blk;
    s.nbr = p->d_nbr;
    s.slotc = p->d_slotc;
    s.W = (const TF *)p->d_W;
    s.EF = (const TF *)p->d_EF;
    s.GB = (const TF *)p->d_GB;
    s.rhs = p->d_rhs;
    s.U = p->d_U;
    s.g = (TF *)p->d_g;
    s.dUp = (TF *)p->d_dUp;
    s.R = p->d_R;
    s.gpart = p->d_gpart;
    s.hist = p->d_hist;
    s.state = p->d_state;
    s.n = p->nle;
    s.ngroups = p->ngroups;
    s.F = p->F;
    return s;
}

size_t tf_size(int prec) { return prec == TBS_FP32 ? sizeof(float) : sizeof(double); }

// Thomas factor-window stride in values (tbs::WinLayout::STRIDE for this nvar/precision).
size_t win_stride(int nv, int prec) {
    const size_t tf = tf_size(prec), lw = (size_t)(64 / nv) * nv;
    return (nv * lw * tf + 15) / 16 * 16 / tf;
}

int check_mesh(const tbs_mesh_host *m) {
    TBS_REQUIRE(m != nullptr, TBS_ERR_ARG, "null mesh");
    TBS_REQUIRE(m->nelem >= 1 && m->nline >= 1 && m->nintface >= 0 && m->eftot >= 0, TBS_ERR_ARG,
                "bad mesh sizes");
    TBS_REQUIRE(m->epoint && m->ef && m->linesize && m->linepoint && m->lines && m->lineface,
                TBS_ERR_ARG, "null mesh array");
    TBS_REQUIRE(m->nintface == 0 || m->fc, TBS_ERR_ARG, "null fc");
    TBS_REQUIRE(m->nlineelem == m->nelem, TBS_ERR_ARG,
                "the fused path needs every element in exactly one line (nlineelem %d != nelem %d)",
                m->nlineelem, m->nelem);
    TBS_REQUIRE(m->linepoint[0] == 0, TBS_ERR_ARG, "linepoint[0] must be 0");
    for (int l = 0; l < m->nline; ++l)
        TBS_REQUIRE(m->linesize[l] >= 1 && m->linepoint[l + 1] == m->linepoint[l] + m->linesize[l],
                    TBS_ERR_ARG, "line %d: linepoint/linesize not contiguous", l);
    TBS_REQUIRE(m->linepoint[m->nline] == m->nlineelem, TBS_ERR_ARG, "linepoint[nline] != nlineelem");
    std::vector<char> seen(m->nelem, 0);
    for (int i = 0; i < m->nlineelem; ++i) {
        const int e = m->lines[i];
        TBS_REQUIRE(e >= 0 && e < m->nelem && !seen[e], TBS_ERR_ARG, "lines[] is not a permutation");
        seen[e] = 1;
    }
    TBS_REQUIRE(m->epoint[0] == 0 && m->epoint[m->nelem] == m->eftot, TBS_ERR_ARG, "bad epoint");
    for (int e = 0; e < m->nelem; ++e)
        TBS_REQUIRE(m->epoint[e + 1] >= m->epoint[e], TBS_ERR_ARG, "epoint not monotone");
    for (int k = 0; k < m->eftot; ++k)
        TBS_REQUIRE(m->ef[k] < m->nintface, TBS_ERR_ARG, "ef[%d] out of range", k);
    for (int f = 0; f < 2 * m->nintface; ++f)
        TBS_REQUIRE(m->fc[f] >= 0 && m->fc[f] < m->nelem, TBS_ERR_ARG, "fc[%d] out of range", f);
    for (int l = 0; l < m->nline; ++l)
        for (int k = 1; k < m->linesize[l]; ++k) {
            const int mm = m->linepoint[l] + k;
            const int f = m->lineface[mm];
            TBS_REQUIRE(f >= 0 && f < m->nintface, TBS_ERR_ARG, "lineface[%d] not an interior face", mm);
            const int e = m->lines[mm], ep = m->lines[mm - 1];
            const int a = m->fc[2 * f], b = m->fc[2 * f + 1];
            TBS_REQUIRE((a == e && b == ep) || (a == ep && b == e), TBS_ERR_ARG,
                        "lineface[%d] does not join consecutive line elements", mm);
        }
    return TBS_OK;
}

struct Layout {
    int F, S, ngroups, nwin;
    std::vector<int> woff;
};

Layout layout_of(const tbs_mesh_host *m) {
    Layout L;
    int F = 0;
    for (int e = 0; e < m->nelem; ++e) F = std::max(F, m->epoint[e + 1] - m->epoint[e]);
    L.F = F;
    L.S = F + 1;
    L.ngroups = (m->nlineelem + kGroup - 1) / kGroup;
    const int spw = 64 / m->nvar;
    L.woff.resize(m->nline + 1);
    int w = 0;
    for (int l = 0; l < m->nline; ++l) {
        L.woff[l] = w;
        w += line_windows(m->linesize[l], spw);
    }
    L.woff[m->nline] = w;
    L.nwin = w;
    return L;
}

// Packed integer tables of a plan (one device array): byte offsets, each
// 16-byte aligned, and the total.
// The arrays up to host_end come from the host (mesh arrays as given, the
// window offsets and the system descriptor); the rest are built on the device.
struct TableLayout {
    size_t lines, linepoint, linesize, woff, fc, lineface, epoint, ef, sys, host_end;
    size_t mline, src, fidx, nbr, slotc, ltab, gtab, total;
};

TableLayout table_layout(const tbs_mesh_host *m, const Layout &L, int prec) {
    TableLayout t{};
    size_t o = 0;
    auto take = [&o](size_t bytes) { const size_t at = o; o += (bytes + 15) / 16 * 16; return at; };
    const size_t G = (size_t)L.ngroups * kGroup, Fm = (size_t)std::max(L.F, 1);
    const size_t nle = m->nlineelem, nl = m->nline;
    t.lines = take(4 * nle);
    t.linepoint = take(4 * (nl + 1));
    t.linesize = take(4 * nl);
    t.woff = take(4 * (nl + 1));
    t.fc = take(4 * std::max<size_t>(2 * (size_t)m->nintface, 1));
    t.lineface = take(4 * nle);
    t.epoint = take(4 * ((size_t)m->nelem + 1));
    t.ef = take(4 * std::max<size_t>((size_t)m->eftot, 1));
    t.sys = take(prec == TBS_FP32 ? sizeof(SysDev<float>) : sizeof(SysDev<double>));
    t.host_end = o;
    t.mline = take(4 * nle);
    t.src = take(sizeof(int2) * G * L.S);
    t.fidx = take(sizeof(int4) * nle);
    t.nbr = take(4 * G * Fm);
    t.slotc = take(4 * (size_t)L.ngroups * Fm);
    t.ltab = take(sizeof(int4) * nl);
    t.gtab = take(sizeof(int2) * L.ngroups);
    t.total = o;
    return t;
}

// Device arrays of a plan, in allocation order.
enum { A_BLK, A_W, A_EF, A_GB, A_RHS, A_U, A_R, A_G, A_DUP, A_GPART, A_HIST, A_STATE, A_TABLES, A_COUNT };

std::array<size_t, A_COUNT> array_sizes(const tbs_mesh_host *m, const Layout &L, int prec) {
    const size_t nv = m->nvar, bs = nv * nv, tf = tf_size(prec);
    const size_t G = (size_t)L.ngroups * kGroup, nle = m->nlineelem;
    std::array<size_t, A_COUNT> a{};
    a[A_BLK] = G * L.S * bs * 8;
    a[A_W] = G * bs * tf;
    a[A_EF] = a[A_GB] = (size_t)L.nwin * win_stride((int)nv, prec) * tf;
    a[A_RHS] = a[A_U] = a[A_R] = (nle + 2 * (size_t)kPad) * nv * 8;
    a[A_G] = a[A_DUP] = nle * nv * tf;
    a[A_GPART] = (size_t)L.ngroups * 8;
    a[A_HIST] = (size_t)kHistCap * 8;
    a[A_STATE] = 4 * sizeof(int);
    a[A_TABLES] = table_layout(m, L, prec).total;
    return a;
}

size_t aligned_bytes(size_t b) { return (std::max<size_t>(b, 1) + tbs::kRegionAlign - 1) / tbs::kRegionAlign * tbs::kRegionAlign; }

// Device bytes a plan holds (each array rounded to the region alignment).
size_t footprint_of(const tbs_mesh_host *m, const Layout &L, int prec) {
    size_t b = 0;
    for (size_t x : array_sizes(m, L, prec)) b += aligned_bytes(x);
    return b;
}

// Bytes of scratch the factorisation needs beyond the plan (none: the fused
// factor kernel writes the sweep-form factors directly).
size_t factor_workspace_bytes(int, int) { return 0; }

// Algorithmic bytes of one sweep, BASELINE model: factor-side D', A, DinvC
// (tf bytes each), D + Of1 + Of2 at FP64, 6 vector passes at FP64.
double model_bytes(const tbs_mesh_host *m, int prec) {
    const double bs = (double)m->nvar * m->nvar, nv = m->nvar;
    const double tf = (double)tf_size(prec);
    return tf * bs * (3.0 * m->nelem - 2.0 * m->nline) + 8.0 * bs * (m->nelem + 2.0 * m->nintface) +
           6.0 * 8.0 * nv * m->nelem;
}

// The reference's own per-sweep byte model, src/main.cxx:149-175 / iters.
double reference_model_bytes(const tbs_mesh_host *m) {
    const double nv = m->nvar, bs = nv * nv;
    const double du = 8.0 * (bs * m->nelem + bs * m->nelem + bs * m->nlineelem + 4.0 * nv * m->nelem +
                             nv * m->nelem) +
                      4.0 * (m->nline + (m->nline + 1.0) + 2.0 * m->nlineelem);
    const double cp = 5.0 * nv * m->nelem * 8.0;
    const double lr = 8.0 * (bs * m->nelem + 2.0 * bs * m->nintface + 5.0 * m->nelem * nv +
                             2.0 * nv * m->nelem) +
                      4.0 * ((m->nelem + 1.0) + m->eftot + 2.0 * m->nintface + m->nline +
                             (m->nline + 1.0) + 2.0 * m->nlineelem);
    return du + cp + lr;
}

int plan_alloc(tbs_plan *p, size_t bytes, void **ptr) {
    bool reg = false;
    TBS_TRY(tbs::dev_alloc(p->ctx, bytes, ptr, &reg));
    if (reg) p->in_region = true;
    else p->allocs.push_back(*ptr);
    p->bytes += aligned_bytes(bytes);
    return TBS_OK;
}

void free_plan(tbs_plan *p) {
    if (!p) return;
    hipSetDevice(p->ctx->device);
    for (auto &kv : p->graphs) hipGraphExecDestroy(kv.second);
    for (void *q : p->allocs) hipFree(q);
    if (p->h_state) {
        if (p->in_region) delete[] p->h_state;
        else hipHostFree(p->h_state);
    }
    delete p;
}

// Launch on st; with a start event, through hipExtLaunchKernelGGL, which
// records the two events at the kernel's own start and end (timing of exactly
// this kernel, without the dispatch or the completion signal around it).
template <typename K, typename... Args>
void launch_k(K kernel, dim3 grid, dim3 block, hipStream_t st, hipEvent_t e0, hipEvent_t e1, Args... args) {
    if (e0) hipExtLaunchKernelGGL(kernel, grid, block, 0, st, e0, e1, 0, args...);
    else hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
}

template <int NV, typename TF>
int launch_residual(hipStream_t st, const void *sys, const int2 *gtab, int ng, int want_R, int force,
                    hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    if (ng == 0) return TBS_OK;
    const int nblk = (ng + 3) / 4;
    launch_k(fres_kernel<NV, TF>, dim3(8 * ((nblk + 7) / 8)), dim3(256), st, e0, e1,
             (const SysDev<TF> *)sys, gtab, ng, want_R, force);
    TBS_LAUNCHED("fres_kernel");
    return TBS_OK;
}

// Thomas variants the autotuner picks from: {prefetch windows PF, ring
// cycles K per loop iteration}; index 0 is the default.
struct TV { int pf, k; };
template <int NV, typename TF> struct ThomasV { static constexpr int n = 1; static constexpr TV v[1] = {{2, 2}}; };
// (LDS per two-wave workgroup stays <= 32 KiB so five lines fit a CU.)
template <> struct ThomasV<9, double> {
    static constexpr int n = 4;
    static constexpr TV v[4] = {{2, 2}, {3, 1}, {1, 3}, {1, 2}};
};
template <> struct ThomasV<9, float> {
    static constexpr int n = 4;
    static constexpr TV v[4] = {{3, 2}, {2, 2}, {5, 1}, {3, 1}};
};

template <int NV, typename TF, int PF, int K>
int launch_thomas_pf(hipStream_t st, const void *sys, const int4 *ltab, int nl, hipEvent_t e0, hipEvent_t e1) {
    launch_k(fthomas_kernel<NV, TF, PF, K>, dim3(nl), dim3(128), st, e0, e1, (const SysDev<TF> *)sys, ltab, nl);
    TBS_LAUNCHED("fthomas_kernel");
    return TBS_OK;
}

template <int NV, typename TF>
int launch_thomas(hipStream_t st, const void *sys, const int4 *ltab, int nl, int variant, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr) {
    if (nl == 0) return TBS_OK;
    using P = ThomasV<NV, TF>;
    if (variant < 0 || variant >= P::n) variant = 0;
    constexpr TV a = P::v[0];
    constexpr TV b = P::v[P::n > 1 ? 1 : 0];
    constexpr TV c = P::v[P::n > 2 ? 2 : 0];
    constexpr TV d = P::v[P::n > 3 ? 3 : 0];
    switch (variant) {
        case 1: return launch_thomas_pf<NV, TF, b.pf, b.k>(st, sys, ltab, nl, e0, e1);
        case 2: return launch_thomas_pf<NV, TF, c.pf, c.k>(st, sys, ltab, nl, e0, e1);
        case 3: return launch_thomas_pf<NV, TF, d.pf, d.k>(st, sys, ltab, nl, e0, e1);
        default: return launch_thomas_pf<NV, TF, a.pf, a.k>(st, sys, ltab, nl, e0, e1);
    }
}

int thomas_variants(int nv, int prec) {
    if (nv == 9) return prec == TBS_FP32 ? ThomasV<9, float>::n : ThomasV<9, double>::n;
    return 1;
}

// encoded as 100 * PF + K
int thomas_pf_of(int nv, int prec, int variant) {
    TV t = {2, 2};
    if (nv == 9) t = prec == TBS_FP32 ? ThomasV<9, float>::v[variant % ThomasV<9, float>::n]
                                      : ThomasV<9, double>::v[variant % ThomasV<9, double>::n];
    return 100 * t.pf + t.k;
}

template <typename TF>
int launch_final(hipStream_t st, const void *sys, int nsys, int max_sweeps, double rtol, int force,
                 hipEvent_t e0, hipEvent_t e1) {
    launch_k(ffinal_kernel<TF>, dim3(nsys), dim3(kFinalThreads), st, e0, e1, (const SysDev<TF> *)sys, nsys,
             max_sweeps, rtol, force);
    TBS_LAUNCHED("ffinal_kernel");
    return TBS_OK;
}

// Dispatch on (nvar, precision) for the fused kernels.
#define TBS_PLAN_NVARS(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

FactorRec factor_rec(const tbs_plan *p, const double *D, const double *O1, const double *O2) {
    FactorRec r{};
    r.linepoint = p->d_linepoint;
    r.linesize = p->d_linesize;
    r.woff = p->d_woff;
    r.fidx = p->d_fidx;
    r.D = D;
    r.O1 = O1;
    r.O2 = O2;
    r.zero = p->a_rhs;   // the rhs allocation starts with kPad * nv >= nv * nv zeros
    r.W = p->d_W;
    r.EF = p->d_EF;
    r.GB = p->d_GB;
    return r;
}

int launch_factor(int nv, int prec, hipStream_t st, const FactorRec &one, int nline, const FactorRec *recs,
                  const int2 *ltab) {
    if (nline == 0) return TBS_OK;
    switch (nv) {
#define X(v) case v: \
        if (prec == TBS_FP32) hipLaunchKernelGGL((ffactor_kernel<v, float>), dim3(nline), dim3(128), 0, st, one, \
                                                 nline, recs, ltab); \
        else hipLaunchKernelGGL((ffactor_kernel<v, double>), dim3(nline), dim3(128), 0, st, one, nline, recs, ltab); \
        break;
        TBS_PLAN_NVARS(X)
#undef X
    }
    TBS_LAUNCHED("ffactor_kernel");
    return TBS_OK;
}

int launch_pack(const tbs_plan *p, hipStream_t pst, const double *jacD, const double *jacO1, const double *jacO2) {
    const int nv = p->nv, BS = nv * nv;
    const long nws = (long)p->ngroups * p->S;
    if (2 * nws < INT_MAX) {
        switch (nv) {
#define X(v) case v: hipLaunchKernelGGL(pack_lds_kernel<v>, dim3((unsigned)(2 * nws)), dim3(128), 0, pst, \
                   p->S, p->d_src, jacD, jacO1, jacO2, p->d_blk); break;
            TBS_PLAN_NVARS(X)
#undef X
        }
        TBS_LAUNCHED("pack_lds_kernel");
    } else {
        const long total = (long)p->ngroups * kGroup * p->S * BS;
        switch (nv) {
#define X(v) case v: hipLaunchKernelGGL(pack_blocks_kernel<v>, dim3(gs_grid(total)), dim3(256), 0, pst, \
                   total, p->S, p->d_src, jacD, jacO1, jacO2, p->d_blk); break;
            TBS_PLAN_NVARS(X)
#undef X
        }
        TBS_LAUNCHED("pack_blocks_kernel");
    }
    return TBS_OK;
}

int k_residual(int nv, int prec, hipStream_t st, const void *sys, const int2 *gtab, int ng,
               int want_R, int force, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    switch (nv) {
#define X(v) case v: return prec == TBS_FP32 ? launch_residual<v, float>(st, sys, gtab, ng, want_R, force, e0, e1) \
                                         : launch_residual<v, double>(st, sys, gtab, ng, want_R, force, e0, e1);
        TBS_PLAN_NVARS(X)
#undef X
    }
    tbs::set_error("nvar %d not built for the fused path", nv);
    return TBS_ERR_UNSUPPORTED;
}
int k_thomas(int nv, int prec, hipStream_t st, const void *sys, const int4 *ltab, int nl, int variant,
             hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    switch (nv) {
#define X(v) case v: return prec == TBS_FP32 ? launch_thomas<v, float>(st, sys, ltab, nl, variant, e0, e1) \
                                         : launch_thomas<v, double>(st, sys, ltab, nl, variant, e0, e1);
        TBS_PLAN_NVARS(X)
#undef X
    }
    tbs::set_error("nvar %d not built for the fused path", nv);
    return TBS_ERR_UNSUPPORTED;
}
int k_final(int prec, hipStream_t st, const void *sys, int nsys, int max_sweeps, double rtol, int force,
            hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    return prec == TBS_FP32 ? launch_final<float>(st, sys, nsys, max_sweeps, rtol, force, e0, e1)
                            : launch_final<double>(st, sys, nsys, max_sweeps, rtol, force, e0, e1);
}
int k_reset(int prec, hipStream_t st, const void *sys, int nsys) {
    if (prec == TBS_FP32)
        hipLaunchKernelGGL((reset_state_kernel<float>), dim3((nsys + 63) / 64), dim3(64), 0, st,
                           (const SysDev<float> *)sys, nsys);
    else
        hipLaunchKernelGGL((reset_state_kernel<double>), dim3((nsys + 63) / 64), dim3(64), 0, st,
                           (const SysDev<double> *)sys, nsys);
    TBS_LAUNCHED("reset_state_kernel");
    return TBS_OK;
}

bool plan_nvar_ok(int nv) {
    switch (nv) {
#define X(v) case v: return true;
        TBS_PLAN_NVARS(X)
#undef X
    }
    return false;
}

// One sweep: Thomas (+U update), residual, norm/convergence.
int batch_sweeps(tbs_batch *b, hipStream_t st, int nsweeps, int max_sweeps, double rtol, bool split);

// One sweep of the systems the tables ltab[0, nl) and gtab[0, ng) cover
// (they index sys by system); the norm step runs over sys_final[0, nsys).
int enqueue_sweep(int nv, int prec, hipStream_t st, const void *sys, const void *sys_final, int nsys,
                  const int4 *ltab, int nl, const int2 *gtab, int ng, int max_sweeps, double rtol, int variant) {
    TBS_TRY(k_thomas(nv, prec, st, sys, ltab, nl, variant));
    TBS_TRY(k_residual(nv, prec, st, sys, gtab, ng, 0, 0));
    TBS_TRY(k_final(prec, st, sys_final, nsys, max_sweeps, rtol, 0));
    return TBS_OK;
}

int capture_sweeps(hipStream_t st, int nsweeps, hipGraphExec_t *out,
                   const std::function<int()> &one_sweep) {
    TBS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = TBS_OK;
    for (int i = 0; i < nsweeps && rc == TBS_OK; ++i) rc = one_sweep();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) { if (graph) hipGraphDestroy(graph); return rc; }
    TBS_HIP(e);
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    TBS_HIP(e);
    return TBS_OK;
}

}  // namespace

namespace tbs {

int plan_scatter_U(tbs_plan *p, hipStream_t st, double *U) {
    const long total = (long)p->nle * p->nv;
    hipLaunchKernelGGL(scatter_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv, p->d_lines,
                       p->d_U, U);
    TBS_LAUNCHED("scatter U");
    return TBS_OK;
}

const double *plan_hist(const tbs_plan *p) { return p->d_hist; }

int plan_vecs(tbs_plan *p, PlanVecs *out) {
    TBS_REQUIRE(p && out && p->factored, TBS_ERR_STATE, "plan not factored");
    out->ctx = p->ctx;
    out->rhs = p->d_rhs;
    out->U = p->d_U;
    out->R = p->d_R;
    out->n = p->nle;
    out->nv = p->nv;
    out->lines = p->d_lines;
    return TBS_OK;
}

int plan_activate(tbs_plan *p, hipStream_t st) { return k_reset(p->prec, st, p->d_sys, 1); }

int plan_residual(tbs_plan *p, hipStream_t st, int want_R, bool u_zero) {
    return k_residual(p->nv, p->prec, st, p->d_sys, p->d_gtab, p->ngroups, want_R, u_zero ? 3 : 1);
}

int plan_thomas(tbs_plan *p, hipStream_t st) {
    return k_thomas(p->nv, p->prec, st, p->d_sys, p->d_ltab, p->nline, p->variant);
}

int plan_gather(tbs_plan *p, hipStream_t st, const double *src, double *dst) {
    const long total = (long)p->nle * p->nv;
    hipLaunchKernelGGL(gather_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv, p->d_lines, src,
                       dst);
    TBS_LAUNCHED("gather");
    return TBS_OK;
}

int plan_scatter(tbs_plan *p, hipStream_t st, const double *src, double *dst) {
    const long total = (long)p->nle * p->nv;
    hipLaunchKernelGGL(scatter_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv, p->d_lines,
                       src, dst);
    TBS_LAUNCHED("scatter");
    return TBS_OK;
}

size_t plan_factor_workspace(int nvar, int nelem) { return factor_workspace_bytes(nvar, nelem); }

}  // namespace tbs

extern "C" {

int tbs_plan_supported_nvar(int *out, int cap) {
    int n = 0;
#define X(v) if (out && n < cap) out[n] = v; ++n;
    TBS_PLAN_NVARS(X)
#undef X
    return n;
}

int tbs_plan_footprint(const tbs_mesh_host *mesh, int precision, size_t *bytes) {
    return tbs::plan_footprint(mesh, precision, bytes, true);
}

int tbs_plan_create(tbs_ctx *ctx, const tbs_mesh_host *m, int precision, tbs_plan **out) {
    return tbs::plan_create(ctx, m, precision, out, true);
}

}  // extern "C"

int tbs::plan_footprint(const tbs_mesh_host *mesh, int precision, size_t *bytes, bool validate) {
    TBS_REQUIRE(bytes, TBS_ERR_ARG, "null output");
    if (validate) TBS_TRY(check_mesh(mesh));
    else TBS_REQUIRE(mesh != nullptr && mesh->nelem >= 1 && mesh->nline >= 1, TBS_ERR_ARG, "bad mesh sizes");
    TBS_REQUIRE(precision == TBS_FP64 || precision == TBS_FP32, TBS_ERR_ARG, "bad precision");
    Layout L = layout_of(mesh);
    *bytes = footprint_of(mesh, L, precision);
    return TBS_OK;
}

// Plan creation; validate = false skips the O(mesh) consistency checks for
// meshes the library built itself (a served request's structured grid).
int tbs::plan_create(tbs_ctx *ctx, const tbs_mesh_host *m, int precision, tbs_plan **out, bool validate) {
    TBS_TRY(tbs::use_ctx(ctx));
    TBS_REQUIRE(out, TBS_ERR_ARG, "null output");
    *out = nullptr;
    if (validate) TBS_TRY(check_mesh(m));
    else TBS_REQUIRE(m != nullptr && m->nelem >= 1 && m->nline >= 1 && m->nlineelem == m->nelem, TBS_ERR_ARG,
                     "bad mesh sizes");
    TBS_REQUIRE(m->nvar == ctx->nvar, TBS_ERR_ARG, "mesh nvar %d != context nvar %d", m->nvar, ctx->nvar);
    TBS_REQUIRE(plan_nvar_ok(m->nvar), TBS_ERR_UNSUPPORTED, "nvar %d not built for the fused path", m->nvar);
    TBS_REQUIRE(precision == TBS_FP64 || precision == TBS_FP32, TBS_ERR_ARG, "bad precision");
    // whole-system vectors are addressed through buffer descriptors with a
    // 32-bit byte range
    TBS_REQUIRE(((size_t)m->nlineelem + 2 * kPad) * m->nvar * 8 < ((size_t)1 << 32) - 4096, TBS_ERR_UNSUPPORTED,
                "system too large for one plan: %d elements x %d values exceeds 4 GiB per vector",
                m->nlineelem, m->nvar);
    const int nv = m->nvar;
    Layout L = layout_of(m);
    std::unique_ptr<tbs_plan, void (*)(tbs_plan *)> p(new tbs_plan(), free_plan);
    p->ctx = ctx;
    p->nv = nv;
    p->prec = precision;
    p->nelem = m->nelem;
    p->nintface = m->nintface;
    p->nline = m->nline;
    p->nle = m->nlineelem;
    p->F = L.F;
    p->S = L.S;
    p->ngroups = L.ngroups;
    p->nwin = L.nwin;
    p->h_linesize.assign(m->linesize, m->linesize + m->nline);
    p->h_linepoint.assign(m->linepoint, m->linepoint + m->nline + 1);
    p->h_woff = L.woff;
    const size_t G = (size_t)L.ngroups * kGroup;

    tbs_plan *P = p.get();
    const std::array<size_t, A_COUNT> sz = array_sizes(m, L, precision);
    void *arr[A_COUNT];
    for (int i = 0; i < A_COUNT; ++i) TBS_TRY(plan_alloc(P, sz[i], &arr[i]));
    P->d_blk = (double *)arr[A_BLK];
    P->d_W = arr[A_W];
    P->d_EF = arr[A_EF];
    P->d_GB = arr[A_GB];
    P->a_rhs = (double *)arr[A_RHS];
    P->a_U = (double *)arr[A_U];
    P->a_R = (double *)arr[A_R];
    P->d_rhs = P->a_rhs + (size_t)kPad * nv;
    P->d_U = P->a_U + (size_t)kPad * nv;
    P->d_R = P->a_R + (size_t)kPad * nv;
    P->d_g = arr[A_G];
    P->d_dUp = arr[A_DUP];
    P->d_gpart = (double *)arr[A_GPART];
    P->d_hist = (double *)arr[A_HIST];
    P->d_state = (int *)arr[A_STATE];
    char *tb = (char *)arr[A_TABLES];
    const TableLayout T = table_layout(m, L, precision);
    P->d_lines = (int *)(tb + T.lines);
    P->d_mline = (int *)(tb + T.mline);
    P->d_linepoint = (int *)(tb + T.linepoint);
    P->d_linesize = (int *)(tb + T.linesize);
    P->d_woff = (int *)(tb + T.woff);
    P->d_fc = (int *)(tb + T.fc);
    P->d_lineface = (int *)(tb + T.lineface);
    P->d_src = (int2 *)(tb + T.src);
    P->d_fidx = (int4 *)(tb + T.fidx);
    P->d_nbr = (int *)(tb + T.nbr);
    P->d_slotc = (int *)(tb + T.slotc);
    P->d_ltab = (int4 *)(tb + T.ltab);
    P->d_gtab = (int2 *)(tb + T.gtab);
    P->d_sys = tb + T.sys;
    if (P->in_region) P->h_state = new int[4]();
    else TBS_HIP(hipHostMalloc((void **)&P->h_state, 4 * sizeof(int), hipHostMallocDefault));

    // host part of the tables (the mesh arrays as given): small plans (a
    // served request's) go through the context's pinned staging buffer,
    // reused from plan to plan, so the upload is one DMA; large ones through
    // a pageable image kept by the plan so the upload stays asynchronous
    constexpr size_t kStageMax = (size_t)64 << 20;
    const bool staged = T.host_end <= kStageMax;
    char *h = nullptr;
    if (staged) {
        TBS_TRY(tbs::stage_acquire(ctx, T.host_end, &h));
    } else {
        P->h_tables.reset(new char[T.host_end]);
        h = P->h_tables.get();
    }
    const int nle = P->nle, F = L.F, Fm = std::max(F, 1);
    auto put = [h](size_t off, const void *src_, size_t bytes) { if (bytes) std::memcpy(h + off, src_, bytes); };
    put(T.lines, m->lines, (size_t)nle * 4);
    put(T.linepoint, m->linepoint, (size_t)(P->nline + 1) * 4);
    put(T.linesize, m->linesize, (size_t)P->nline * 4);
    put(T.woff, L.woff.data(), (size_t)(P->nline + 1) * 4);
    put(T.fc, m->fc, (size_t)2 * m->nintface * 4);
    put(T.lineface, m->lineface, (size_t)nle * 4);
    put(T.epoint, m->epoint, (size_t)(P->nelem + 1) * 4);
    put(T.ef, m->ef, (size_t)m->eftot * 4);
    if (precision == TBS_FP32) {
        SysDev<float> sd = make_sysdev<float>(P);
        put(T.sys, &sd, sizeof(sd));
    } else {
        SysDev<double> sd = make_sysdev<double>(P);
        put(T.sys, &sd, sizeof(sd));
    }
    hipStream_t st = ctx->stream;
    TBS_HIP(hipMemcpyAsync(tb, h, T.host_end, hipMemcpyHostToDevice, st));
    if (staged) TBS_TRY(tbs::stage_release(ctx));
    // derived tables on the device; the R vector (zeroed below) holds the
    // inverse permutation meanwhile
    {
        const int *d_epoint = (const int *)(tb + T.epoint), *d_ef = (const int *)(tb + T.ef);
        int *inv = (int *)P->a_R;
        hipLaunchKernelGGL(tab_lines_kernel, dim3(P->nline), dim3(64), 0, st, P->d_linepoint, P->d_linesize,
                           P->d_woff, P->d_lines, P->d_lineface, P->d_fc, P->d_mline, P->d_fidx, P->d_ltab);
        TBS_LAUNCHED("tab_lines_kernel");
        hipLaunchKernelGGL(tab_inv_kernel, dim3(gs_grid(nle)), dim3(256), 0, st, nle, P->d_lines, inv);
        TBS_LAUNCHED("tab_inv_kernel");
        hipLaunchKernelGGL(tab_src_kernel, dim3(gs_grid((long)G)), dim3(256), 0, st, (long)G, nle, L.S, Fm,
                           P->d_lines, d_epoint, d_ef, P->d_fc, inv, P->d_src, P->d_nbr);
        TBS_LAUNCHED("tab_src_kernel");
        hipLaunchKernelGGL(tab_slot_kernel, dim3(gs_grid((long)L.ngroups * Fm)), dim3(256), 0, st, L.ngroups, F,
                           Fm, nle, P->d_nbr, P->d_slotc, P->d_gtab);
        TBS_LAUNCHED("tab_slot_kernel");
    }
    for (double *a : {P->a_rhs, P->a_U, P->a_R}) TBS_HIP(hipMemsetAsync(a, 0, sz[A_RHS], st));
    *out = p.release();
    return TBS_OK;
}

extern "C" {

int tbs_plan_destroy(tbs_plan *plan) {
    if (!plan) return TBS_OK;
    hipSetDevice(plan->ctx->device);
    // a region-backed plan frees no device memory; its owner retires it only
    // after the plan's work has completed
    if (!plan->in_region) hipStreamSynchronize(plan->ctx->stream);
    free_plan(plan);
    return TBS_OK;
}

int tbs_plan_info_get(tbs_plan *p, tbs_plan_info *info) {
    TBS_REQUIRE(p && info, TBS_ERR_ARG, "null argument");
    tbs_mesh_host mh{};
    mh.nvar = p->nv;
    mh.nelem = p->nelem;
    mh.nintface = p->nintface;
    mh.nline = p->nline;
    mh.nlineelem = p->nle;
    mh.eftot = 0;
    info->nvar = p->nv;
    info->nelem = p->nelem;
    info->nline = p->nline;
    info->max_faces = p->F;
    info->precision = p->prec;
    info->device_bytes = p->bytes;
    info->alg_bytes_per_sweep = model_bytes(&mh, p->prec);
    info->ref_bytes_per_sweep = 0.0;
    // per-kernel algorithmic bytes of the fused form:
    //   Thomas: E + Gamma factor blocks (tf), g in, d' out + in, U in + out
    //   residual: D + Of1 + Of2 (FP64), W (tf), U, rhs in, g out
    const double bs = (double)p->nv * p->nv, nv = p->nv, tf = (double)tf_size(p->prec);
    const double n = p->nle;
    info->thomas_bytes = tf * bs * 2.0 * (n - p->nline) + nv * n * (tf * 3.0 + 16.0);
    info->residual_bytes = 8.0 * bs * (n + 2.0 * p->nintface) + tf * bs * n + nv * n * (16.0 + tf);
    return TBS_OK;
}

int tbs_plan_factor(tbs_plan *p, const double *jacD, const double *jacO1, const double *jacO2) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    return tbs_plans_factor(p->ctx, &p, 1, &jacD, &jacO1, &jacO2);
}

int tbs_plans_factor(tbs_ctx *ctx, tbs_plan *const *plans, int n, const double *const *jacD,
                     const double *const *jacO1, const double *const *jacO2) {
    TBS_REQUIRE(ctx && plans && n >= 1 && jacD && jacO1 && jacO2, TBS_ERR_ARG, "bad arguments");
    TBS_TRY(tbs::use_ctx(ctx));
    const int nv = plans[0] ? plans[0]->nv : 0, prec = plans[0] ? plans[0]->prec : 0;
    long nl = 0;
    for (int i = 0; i < n; ++i) {
        const tbs_plan *p = plans[i];
        TBS_REQUIRE(p && p->ctx == ctx, TBS_ERR_ARG, "plan %d is null or on another context", i);
        TBS_REQUIRE(p->nv == nv && p->prec == prec, TBS_ERR_ARG, "plan %d: block size or precision differs", i);
        TBS_REQUIRE(jacD[i] && (p->nintface == 0 || (jacO1[i] && jacO2[i])), TBS_ERR_ARG,
                    "plan %d: null Jacobian array", i);
        nl += p->nline;
    }
    TBS_REQUIRE(nl < INT_MAX, TBS_ERR_ARG, "too many lines");
    hipStream_t st = ctx->stream;
    // sweep-form factors W, E, Gamma straight from the blocks: the factor
    // kernel writes every window entry the sweep reads; of W only the last
    // group's slot has lanes past the last element, zeroed here
    for (int i = 0; i < n; ++i) {
        const tbs_plan *p = plans[i];
        const size_t slot = (size_t)kGroup * nv * nv * tf_size(prec);
        TBS_HIP(hipMemsetAsync((char *)p->d_W + (size_t)(p->ngroups - 1) * slot, 0, slot, st));
    }
    hipStream_t pst;
    TBS_TRY(tbs::aux_stream(ctx, &pst));
    TBS_HIP(hipEventRecord(ctx->ev_fork, st));
    TBS_HIP(hipStreamWaitEvent(pst, ctx->ev_fork, 0));
    if (n == 1) {
        TBS_TRY(launch_factor(nv, prec, st, factor_rec(plans[0], jacD[0], jacO1[0], jacO2[0]), plans[0]->nline,
                              nullptr, nullptr));
    } else {
        // one launch over every plan's lines, longest first; the records and
        // the line table go up through the staging ring into a device buffer
        // owned by the context (stream-ordered reuse)
        std::vector<std::pair<int, int2>> order;
        order.reserve(nl);
        for (int i = 0; i < n; ++i)
            for (int l = 0; l < plans[i]->nline; ++l) order.push_back({plans[i]->h_linesize[l], make_int2(i, l)});
        std::stable_sort(order.begin(), order.end(),
                         [](const std::pair<int, int2> &a, const std::pair<int, int2> &b) { return a.first > b.first; });
        const size_t rec_bytes = (size_t)n * sizeof(FactorRec), lt_off = (rec_bytes + 15) / 16 * 16;
        const size_t bytes = lt_off + (size_t)nl * sizeof(int2);
        char *h = nullptr;
        TBS_TRY(tbs::stage_acquire(ctx, bytes, &h));
        FactorRec *rec = (FactorRec *)h;
        for (int i = 0; i < n; ++i) rec[i] = factor_rec(plans[i], jacD[i], jacO1[i], jacO2[i]);
        int2 *lt = (int2 *)(h + lt_off);
        for (long b = 0; b < nl; ++b) lt[b] = order[b].second;
        TBS_TRY(tbs::dev_stage(ctx, bytes));
        TBS_HIP(hipMemcpyAsync(ctx->dstage, h, bytes, hipMemcpyHostToDevice, st));
        TBS_TRY(tbs::stage_release(ctx));
        TBS_TRY(launch_factor(nv, prec, st, FactorRec{}, (int)nl, (const FactorRec *)ctx->dstage,
                              (const int2 *)(ctx->dstage + lt_off)));
    }
    // residual blocks, packed on the context's second stream while the
    // factor kernel (latency-bound, two waves per line) runs on the main one:
    // queued after the factor kernel (its waves take their CUs first) and
    // forked after the memsets; the main stream waits for the packing before
    // anything after the call
    for (int i = 0; i < n; ++i) TBS_TRY(launch_pack(plans[i], pst, jacD[i], jacO1[i], jacO2[i]));
    TBS_HIP(hipEventRecord(ctx->ev_join, pst));
    TBS_HIP(hipStreamWaitEvent(st, ctx->ev_join, 0));
    for (int i = 0; i < n; ++i) plans[i]->factored = true;
    return TBS_OK;
}

int tbs_plan_begin(tbs_plan *p, const double *rhs, const double *U0) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    TBS_TRY(tbs::use_ctx(p->ctx));
    TBS_REQUIRE(p->factored, TBS_ERR_STATE, "tbs_plan_begin before tbs_plan_factor");
    TBS_REQUIRE(rhs, TBS_ERR_ARG, "null rhs");
    hipStream_t st = p->ctx->stream;
    const long total = (long)p->nle * p->nv;
    hipLaunchKernelGGL(gather_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                       p->d_lines, rhs, p->d_rhs);
    TBS_LAUNCHED("gather rhs");
    hipLaunchKernelGGL(gather_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                       p->d_lines, U0, p->d_U);
    TBS_LAUNCHED("gather U0");
    TBS_TRY(k_reset(p->prec, st, p->d_sys, 1));
    // R_0 = rhs + J U_0 (= rhs when U_0 = 0, the reference's start; then J U
    // is skipped) and ||R_0||
    TBS_TRY(k_residual(p->nv, p->prec, st, p->d_sys, p->d_gtab, p->ngroups, 0, U0 ? 1 : 3));
    TBS_TRY(k_final(p->prec, st, p->d_sys, 1, p->max_sweeps, p->rtol, 1));
    return TBS_OK;
}

int tbs_plan_launch_thomas(tbs_plan *p) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    TBS_TRY(tbs::use_ctx(p->ctx));
    return k_thomas(p->nv, p->prec, p->ctx->stream, p->d_sys, p->d_ltab, p->nline, p->variant);
}

int tbs_plan_launch_residual(tbs_plan *p) {
    TBS_TRY(tbs_plan_launch_kernel(p, TBS_STEP_RESIDUAL, nullptr, nullptr));
    return tbs_plan_launch_kernel(p, TBS_STEP_NORM, nullptr, nullptr);
}

int tbs_plan_launch_kernel(tbs_plan *p, int step, void *ev_start, void *ev_stop) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    TBS_REQUIRE((ev_start == nullptr) == (ev_stop == nullptr), TBS_ERR_ARG, "give both events or neither");
    TBS_TRY(tbs::use_ctx(p->ctx));
    hipStream_t st = p->ctx->stream;
    hipEvent_t e0 = (hipEvent_t)ev_start, e1 = (hipEvent_t)ev_stop;
    switch (step) {
        case TBS_STEP_THOMAS: return k_thomas(p->nv, p->prec, st, p->d_sys, p->d_ltab, p->nline, p->variant, e0, e1);
        case TBS_STEP_RESIDUAL: return k_residual(p->nv, p->prec, st, p->d_sys, p->d_gtab, p->ngroups, 0, 0, e0, e1);
        case TBS_STEP_NORM: return k_final(p->prec, st, p->d_sys, 1, p->max_sweeps, p->rtol, 0, e0, e1);
    }
    tbs::set_error("unknown sweep step %d", step);
    return TBS_ERR_ARG;
}

int tbs_plan_launch_sweeps(tbs_plan *p, int nsweeps, int use_graph) {
    TBS_REQUIRE(p && nsweeps >= 0, TBS_ERR_ARG, "bad arguments");
    TBS_TRY(tbs::use_ctx(p->ctx));
    hipStream_t st = p->ctx->stream;
    auto one = [p, st]() {
        return enqueue_sweep(p->nv, p->prec, st, p->d_sys, p->d_sys, 1, p->d_ltab, p->nline, p->d_gtab,
                             p->ngroups, p->max_sweeps, p->rtol, p->variant);
    };
    if (!use_graph) {
        for (int i = 0; i < nsweeps; ++i) TBS_TRY(one());
        return TBS_OK;
    }
    if (nsweeps == 0) return TBS_OK;
    auto it = p->graphs.find(nsweeps);
    if (it == p->graphs.end()) {
        hipGraphExec_t ge = nullptr;
        TBS_TRY(capture_sweeps(st, nsweeps, &ge, one));
        it = p->graphs.emplace(nsweeps, ge).first;
    }
    TBS_HIP(hipGraphLaunch(it->second, st));
    return TBS_OK;
}

int tbs_plan_end(tbs_plan *p, double *U, double *R_final) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    TBS_TRY(tbs::use_ctx(p->ctx));
    hipStream_t st = p->ctx->stream;
    const long total = (long)p->nle * p->nv;
    if (U) {
        hipLaunchKernelGGL(scatter_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                           p->d_lines, p->d_U, U);
        TBS_LAUNCHED("scatter U");
    }
    if (R_final) {
        TBS_TRY(k_residual(p->nv, p->prec, st, p->d_sys, p->d_gtab, p->ngroups, 1, 1));
        hipLaunchKernelGGL(scatter_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                           p->d_lines, p->d_R, R_final);
        TBS_LAUNCHED("scatter R");
    }
    return TBS_OK;
}

int tbs_plan_history(tbs_plan *p, double *hist, int cap, int *count) {
    TBS_REQUIRE(p && count, TBS_ERR_ARG, "null argument");
    TBS_TRY(tbs::use_ctx(p->ctx));
    hipStream_t st = p->ctx->stream;
    TBS_HIP(hipMemcpyAsync(p->h_state, p->d_state, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
    TBS_HIP(hipStreamSynchronize(st));
    int k = std::min(p->h_state[0], kHistCap);
    *count = k;
    if (hist && cap > 0 && k > 0) {
        TBS_HIP(hipMemcpy(hist, p->d_hist, std::min(k, cap) * sizeof(double), hipMemcpyDeviceToHost));
    }
    return TBS_OK;
}

int tbs_plan_internal(tbs_plan *p, int which, void **dptr, size_t *bytes) {
    TBS_REQUIRE(p && dptr && bytes, TBS_ERR_ARG, "null argument");
    const size_t tf = tf_size(p->prec), nv = p->nv, n = p->nle, G = (size_t)p->ngroups * kGroup;
    switch (which) {
        case 0: *dptr = p->d_g; *bytes = n * nv * tf; break;
        case 1: *dptr = p->d_dUp; *bytes = n * nv * tf; break;
        case 2: *dptr = p->d_U; *bytes = n * nv * 8; break;
        case 3: *dptr = p->d_rhs; *bytes = n * nv * 8; break;
        case 4: *dptr = p->d_R; *bytes = n * nv * 8; break;
        case 5: *dptr = p->d_W; *bytes = G * nv * nv * tf; break;
        case 6: *dptr = p->d_EF; *bytes = (size_t)p->nwin * win_stride((int)nv, p->prec) * tf; break;
        case 7: *dptr = p->d_GB; *bytes = (size_t)p->nwin * win_stride((int)nv, p->prec) * tf; break;
        case 8: *dptr = p->d_blk; *bytes = G * p->S * nv * nv * 8; break;
        default: TBS_REQUIRE(false, TBS_ERR_ARG, "unknown internal array %d", which);
    }
    return TBS_OK;
}

}  // extern "C"

namespace {
// Time each Thomas variant on zeroed vectors (the kernel's cost does not
// depend on the values) and return the fastest.  Leaves the systems' state
// reset; callers start a solve with begin() afterwards.
int autotune_thomas(tbs_ctx *ctx, int nv, int prec, const void *sys, int nsys, const int4 *ltab,
                    int nl, const std::vector<std::pair<double *, long>> &zero_vecs, int reps,
                    int *chosen, double *us, int cap) {
    hipStream_t st = ctx->stream;
    for (auto &zv : zero_vecs) {
        hipLaunchKernelGGL(zero_kernel, dim3(gs_grid(zv.second)), dim3(256), 0, st, zv.second, zv.first);
        TBS_LAUNCHED("zero_kernel");
    }
    TBS_TRY(k_reset(prec, st, sys, nsys));
    const int nvar = thomas_variants(nv, prec);
    hipEvent_t e0, e1;
    TBS_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));   // timing only
    TBS_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
    int best = 0;
    double best_t = 1e300;
    int rc = TBS_OK;
    for (int v = 0; v < nvar && rc == TBS_OK; ++v) {
        rc = k_thomas(nv, prec, st, sys, ltab, nl, v);     // warm-up
        if (rc) break;
        if (hipEventRecord(e0, st)) { rc = TBS_ERR_STATE; break; }
        for (int r = 0; r < reps && rc == TBS_OK; ++r) rc = k_thomas(nv, prec, st, sys, ltab, nl, v);
        if (rc) break;
        if (hipEventRecord(e1, st) || hipEventSynchronize(e1)) { rc = TBS_ERR_STATE; break; }
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        const double t = 1e3 * ms / reps;
        if (us && v < cap) us[v] = t;
        if (t < best_t) { best_t = t; best = v; }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    TBS_TRY(k_reset(prec, st, sys, nsys));
    TBS_HIP(hipStreamSynchronize(st));
    if (chosen) *chosen = best;
    return TBS_OK;
}

// nsweeps sweeps of a batch, enqueued on st.  With split, a batch in two
// halves forks the second half onto the context's second stream once and
// joins it back after the last sweep, so the halves drift apart within the
// run: each half's kernels follow each other, and every system's arithmetic
// is the same as on one stream.  (Captured into a graph, the two branches
// made a solve slower, so graphs are captured on one stream.)
int batch_sweeps(tbs_batch *b, hipStream_t st, int nsweeps, int max_sweeps, double rtol, bool split) {
    const int ns = (int)b->plans.size();
    if (ns == 0 || nsweeps == 0) return TBS_OK;
    if (!split || b->nsA <= 0 || b->nsA >= ns) {
        for (int i = 0; i < nsweeps; ++i)
            TBS_TRY(enqueue_sweep(b->nv, b->prec, st, b->d_sys, b->d_sys, ns, b->d_ltab, b->nl, b->d_gtab, b->ng,
                                  max_sweeps, rtol, b->variant));
        return TBS_OK;
    }
    hipStream_t aux;
    TBS_TRY(tbs::aux_stream(b->ctx, &aux));
    const size_t ssz = b->prec == TBS_FP32 ? sizeof(SysDev<float>) : sizeof(SysDev<double>);
    const void *sysB = (const char *)b->d_sys + ssz * (size_t)b->nsA;
    TBS_HIP(hipEventRecord(b->ctx->ev_fork, st));
    TBS_HIP(hipStreamWaitEvent(aux, b->ctx->ev_fork, 0));
    int rc = TBS_OK;
    for (int i = 0; i < nsweeps && rc == TBS_OK; ++i) {
        rc = enqueue_sweep(b->nv, b->prec, st, b->d_sys, b->d_sys, b->nsA, b->d_ltab, b->nlA, b->d_gtab, b->ngA,
                           max_sweeps, rtol, b->variant);
        if (rc == TBS_OK)
            rc = enqueue_sweep(b->nv, b->prec, aux, b->d_sys, sysB, ns - b->nsA, b->d_ltab + b->nlA, b->nl - b->nlA,
                               b->d_gtab + b->ngA, b->ng - b->ngA, max_sweeps, rtol, b->variant);
    }
    // join even after a failed launch, so that a capture on st can end
    TBS_HIP(hipEventRecord(b->ctx->ev_join, aux));
    TBS_HIP(hipStreamWaitEvent(st, b->ctx->ev_join, 0));
    return rc;
}
}  // namespace

extern "C" {

int tbs_plan_autotune(tbs_plan *p, int reps, int *chosen, double *us, int cap) {
    TBS_REQUIRE(p && reps >= 1, TBS_ERR_ARG, "bad arguments");
    TBS_TRY(tbs::use_ctx(p->ctx));
    TBS_REQUIRE(p->factored, TBS_ERR_STATE, "autotune before factor");
    const long nvec = (long)p->nle * p->nv;
    const size_t tf = tf_size(p->prec);
    std::vector<std::pair<double *, long>> zv = {{p->d_U, nvec}, {p->d_rhs, nvec},
                                                 {(double *)p->d_g, (long)(nvec * tf / 8)}};
    int best = 0;
    TBS_TRY(autotune_thomas(p->ctx, p->nv, p->prec, p->d_sys, 1, p->d_ltab, p->nline, zv, reps, &best, us, cap));
    if (best != p->variant) {
        for (auto &kv : p->graphs) hipGraphExecDestroy(kv.second);
        p->graphs.clear();
        p->variant = best;
    }
    if (chosen) *chosen = best;
    return TBS_OK;
}

int tbs_plan_set_variant(tbs_plan *p, int variant) {
    TBS_REQUIRE(p && variant >= 0 && variant < thomas_variants(p->nv, p->prec), TBS_ERR_ARG,
                "variant %d not built", variant);
    if (variant != p->variant) {
        for (auto &kv : p->graphs) hipGraphExecDestroy(kv.second);
        p->graphs.clear();
        p->variant = variant;
    }
    return TBS_OK;
}

int tbs_plan_variants(tbs_plan *p, int *pf, int cap) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    const int n = thomas_variants(p->nv, p->prec);
    for (int v = 0; v < n && v < cap; ++v) if (pf) pf[v] = thomas_pf_of(p->nv, p->prec, v);
    return n;
}

}  // extern "C"

extern "C" {

int tbs_plan_solve(tbs_plan *p, const double *rhs, double *U, int max_sweeps, double rtol,
                   double *hist, int *nsweeps, double *R_final) {
    TBS_REQUIRE(p && U, TBS_ERR_ARG, "null argument");
    TBS_REQUIRE(max_sweeps >= 0 && max_sweeps < kHistCap, TBS_ERR_ARG,
                "max_sweeps %d outside [0, %d)", max_sweeps, kHistCap);
    TBS_REQUIRE(rtol >= 0.0, TBS_ERR_ARG, "rtol < 0");
    TBS_TRY(tbs::use_ctx(p->ctx));
    if (p->max_sweeps != max_sweeps || p->rtol != rtol) {
        for (auto &kv : p->graphs) hipGraphExecDestroy(kv.second);
        p->graphs.clear();
        p->max_sweeps = max_sweeps;
        p->rtol = rtol;
    }
    TBS_TRY(tbs_plan_begin(p, rhs, U));
    hipStream_t st = p->ctx->stream;
    int done = 0;
    const int chunk = 4;
    // a plan's first solve launches its sweeps one kernel at a time (a graph
    // capture costs more than the launches it saves over a few sweeps);
    // later solves replay captured graphs
    const int use_graph = p->solves++ > 0 ? 1 : 0;
    while (done < max_sweeps) {
        const int k = std::min(chunk, max_sweeps - done);
        TBS_TRY(tbs_plan_launch_sweeps(p, k, use_graph));
        done += k;
        TBS_HIP(hipMemcpyAsync(p->h_state, p->d_state, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
        TBS_HIP(hipStreamSynchronize(st));
        if (!p->h_state[1]) break;
    }
    TBS_TRY(tbs_plan_end(p, U, R_final));
    int count = 0;
    TBS_TRY(tbs_plan_history(p, hist, max_sweeps + 1, &count));
    if (nsweeps) *nsweeps = count > 0 ? count - 1 : 0;
    if (p->h_state[2] == 2 || p->h_state[2] == 4) {
        tbs::set_error(p->h_state[2] == 2 ? "non-finite residual" : "residual diverging");
        return TBS_ERR_NUMERIC;
    }
    return TBS_OK;
}

// ------------------------------------------------------------ batch --
int tbs_batch_create(tbs_ctx *ctx, tbs_plan *const *plans, int nplans, tbs_batch **out) {
    TBS_TRY(tbs::use_ctx(ctx));
    TBS_REQUIRE(out && plans && nplans >= 1 && plans[0], TBS_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    std::unique_ptr<tbs_batch, int (*)(tbs_batch *)> b(new tbs_batch(), tbs_batch_destroy);
    b->ctx = ctx;
    b->nv = plans[0]->nv;
    b->prec = plans[0]->prec;
    TBS_HIP(hipEventCreateWithFlags(&b->staged, hipEventDisableTiming));
    TBS_TRY(tbs_batch_set_plans(b.get(), plans, nplans));
    *out = b.release();
    return TBS_OK;
}

int tbs_batch_set_plans(tbs_batch *b, tbs_plan *const *plans, int nplans) {
    TBS_REQUIRE(b && nplans >= 0 && (plans || nplans == 0), TBS_ERR_ARG, "bad batch arguments");
    TBS_TRY(tbs::use_ctx(b->ctx));
    for (int i = 0; i < nplans; ++i) {
        TBS_REQUIRE(plans[i] && plans[i]->ctx->device == b->ctx->device, TBS_ERR_ARG,
                    "plan %d is on another device", i);
        TBS_REQUIRE(plans[i]->nv == b->nv && plans[i]->prec == b->prec, TBS_ERR_ARG,
                    "plans in a batch must share nvar and precision");
        TBS_REQUIRE(plans[i]->factored, TBS_ERR_STATE, "plan %d not factored", i);
    }
    hipStream_t st = b->ctx->stream;
    if (!b->graphs.empty()) {
        TBS_HIP(hipStreamSynchronize(st));
        for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
        b->graphs.clear();
    }
    // FP64 factors: the systems in two halves of about equal element count,
    // run on two streams (each half's Thomas pass overlaps the other's
    // residual; with FP32 factors one stream is as fast)
    int nsA = 0;
    if (b->prec == TBS_FP64 && nplans >= 2) {
        long tot = 0, acc = 0;
        for (int s2 = 0; s2 < nplans; ++s2) tot += plans[s2]->nle;
        while (nsA < nplans - 1 && 2 * (acc + plans[nsA]->nle) <= tot) acc += plans[nsA++]->nle;
        nsA = std::max(nsA, 1);
    }
    // tables: each half's lines longest-first so the longest chains start
    // first; groups in system order
    std::vector<std::pair<int, int4>> lines;
    std::vector<int2> gtab;
    size_t nlA = 0, ngA = 0;
    auto longest_first = [](const std::pair<int, int4> &x, const std::pair<int, int4> &y) { return x.first > y.first; };
    for (int s2 = 0; s2 < nplans; ++s2) {
        if (s2 == nsA && nsA > 0) {
            std::stable_sort(lines.begin(), lines.end(), longest_first);
            nlA = lines.size();
            ngA = gtab.size();
        }
        const tbs_plan *p = plans[s2];
        for (int l = 0; l < p->nline; ++l)
            lines.push_back({p->h_linesize[l], make_int4(s2, p->h_linepoint[l], p->h_linesize[l], p->h_woff[l])});
        for (int g = 0; g < p->ngroups; ++g) gtab.push_back(make_int2(s2, g));
    }
    std::stable_sort(lines.begin() + nlA, lines.end(), longest_first);
    const size_t ssz = b->prec == TBS_FP32 ? sizeof(SysDev<float>) : sizeof(SysDev<double>);
    const size_t nl = lines.size(), ng = gtab.size(), ns = (size_t)nplans;
    // grow the device tables (after the work that reads them has finished);
    // on any failure below the batch is left empty, so launches are no-ops
    b->plans.clear();
    b->nl = b->ng = 0;
    if (ns > b->cap_sys || nl > b->cap_l || ng > b->cap_g) {
        TBS_HIP(hipStreamSynchronize(st));
        if (ns > b->cap_sys) {
            const size_t c = std::max<size_t>(2 * ns, 16);
            if (b->d_sys) TBS_HIP(hipFree(b->d_sys));
            if (b->d_states) TBS_HIP(hipFree(b->d_states));
            if (b->h_states) TBS_HIP(hipHostFree(b->h_states));
            b->d_sys = nullptr; b->d_states = nullptr; b->h_states = nullptr; b->cap_sys = 0;
            TBS_HIP(hipMalloc(&b->d_sys, c * ssz));
            TBS_HIP(hipMalloc((void **)&b->d_states, c * 4 * sizeof(int)));
            TBS_HIP(hipHostMalloc((void **)&b->h_states, c * 4 * sizeof(int), hipHostMallocDefault));
            b->cap_sys = c;
        }
        if (nl > b->cap_l) {
            const size_t c = std::max<size_t>(2 * nl, 256);
            if (b->d_ltab) TBS_HIP(hipFree(b->d_ltab));
            b->d_ltab = nullptr; b->cap_l = 0;
            TBS_HIP(hipMalloc((void **)&b->d_ltab, c * sizeof(int4)));
            b->cap_l = c;
        }
        if (ng > b->cap_g) {
            const size_t c = std::max<size_t>(2 * ng, 256);
            if (b->d_gtab) TBS_HIP(hipFree(b->d_gtab));
            b->d_gtab = nullptr; b->cap_g = 0;
            TBS_HIP(hipMalloc((void **)&b->d_gtab, c * sizeof(int2)));
            b->cap_g = c;
        }
    }
    // stage through pinned memory; the previous upload must have left it
    const size_t o_l = (ns * ssz + 15) / 16 * 16, o_g = o_l + nl * sizeof(int4), total = o_g + ng * sizeof(int2);
    TBS_HIP(hipEventSynchronize(b->staged));
    if (total > b->cap_stage) {
        if (b->h_stage) TBS_HIP(hipHostFree(b->h_stage));
        b->h_stage = nullptr; b->cap_stage = 0;
        const size_t c = std::max<size_t>(2 * total, 4096);
        TBS_HIP(hipHostMalloc((void **)&b->h_stage, c, hipHostMallocDefault));
        b->cap_stage = c;
    }
    for (size_t s2 = 0; s2 < ns; ++s2) {
        if (b->prec == TBS_FP32) { SysDev<float> d = make_sysdev<float>(plans[s2]); std::memcpy(b->h_stage + ssz * s2, &d, ssz); }
        else { SysDev<double> d = make_sysdev<double>(plans[s2]); std::memcpy(b->h_stage + ssz * s2, &d, ssz); }
    }
    for (size_t i = 0; i < nl; ++i) std::memcpy(b->h_stage + o_l + i * sizeof(int4), &lines[i].second, sizeof(int4));
    if (ng) std::memcpy(b->h_stage + o_g, gtab.data(), ng * sizeof(int2));
    if (ns) TBS_HIP(hipMemcpyAsync(b->d_sys, b->h_stage, ns * ssz, hipMemcpyHostToDevice, st));
    if (nl) TBS_HIP(hipMemcpyAsync(b->d_ltab, b->h_stage + o_l, nl * sizeof(int4), hipMemcpyHostToDevice, st));
    if (ng) TBS_HIP(hipMemcpyAsync(b->d_gtab, b->h_stage + o_g, ng * sizeof(int2), hipMemcpyHostToDevice, st));
    TBS_HIP(hipEventRecord(b->staged, st));
    b->plans.assign(plans, plans + nplans);
    b->nl = (int)nl;
    b->ng = (int)ng;
    b->nsA = nsA;
    b->nlA = (int)nlA;
    b->ngA = (int)ngA;
    return TBS_OK;
}

int tbs_batch_autotune(tbs_batch *b, int reps, int *chosen, double *us, int cap) {
    TBS_REQUIRE(b && reps >= 1, TBS_ERR_ARG, "bad arguments");
    TBS_REQUIRE(!b->plans.empty(), TBS_ERR_STATE, "batch has no systems");
    TBS_TRY(tbs::use_ctx(b->ctx));
    std::vector<std::pair<double *, long>> zv;
    const size_t tf = tf_size(b->prec);
    for (tbs_plan *p : b->plans) {
        const long nvec = (long)p->nle * p->nv;
        zv.push_back({p->d_U, nvec});
        zv.push_back({p->d_rhs, nvec});
        zv.push_back({(double *)p->d_g, (long)(nvec * tf / 8)});
    }
    int best = 0;
    TBS_TRY(autotune_thomas(b->ctx, b->nv, b->prec, b->d_sys, (int)b->plans.size(), b->d_ltab, b->nl, zv,
                            reps, &best, us, cap));
    if (best != b->variant) {
        for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        b->variant = best;
    }
    if (chosen) *chosen = best;
    return TBS_OK;
}

int tbs_batch_launch_sweeps(tbs_batch *b, int nsweeps, int max_sweeps, double rtol) {
    TBS_REQUIRE(b && nsweeps >= 0 && max_sweeps >= 0 && max_sweeps < kHistCap, TBS_ERR_ARG, "bad arguments");
    TBS_TRY(tbs::use_ctx(b->ctx));
    if (b->plans.empty()) return TBS_OK;
    return batch_sweeps(b, b->ctx->stream, nsweeps, max_sweeps, rtol, true);
}

int tbs_batch_states(tbs_batch *b, int *states) {
    TBS_REQUIRE(b && (states || b->plans.empty()), TBS_ERR_ARG, "null argument");
    TBS_TRY(tbs::use_ctx(b->ctx));
    const int ns = (int)b->plans.size();
    if (ns == 0) return TBS_OK;
    hipStream_t st = b->ctx->stream;
    if (b->prec == TBS_FP32)
        hipLaunchKernelGGL((gather_states_kernel<float>), dim3((4 * ns + 255) / 256), dim3(256), 0, st,
                           (const SysDev<float> *)b->d_sys, ns, b->d_states);
    else
        hipLaunchKernelGGL((gather_states_kernel<double>), dim3((4 * ns + 255) / 256), dim3(256), 0, st,
                           (const SysDev<double> *)b->d_sys, ns, b->d_states);
    TBS_LAUNCHED("gather_states_kernel");
    TBS_HIP(hipMemcpyAsync(b->h_states, b->d_states, 4 * sizeof(int) * ns, hipMemcpyDeviceToHost, st));
    TBS_HIP(hipStreamSynchronize(st));
    std::memcpy(states, b->h_states, 4 * sizeof(int) * ns);
    return TBS_OK;
}

int tbs_batch_destroy(tbs_batch *b) {
    if (!b) return TBS_OK;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
    if (b->d_sys) hipFree(b->d_sys);
    if (b->d_ltab) hipFree(b->d_ltab);
    if (b->d_gtab) hipFree(b->d_gtab);
    if (b->d_states) hipFree(b->d_states);
    if (b->h_states) hipHostFree(b->h_states);
    if (b->h_stage) hipHostFree(b->h_stage);
    if (b->staged) hipEventDestroy(b->staged);
    delete b;
    return TBS_OK;
}

int tbs_batch_solve(tbs_batch *b, const double *const *rhs, double *const *U, int max_sweeps,
                    double rtol, int sweeps_per_check, double *hist, int *nsweeps) {
    TBS_REQUIRE(b && rhs && U, TBS_ERR_ARG, "null argument");
    TBS_REQUIRE(!b->plans.empty(), TBS_ERR_STATE, "batch has no systems");
    TBS_REQUIRE(max_sweeps >= 0 && max_sweeps < kHistCap, TBS_ERR_ARG, "bad max_sweeps");
    TBS_TRY(tbs::use_ctx(b->ctx));
    const int ns = (int)b->plans.size();
    hipStream_t st = b->ctx->stream;
    if (sweeps_per_check < 1) sweeps_per_check = 1;
    for (int s = 0; s < ns; ++s) {
        tbs_plan *p = b->plans[s];
        TBS_REQUIRE(rhs[s] && U[s], TBS_ERR_ARG, "null vector for system %d", s);
        const long total = (long)p->nle * p->nv;
        hipLaunchKernelGGL(gather_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                           p->d_lines, rhs[s], p->d_rhs);
        hipLaunchKernelGGL(gather_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                           p->d_lines, (const double *)U[s], p->d_U);
        TBS_LAUNCHED("batch gather");
    }
    TBS_TRY(k_reset(b->prec, st, b->d_sys, ns));
    TBS_TRY(k_residual(b->nv, b->prec, st, b->d_sys, b->d_gtab, b->ng, 0, 1));
    TBS_TRY(k_final(b->prec, st, b->d_sys, ns, max_sweeps, rtol, 1));
    auto key = std::make_pair(sweeps_per_check, std::make_pair(max_sweeps, rtol));
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        hipGraphExec_t ge = nullptr;
        auto all = [b, st, sweeps_per_check, max_sweeps, rtol]() {
            return batch_sweeps(b, st, sweeps_per_check, max_sweeps, rtol, false);
        };
        TBS_TRY(capture_sweeps(st, 1, &ge, all));
        it = b->graphs.emplace(key, ge).first;
    }
    int done = 0;
    std::vector<int> st_h(4 * ns);
    while (done < max_sweeps) {
        TBS_HIP(hipGraphLaunch(it->second, st));
        done += sweeps_per_check;
        bool any = false;
        TBS_TRY(tbs_batch_states(b, st_h.data()));
        for (int s = 0; s < ns; ++s) any |= st_h[4 * s + 1] != 0;
        if (!any) break;
    }
    for (int s = 0; s < ns; ++s) {
        TBS_TRY(tbs_plan_end(b->plans[s], U[s], nullptr));
        int count = 0;
        TBS_TRY(tbs_plan_history(b->plans[s], hist ? hist + (size_t)s * (max_sweeps + 1) : nullptr,
                                 max_sweeps + 1, &count));
        if (nsweeps) nsweeps[s] = count > 0 ? count - 1 : 0;
    }
    return TBS_OK;
}

}  // extern "C"
