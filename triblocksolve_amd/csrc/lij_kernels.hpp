// lij_kernels.hpp -- the fused LIJ path's kernels and their launch
// templates, shared by the translation units that instantiate them for the
// block sizes of the library (lij_fast.hip: 2..16 and the host side;
// lij_wide_lo.hip / lij_wide_hi.hip: 17..24 / 25..32).
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
#pragma once
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

namespace tbs {
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
}  // namespace tbs

namespace {

using tbs::rdlane;

constexpr int kGroup = 64;      // elements per residual group (one wave)
constexpr int kHistCap = 4096;  // sweeps of history kept per system
constexpr int kTuneThomas = TBS_TUNE_THOMAS_MAX, kTuneResidual = TBS_TUNE_RESIDUAL_MAX;
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

// RW waves (64-element groups) per workgroup.
template <int NV, typename TF, int RW>
__global__ __launch_bounds__(64 * RW) void fres_kernel(const SysDev<TF> *__restrict__ sys,
                                                       const int2 *__restrict__ gtab, int ng,
                                                       int want_R, int force) {
    constexpr int BS = NV * NV;
    __shared__ double tiles[RW][kGroup * NV];
    // XCD-aware order: hardware block b runs on XCD b % 8; give each XCD a
    // contiguous run of groups so neighbouring lines share an L2.
    const int nblk = (ng + RW - 1) / RW;
    const int per = (nblk + 7) / 8;
    const int blk = (int)(blockIdx.x % 8u) * per + (int)(blockIdx.x / 8u);
    if (blk >= nblk) return;
    const int wave = threadIdx.x >> 6;
    const int wid = __builtin_amdgcn_readfirstlane(blk * RW + wave);
    if (wid >= ng) return;
    double *tile = tiles[wave];
    const int lane = threadIdx.x & 63;
    const int2 sg = gtab[wid];
    const SysDev<TF> &S = sys[sg.x];
    // force bit 0: run even if the system is inactive; bit 1: U is zero
    // (the first residual of a solve from zero is R = rhs: skip J U); bit 2:
    // U is zero unless the gather of the initial U flagged a nonzero value
    // (state[3])
    if (!(force & 1) && !S.state[1]) return;
    const bool u_zero = (force & 2) != 0 || ((force & 4) != 0 && S.state[3] == 0);
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
            if constexpr (NV > 16) {
                // large blocks: the face product goes straight into the sum
                // (one NV-vector fewer in registers)
                slot_matvec<NV, double>(rB, (uint32_t)(1 + f) * kSlot, lane, un, acc);
            } else {
                double t[NV];
#pragma unroll
                for (int i = 0; i < NV; ++i) t[i] = 0.0;
                slot_matvec<NV, double>(rB, (uint32_t)(1 + f) * kSlot, lane, un, t);
#pragma unroll
                for (int i = 0; i < NV; ++i) acc[i] += t[i];
            }
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
// A workgroup holds LPW lines (2 LPW waves, consecutive table lines); a line
// past the table's end or of an inactive system only takes part in the
// workgroup's two barriers.
template <int NV, typename TF, int PF, int K, int LPW>
__global__ __launch_bounds__(128 * LPW) void fthomas_kernel(const SysDev<TF> *__restrict__ sys,
                                                            const int4 *__restrict__ ltab, int nl) {
    constexpr int SPW = 64 / NV;      // steps per window
    constexpr int LW = SPW * NV;      // lanes carrying data
    constexpr int NB = PF + 1;        // register ring depth
    constexpr int G = K * NB;         // windows per group (loop iteration)
    constexpr int VG = G * LW;        // vector values per group
    using IOT = GroupIO<VG, TF>;
    using IOD = GroupIO<VG, double>;
    __shared__ TF s_vin[2 * LPW][2][IOT::LDS];    // [wave] forward: g; backward: d'
    __shared__ TF s_out[2 * LPW][IOT::LDS];       // [wave] forward: d' staging
    __shared__ double s_u[2 * LPW][2][IOD::LDS];  // [wave] backward: U (updated in place)
    __shared__ TF s_xall[LPW][3][NV];             // d' at step m-1, d' at step m+1, d at step m
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);   // wave of the workgroup
    const int l = (int)blockIdx.x * LPW + (wave >> 1);
    const bool bot = (wave & 1) == 1;
    auto s_x = s_xall[wave >> 1];
    // {system, m0, n, window offset}, wave-uniform by construction: the
    // group loops below must be scalar loops (the chain broadcasts read
    // lanes that every step writes)
    int4 lt = l < nl ? ltab[l] : make_int4(0, 0, 0, 0);
    lt.x = __builtin_amdgcn_readfirstlane(lt.x);
    lt.y = __builtin_amdgcn_readfirstlane(lt.y);
    lt.z = __builtin_amdgcn_readfirstlane(lt.z);
    lt.w = __builtin_amdgcn_readfirstlane(lt.w);
    const SysDev<TF> &S = sys[lt.x];
    // (for LPW = 1 the whole workgroup; else the line's two waves, which still
    // meet the workgroup's barriers)
    const bool live = l < nl && __builtin_amdgcn_readfirstlane(S.state[1]) != 0;
    if (LPW == 1 && !live) return;
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

    if (lane < NV) s_x[bot ? 1 : 0][lane] = TF(0);                      // an empty side contributes 0
    TF d[NV];
    TF eb[NB][NV];
    if (live) {
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
            if (w == nw - 1 && st_of_lane == SPW - 1 && lane < LW) s_x[bot ? 1 : 0][row] = out;
        }
        bulk_store<VG, TF>(rd, vbytes(gbase(jg), sizeof(TF)), lane, s_out[wave]);
        if (jg + 1 < ngrp) bulk_commit<VG, TF>(vr, lane, s_vin[wave][cur ^ 1]);
    }
    }
    __syncthreads();
    // ---- step m: d_m = g_m - E_m d'_{m-1} - F_m d'_{m+1}, U_m += d_m
    if (live && !bot && lane < NV) {
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
    if (!live) return;

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
        sys[i].state[3] = 0;
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

// Same packing through LDS, one wave per quarter of a (group, slot): the 16
// source blocks of lanes q*16 .. q*16+15 come into LDS whole in 16-byte
// pieces (consecutive threads read consecutive pieces of a block, which is
// contiguous in the source array and 8-byte aligned; the last piece of an odd
// block is read as the 16 bytes that end the block) and leave as those lanes'
// part of the SoA-64 slot: 256-byte runs of 16-byte pieces.  Each thread
// issues its loads NB at a time before writing any to LDS, so a wave keeps NB
// KiB in flight; 10.5 KB of LDS and 41 VGPRs (9x9 blocks) let the waves sit
// beside the factor kernel's.  The grid strides over the units (any grid size
// is right).
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));
constexpr int kPackLanes = 16;
template <int NV, int NB = 6>
__global__ __launch_bounds__(64) void pack_lds_kernel(long nunits, int S, const int2 *__restrict__ src,
                                                      const double *__restrict__ D,
                                                      const double *__restrict__ O1,
                                                      const double *__restrict__ O2,
                                                      double *__restrict__ blk) {
    constexpr int BS = NV * NV;
    constexpr int EL = kPackLanes;
    constexpr int PB = (BS + 1) / 2;           // 16-byte pieces per block
    constexpr int NP = EL * PB;                // pieces per unit
    constexpr int NIT = (NP + 63) / 64;        // pieces per thread
    constexpr int NU = kGroup / EL;            // units per (group, slot)
    using L = tbs::Soa64<BS, double>;
    static_assert(L::VW == 2, "SoA-64 slots of doubles hold 2 entries per 16-byte chunk");
    using gsrc = const __attribute__((address_space(1))) f64x2u *;
    using gdst = __attribute__((address_space(1))) f64x2u *;
    __shared__ double s[EL * BS];              // [element lane - q*EL][entry]
    __shared__ uintptr_t s_src[EL];            // block addresses (0: the zero block)
    const int t = threadIdx.x;
    for (long unit = blockIdx.x; unit < nunits; unit += gridDim.x) {
        const long w = unit / NU;              // (group, slot)
        const int q = (int)(unit - w * NU);
        const long g = w / S;
        const int slot = (int)(w - g * S);
        if (t < EL) {
            const int2 c = src[(g * kGroup + q * EL + t) * S + slot];
            const double *a = c.x == 1 ? D : (c.x == 2 ? O1 : (c.x == 3 ? O2 : nullptr));
            s_src[t] = a ? (uintptr_t)(a + (size_t)BS * c.y) : 0;
        }
        __syncthreads();
#pragma unroll
        for (int b0 = 0; b0 < NIT; b0 += NB) {
            f64x2u v[NB];
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                const int i = t + 64 * (b0 + r);
                if (b0 + r < NIT) {
                    const int e = i < NP ? i / PB : 0, k = i < NP ? i - e * PB : 0;
                    const uintptr_t a = s_src[e];
                    const bool odd_last = (BS & 1) && k == PB - 1;
                    // a missing block reads D's first block and keeps zeros
                    const uintptr_t at = (a ? a : (uintptr_t)D) + (uintptr_t)(odd_last ? 8 * (BS - 2) : 16 * k);
                    v[r] = *(gsrc)at;
                    if (!a) v[r] = f64x2u{0.0, 0.0};
                }
            }
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                const int i = t + 64 * (b0 + r);
                if (b0 + r < NIT && i < NP) {
                    const int e = i / PB, k = i - e * PB;
                    double *d = s + e * BS + 2 * k;
                    if ((BS & 1) && k == PB - 1) {
                        d[0] = v[r][1];
                    } else {
                        d[0] = v[r][0];
                        d[1] = v[r][1];
                    }
                }
            }
        }
        __syncthreads();
        double *out = blk + w * L::ELEMS;
        constexpr int NO = (L::NC * EL + 63) / 64;
#pragma unroll
        for (int r = 0; r < NO; ++r) {         // chunk c, lane q*EL + e: entries 2c, 2c+1
            const int j = t + 64 * r;
            if (j < L::NC * EL) {
                const int c = j / EL, e = j - c * EL;
                const double *x = s + e * BS + 2 * c;
                *(gdst)(out + c * 128 + (q * EL + e) * 2) = f64x2u{x[0], x[1]};
            }
        }
        for (int j = t; j < L::REM * EL; j += 64)   // odd BS: the last entry, [lane]
            out[L::NC * 128 + q * EL + j] = s[j * BS + BS - 1];
        __syncthreads();                        // LDS reused by the next unit
    }
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
// and the outputs are staged in the step's LDS slot and stored a chunk later
// in 16-byte pieces (see the output descriptors below).  The D, A, C blocks of 4
// steps at a time (2 for blocks larger than 10x10) are prefetched into
// double-buffered LDS one chunk ahead;
// their sources come from the per-element table fidx = {e, A source, C
// source of the face before} (source = 2 * face + 1 for Of2, 2 * face for
// Of1, -1 none: the zero block).
using tbs::FactorRec;

// t += (a of lane m of the lane's row of 16) * v[m] for m = 0..NV-1 in
// order: a DPP row broadcast per term (NV <= 16)
template <int NV, int M = 0>
__device__ __forceinline__ void row_bcast_dot(double a, const double (&v)[NV], double &t) {
    if constexpr (M < NV) {
        t += __builtin_amdgcn_mov_dpp(a, 0x150 + M, 0xF, 0xF, true) * v[M];
        row_bcast_dot<NV, M + 1>(a, v, t);
    }
}

// (9x9: at most 168 VGPRs, three waves per SIMD, so that 1,200 lines are
// resident at once: 4.0-4.2 ms against 4.2-4.4 ms for 1,200 lines of 1,280
// with the packing beside it, no registers spilled, the same elsewhere)
template <int NV, typename TF>
__global__ __launch_bounds__(128, NV == 9 ? 3 : 1) void ffactor_kernel(FactorRec one, int nline,
                                                                       const FactorRec *__restrict__ recs,
                                                                       const int2 *__restrict__ ltab) {
    constexpr int BS = NV * NV;
    constexpr int CH = NV <= 10 ? 4 : 2;        // steps per chunk (LDS: two waves, two buffers)
    constexpr int NB = 3 * CH;                  // blocks per chunk: D_k, A_k, C_{k+1} of each step
    constexpr int SB = 3 * BS;                  // doubles per step
    constexpr int NLD = (CH * SB + 63) / 64;    // loads per lane per chunk
    constexpr int SPW = 64 / NV, LW = SPW * NV;
    // How the S_k update gets A_k's entries (below), by what was measured
    // fastest (factorisation at 384,000 elements, and at 1,024-1,600 lines of
    // 1,280 for 9x9; the LDS form = 1): 0 LDS reads at a uniform address
    // (2x2 to 8x8, 10x10, 11x11, 16x16: readlane 1.01-1.10, DPP 0.93-1.08
    // with no consistent gain);
    // 2 DPP row broadcasts (9x9 FP64, 12x12 to 15x15: 0.80-0.96); 1 readlane
    // broadcasts (9x9 FP32 factors: the fewest VGPRs, 168, so three waves
    // share a SIMD and 1,200 lines fit one round, 0.90 against 0.97 with
    // the DPP form)
    constexpr int kAMode = NV == 9 && sizeof(TF) == 4 ? 1 : (NV == 9 || (NV >= 12 && NV <= 15) ? 2 : 0);
    using WL = tbs::WinLayout<NV, LW, TF>;
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
    // outputs: a step's Gamma (X lanes), E (A lanes) and W (W lanes) are
    // staged in the step's own LDS slot once its D, A, C have been read -- W
    // column-major at [0, BS), E and Gamma row-major at [BS, 2 BS) and
    // [2 BS, 3 BS), in TF -- and a chunk's slots are stored at the start of
    // the next chunk, before that chunk's block loads are issued.  The stores
    // and the loads share the wave's memory counter, which completes in
    // order, and the chunk start waits for the loads it needs: stores issued
    // in the middle of a chunk would make that wait include their write
    // latency (stored right after each step, the same outputs took the
    // kernel from 2.33 to 2.66 ms at 1,024 lines of 1,280, 9x9 FP64).  A
    // slot leaves as 16-byte pieces of VW values: W as
    // its element's SoA-64 chunks (q = i + NV j, VW consecutive entries), E
    // and Gamma as the VW-entry chunks of the step's NV window lanes; then
    // the leftover single values.  Lane t's pieces are the same for every
    // step: descriptors kind << 30 | LDS offset << 20 | destination offset,
    // kind 0 W, 1 E, 2 Gamma, 3 none.
    constexpr int VW = 16 / (int)sizeof(TF);
    constexpr int NCW = BS / VW, REMW = BS - NCW * VW;   // W: pieces, leftover values
    constexpr int NCE = NV / VW, REME = NV - NCE * VW;   // one row of E or Gamma
    constexpr int NCEd = NCE > 0 ? NCE : 1, REMEd = REME > 0 ? REME : 1;   // divisors where a branch is unreached
    constexpr int P16 = NCW + 2 * NV * NCE, PR = REMW + 2 * NV * REME;
    constexpr int NI = (P16 + 63) / 64, NR = (PR + 63) / 64;
    static_assert(3 * BS < 1024 && NCW * 64 * VW < (1 << 20), "output descriptor fields");
    uint32_t pdesc[NI + NR];
#pragma unroll
    for (int t = 0; t < NI; ++t) {
        const int r = lane + 64 * t;
        uint32_t d = 3u << 30;
        if (r < NCW) {
            d = (uint32_t)(r * VW) << 20 | (uint32_t)(r * 64 * VW);
        } else if (r < P16) {
            const int r1 = r - NCW, h = r1 / (NV * NCEd), r2 = r1 - h * NV * NCE, i = r2 / NCEd, c = r2 - i * NCE;
            d = (uint32_t)(1 + h) << 30 | (uint32_t)((1 + h) * BS + i * NV + c * VW) << 20 |
                (uint32_t)(c * LW * VW + i * VW);
        }
        pdesc[t] = d;
    }
#pragma unroll
    for (int t = 0; t < NR; ++t) {
        const int r = lane + 64 * t;
        uint32_t d = 3u << 30;
        if (r < REMW) {
            d = (uint32_t)(NCW * VW + r) << 20 | (uint32_t)(r * 64);
        } else if (r < PR) {
            const int r1 = r - REMW, h = r1 / (NV * REMEd), r2 = r1 - h * NV * REME, i = r2 / REMEd, e = r2 - i * REME;
            d = (uint32_t)(1 + h) << 30 | (uint32_t)((1 + h) * BS + i * NV + NCE * VW + e) << 20 |
                (uint32_t)(e * LW + i);
        }
        pdesc[NI + t] = d;
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

    const uintptr_t aEF = (uintptr_t)EF, dGB = (uintptr_t)GB - (uintptr_t)EF;
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
        // S_k = D_k - A_k Gamma_{k-1} on the pivot lanes; the other lanes
        // subtract exact zeros.  Every lane reads all of A_k from LDS at a
        // uniform address (a VGPR pair per entry, and the reads wait for LDS
        // about ten times a step), or A_k is held in the wave and each entry
        // is broadcast: lane 16 r + m holds column m (every row of 16 lanes a
        // copy) and a DPP row broadcast takes entry (i, m) from lane m of
        // the lane's own row (one VGPR-to-VGPR move per entry), or lane t
        // holds entry t + 64 r and readlane broadcasts it (two moves to
        // SGPRs per entry, four VGPRs in all).  The sums are the same, in
        // the same order, every way.
        if constexpr (kAMode == 2) {
            const int cm = lane & 15;
            const double *acol = Ak + NV * (cm < NV ? cm : 0);
            double ac[NV];
#pragma unroll
            for (int i = 0; i < NV; ++i) ac[i] = acol[i];
            double bv[NV];
#pragma unroll
            for (int i = 0; i < NV; ++i) bv[i] = bcol[i];
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                double tot = 0.0;
                row_bcast_dot<NV>(ac[i], vold, tot);
                col[i] = bv[i] - tot;
            }
        } else if constexpr (kAMode == 1) {
            constexpr int NA = (BS + 63) / 64;
            double av[NA];
#pragma unroll
            for (int r = 0; r < NA; ++r) av[r] = Ak[lane + 64 * r < BS ? lane + 64 * r : 0];
            double bv[NV];
#pragma unroll
            for (int i = 0; i < NV; ++i) bv[i] = bcol[i];
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                double tot = 0.0;
#pragma unroll
                for (int m = 0; m < NV; ++m) {
                    const int q = i + NV * m;
                    tot += tbs::rdlane(av[q / 64], q % 64) * vold[m];
                }
                col[i] = bv[i] - tot;
            }
        } else {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                double tot = 0.0;
#pragma unroll
                for (int m = 0; m < NV; ++m) tot += Ak[i + NV * m] * vold[m];
                col[i] = bcol[i] - tot;
            }
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
        // outputs: Gamma_k (X lanes), E_k (A lanes), W_k into the step's LDS
        // slot (its D, A, C have been read); every window entry of a real
        // step is written: E_0, E''_{n-1} and Gamma_{n-1} (no neighbour) as
        // zeros, so the factor arrays need no zero fill
        const bool act = grp == xg || grp == 2 || grp == 3;
        const bool zero = bot ? (grp == 2 && sw == 0) : ((grp == xg && k == n - 1) || (grp == 2 && k == 0));
        TF *stg = (TF *)&sblk[b][s * SB];
        wsync();
        if (act) {
            const int o = grp == 3 ? NV * j : (grp == 2 ? BS : 2 * BS) + j;
            const int os = grp == 3 ? 1 : NV;
#pragma unroll
            for (int i = 0; i < NV; ++i) stg[o + i * os] = zero ? TF(0) : (TF)col[i];
        }
        wsync();
    };

    // store the staged outputs of the first nst steps of chunk c (buffer
    // c & 1) at the wave's window positions of those steps (the bottom
    // wave's windows start with pad empty steps)
    auto flush = [&](int c, int nst) {
        using gdst = __attribute__((address_space(1))) TF *;
        typedef TF tf16 __attribute__((ext_vector_type(VW)));
        using gdst16 = __attribute__((address_space(1))) tf16 *;
        wsync();
        for (int s = 0; s < nst; ++s) {
            const int sw = c * CH + s;
            const int k = bot ? n - 1 - sw : sw;
            const int pos = bot ? sw + sp.pad : sw;
            const int w = pos / SPW, st = pos - w * SPW;
            const int mm = m0 + k;
            const TF *stg = (const TF *)&sblk[c & 1][s * SB];
            const size_t wg = (size_t)(mm / kGroup) * BS * kGroup;
            const long wofs = wbase + (long)w * WL::STRIDE;
            TF *const b16[3] = {W + wg + (mm % kGroup) * VW, EF + wofs + st * NV * VW, GB + wofs + st * NV * VW};
            TF *const b1[3] = {W + wg + NCW * VW * kGroup + mm % kGroup, EF + wofs + NCE * VW * LW + st * NV,
                               GB + wofs + NCE * VW * LW + st * NV};
#pragma unroll
            for (int t = 0; t < NI; ++t) {
                const uint32_t d = pdesc[t], kind = d >> 30;
                if (kind != 3) {
                    const TF *src = stg + ((d >> 20) & 1023u);
                    tf16 x;
#pragma unroll
                    for (int q = 0; q < VW; ++q) x[q] = src[q];
                    TF *base = kind == 0 ? b16[0] : (kind == 1 ? b16[1] : b16[2]);
                    *(gdst16)(base + (d & 0xFFFFFu)) = x;
                }
            }
#pragma unroll
            for (int t = 0; t < NR; ++t) {
                const uint32_t d = pdesc[NI + t], kind = d >> 30;
                if (kind != 3) {
                    TF *base = kind == 0 ? b1[0] : (kind == 1 ? b1[1] : b1[2]);
                    *(gdst)(base + (d & 0xFFFFFu)) = stg[(d >> 20) & 1023u];
                }
            }
        }
        wsync();
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
            flush(c - 1, CH);                           // outputs of chunk c-1
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
    if (nsteps > 0) flush((nsteps - 1) / CH, nsteps - (nsteps - 1) / CH * CH);   // the last chunk
    if (bot) {
        // Gamma''_{m+1}: the X lanes of the last step (group 1 after an even
        // step, 0 after an odd one); zeros when the bottom side is empty
        // (written after the last chunk's outputs have left its buffer)
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
                const long at = WL::index(j, st * NV + i);
                e[at] = TF(0);
                g[at] = TF(0);
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

// The same factorisation for block sizes whose four lane roles do not fit one
// wave (4 * NV > 64, i.e. NV 17..32): every lane holds two columns of
// [S | C | A | I].  The lanes of the pivot half hold S_k (first column) and
// A_k (second column, ends as E_k); the other half holds C_{k+1} (ends as
// Gamma_k, or F_m at the middle step) and the identity (ends as W_k).  Each
// Gauss-Jordan pivot broadcasts the pivot column of S once and updates both
// columns of every lane.  The halves swap roles every step, so the lanes
// that end a step holding Gamma_k form S_{k+1} in place, as in
// ffactor_kernel.  One step's D, A, C blocks (3 NV^2 values) are staged in
// double-buffered LDS one step ahead, the next step's in registers; outputs,
// split, middle step and zero padding are those of ffactor_kernel, so the
// sweep kernels read the factors of every block size alike.
template <int NV, typename TF>
__global__ __launch_bounds__(128) void ffactor_wide_kernel(FactorRec one, int nline,
                                                           const FactorRec *__restrict__ recs,
                                                           const int2 *__restrict__ ltab) {
    constexpr int BS = NV * NV;
    constexpr int SB = 3 * BS;                  // doubles per step: D_k, A_k, C_{k+1}
    constexpr int NLD = (SB + 63) / 64;         // loads per lane per step
    constexpr int SPW = 64 / NV, LW = SPW * NV;
    using WL = tbs::WinLayout<NV, LW, TF>;
    static_assert(2 * NV <= 64 && 4 * NV > 64, "two halves of NV lanes; smaller blocks use ffactor_kernel");
    constexpr bool kDefer = NV <= 24;           // outputs staged in LDS and stored a step later
    __shared__ double sblk_all[2][2][SB];       // [wave][buffer]
    __shared__ double sI[BS];                   // identity: the second column of the non-pivot half
    __shared__ const double *sptr_all[2][2][3];  // [wave][buffer]: sources of a step's three blocks

    if ((int)blockIdx.x >= nline) return;
    int l = blockIdx.x;
    FactorRec fr = one;
    if (ltab) {
        const int2 e = ltab[blockIdx.x];
        fr = recs[__builtin_amdgcn_readfirstlane(e.x)];
        l = __builtin_amdgcn_readfirstlane(e.y);
    }
    const int4 *__restrict__ fidx = fr.fidx;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const bool bot = wave == 1;
    const int lane = threadIdx.x & 63;
    auto sblk = sblk_all[wave];
    auto sptr = sptr_all[wave];
    double *sG = &sblk_all[1][0][0];            // Gamma''_{m+1} for step m (bottom wave's buffer 0)
    // wave-uniform: the step loop must be a scalar loop (the pivot
    // broadcasts read lanes that every step writes)
    const int m0 = __builtin_amdgcn_readfirstlane(fr.linepoint[l]);
    const int n = __builtin_amdgcn_readfirstlane(fr.linesize[l]);
    const Split sp = split_line(n, SPW);
    const int nsteps = bot ? sp.B : sp.m + 1;
    const long wbase = ((long)__builtin_amdgcn_readfirstlane(fr.woff[l]) + (bot ? sp.q + 1 : 0)) * WL::STRIDE;
    auto wsync = []() { __builtin_amdgcn_wave_barrier(); };

    const int half = lane / NV;                 // 0 or 1; 2: idle lanes (NV < 32)
    const bool idle = half >= 2;
    const int j = idle ? 0 : lane - half * NV;  // the lane's column
    if (!bot)
        for (int t = lane; t < BS; t += 64) sI[t] = (t % NV == t / NV) ? 1.0 : 0.0;
    // step entry t = lane + 64 r: block `which` (0 D, 1 A, 2 C) << 12 | entry
    int code[NLD];
#pragma unroll
    for (int r = 0; r < NLD; ++r) {
        const int t = lane + 64 * r;
        const int which = t / BS;
        code[r] = t < SB ? (which << 12) | (t - which * BS) : -1;
    }
    // lane t < 3 holds the source of block t of a step (fidx components:
    // top wave D_k, A_k, C_{k+1}; bottom wave D_k, C_{k+1}, A_k; -1 none)
    const int nextslot = bot ? 1 : 2;
    auto load_idx = [&](int s) -> int {
        const int k = (bot ? n - 1 - s : s) + (lane == nextslot ? 1 : 0);
        const int comp = lane == 0 ? 0 : (lane == nextslot ? 2 : 1);
        int code_src = -1;
        if (lane < 3 && s < nsteps && k < n) code_src = reinterpret_cast<const int *>(fidx + m0 + k)[comp];
        return code_src;
    };
    const uintptr_t aD = (uintptr_t)fr.D, aO1 = (uintptr_t)fr.O1, aZ = (uintptr_t)fr.zero;
    const uintptr_t dO = (uintptr_t)fr.O2 - aO1;
    auto set_ptrs = [&](int b, int cs) {
        if (lane < 3) {
            const uintptr_t face = aO1 + (uintptr_t)(cs & 1) * dO + (uintptr_t)((long)(cs >> 1) * BS * 8);
            const uintptr_t elem = aD + (uintptr_t)((long)cs * BS * 8);
            const uintptr_t p = lane == 0 ? elem : face;
            sptr[b][lane] = (const double *)(cs >= 0 ? p : aZ);
        }
    };
    using gptr = const __attribute__((address_space(1))) double *;
    auto load_blocks = [&](int b, double (&R)[NLD]) {
#pragma unroll
        for (int r = 0; r < NLD; ++r) {
            const int cd = code[r] < 0 ? 0 : code[r];
            const gptr src = (gptr)sptr[b][cd >> 12];
            R[r] = src[cd & 4095];
        }
    };
    auto store_blocks = [&](int b, const double (&R)[NLD]) {
#pragma unroll
        for (int r = 0; r < NLD; ++r) {
            const int t = lane + 64 * r;
            if (r < NLD - 1 || t < SB) sblk[b][t] = R[r];
        }
    };

    double R[NLD];
    set_ptrs(0, load_idx(0));
    __syncthreads();                            // sI, and wave-local sptr
    load_blocks(0, R);
    store_blocks(0, R);
    set_ptrs(1, load_idx(1));
    wsync();
    load_blocks(1, R);
    int xnext = load_idx(2);

    double col1[NV], col2[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) col1[i] = col2[i] = 0.0;

    const uintptr_t aW = (uintptr_t)fr.W, aEF = (uintptr_t)fr.EF, dGB = (uintptr_t)fr.GB - (uintptr_t)fr.EF;
    using gdst = __attribute__((address_space(1))) TF *;
    auto step = [&](auto par, int b, int sw) {
        constexpr int PAR = decltype(par)::value;
        constexpr int P = PAR ? NV : 0;
        const int k = bot ? n - 1 - sw : sw;
        const double *Dk = &sblk[b][0];
        const double *Ak = Dk + BS;
        const double *Ck = Dk + 2 * BS;
        const bool piv = half == PAR;
        const double f = piv ? 1.0 : 0.0;
        // new columns: pivot half D_k and A_k, the other half C_{k+1} and I
        const double *s1 = idle ? sI : (piv ? Dk : Ck);
        const double *s2 = idle ? sI : (piv ? Ak : sI);
        double vold[NV];
#pragma unroll
        for (int m = 0; m < NV; ++m) vold[m] = col1[m] * f;
        // S_k = D_k - A_k Gamma_{k-1} on the pivot half (A_k uniform); the
        // other lanes subtract exact zeros
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double tot = 0.0;
#pragma unroll
            for (int m = 0; m < NV; ++m) tot += Ak[i + NV * m] * vold[m];
            col1[i] = s1[i + NV * j] - tot;
            col2[i] = s2[i + NV * j];
        }
        // Gauss-Jordan on the pivot columns, applied to both columns of every lane
#pragma unroll
        for (int p = 0; p < NV; ++p) {
            double cpv[NV];
#pragma unroll
            for (int i = 0; i < NV; ++i) cpv[i] = tbs::rdlane(col1[i], P + p);
            const double rp = recip(cpv[p]);
            const double x1 = col1[p] * rp, x2 = col2[p] * rp;
#pragma unroll
            for (int i = 0; i < NV; ++i)
                if (i != p) {
                    col1[i] -= cpv[i] * x1;
                    col2[i] -= cpv[i] * x2;
                }
            col1[p] = x1;
            col2[p] = x2;
        }
        // outputs: E_k (pivot half, second column), Gamma_k (other half,
        // first column), W_k (other half, second column); E_0, E''_{n-1} and
        // Gamma_{n-1} (no neighbour) as zeros.  Up to 24x24 blocks they are
        // staged in the step's LDS slot (its D, A, C have been read): W
        // column-major at [0, BS), E and Gamma row-major at [BS, 2 BS) and
        // [2 BS, 3 BS), in TF, and stored during the next step's set-up,
        // after its block loads (flush below); larger blocks store them
        // here, at the wave's window position of the step.
        if constexpr (kDefer) {
            TF *stg = (TF *)&sblk[b][0];
            wsync();
            if (!idle) {
                if (piv) {
                    const bool zero = bot ? sw == 0 : k == 0;
#pragma unroll
                    for (int i = 0; i < NV; ++i) stg[BS + i * NV + j] = zero ? TF(0) : (TF)col2[i];
                } else {
                    const bool zero = !bot && k == n - 1;
#pragma unroll
                    for (int i = 0; i < NV; ++i) {
                        stg[2 * BS + i * NV + j] = zero ? TF(0) : (TF)col1[i];
                        stg[NV * j + i] = (TF)col2[i];
                    }
                }
            }
            wsync();
        } else if (!idle) {
            const int pos = bot ? sw + sp.pad : sw;
            const int w = pos / SPW, st = pos - w * SPW;
            const int mm = m0 + k;
            const long wofs = wbase + (long)w * WL::STRIDE;
            if (piv) {
                const bool zero = bot ? sw == 0 : k == 0;
                const gdst e = (gdst)(TF *)(aEF + (uintptr_t)(wofs * (long)sizeof(TF)));
#pragma unroll
                for (int i = 0; i < NV; ++i) e[WL::index(j, st * NV + i)] = zero ? TF(0) : (TF)col2[i];
            } else {
                const bool zero = !bot && k == n - 1;
                const gdst g = (gdst)(TF *)(aEF + dGB + (uintptr_t)(wofs * (long)sizeof(TF)));
                const gdst wd = (gdst)(TF *)(aW + (uintptr_t)((size_t)(mm / kGroup) * BS * kGroup * sizeof(TF)));
#pragma unroll
                for (int i = 0; i < NV; ++i) {
                    g[WL::index(j, st * NV + i)] = zero ? TF(0) : (TF)col1[i];
                    wd[tbs::Soa64<BS, TF>::index(i + NV * j, mm % kGroup)] = (TF)col2[i];
                }
            }
        }
    };
    // Store the staged outputs of wave step sw (slot b) at the wave's window
    // position of the step, as 16-byte pieces of VW values (W: its element's
    // SoA-64 chunks; E, Gamma: the VW-entry chunks of the step's NV window
    // lanes) and then the leftover single values.  A step's stores are
    // issued after the next step's block loads: the wave's loads and stores
    // complete in order on one memory counter, and each step's set-up waits
    // for the whole counter, so stores issued at the end of a step put their
    // write latency on that wait (at 384,000 elements: 17x17 10.3 -> 8.2 ms,
    // 22x22 13.9 -> 13.2 ms; at 32x32 a step is long enough that the
    // staging's address arithmetic costs more than it saves, 27 -> 31 ms).
    constexpr int VW = 16 / (int)sizeof(TF);
    constexpr int NCW = BS / VW, REMW = BS - NCW * VW;   // W: pieces, leftover values
    constexpr int NCE = NV / VW, REME = NV - NCE * VW;   // one row of E or Gamma
    constexpr int NCEd = NCE > 0 ? NCE : 1, REMEd = REME > 0 ? REME : 1;   // divisors where a branch is unreached
    constexpr int P16 = NCW + 2 * NV * NCE, PR = REMW + 2 * NV * REME;
    auto flush = [&](int b, int sw) {
        typedef TF tf16 __attribute__((ext_vector_type(VW)));
        using gdst16 = __attribute__((address_space(1))) tf16 *;
        const int k = bot ? n - 1 - sw : sw;
        const int pos = bot ? sw + sp.pad : sw;
        const int w = pos / SPW, st = pos - w * SPW;
        const int mm = m0 + k;
        const TF *stg = (const TF *)&sblk[b][0];
        const size_t wg = (size_t)(mm / kGroup) * BS * kGroup;
        const long wofs = wbase + (long)w * WL::STRIDE;
        TF *const Wt = (TF *)aW, *const Et = (TF *)aEF, *const Gt = (TF *)(aEF + dGB);
        wsync();
        for (int r = lane; r < P16; r += 64) {
            int so;
            TF *d;
            if (r < NCW) {
                so = r * VW;
                d = Wt + wg + (size_t)r * 64 * VW + (mm % kGroup) * VW;
            } else {
                const int r1 = r - NCW, h = r1 / (NV * NCEd), r2 = r1 - h * NV * NCE, i = r2 / NCEd, c = r2 - i * NCE;
                so = (1 + h) * BS + i * NV + c * VW;
                d = (h ? Gt : Et) + wofs + (long)c * LW * VW + (long)(st * NV + i) * VW;
            }
            tf16 x;
#pragma unroll
            for (int q = 0; q < VW; ++q) x[q] = stg[so + q];
            *(gdst16)d = x;
        }
        for (int r = lane; r < PR; r += 64) {
            int so;
            TF *d;
            if (r < REMW) {
                so = NCW * VW + r;
                d = Wt + wg + (size_t)(NCW * VW + r) * 64 + mm % kGroup;
            } else {
                const int r1 = r - REMW, h = r1 / (NV * REMEd), r2 = r1 - h * NV * REME, i = r2 / REMEd, e = r2 - i * REME;
                so = (1 + h) * BS + i * NV + NCE * VW + e;
                d = (h ? Gt : Et) + wofs + (long)(NCE * VW + e) * LW + st * NV + i;
            }
            *(gdst)d = stg[so];
        }
        wsync();
    };
    // step m (top wave): wait for the bottom wave's last Gamma'' (the one
    // block-wide barrier of the two waves after the start), then D_m -=
    // C_{m+1} Gamma''_{m+1} in the step's copy of D_m
    auto middle = [&](int b) {
        __syncthreads();
        double *Dm = &sblk[b][0];
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
    // steps in pairs (parity 0, 1); at each step s > 0: step s into LDS,
    // step s+1's blocks and step s+2's sources into flight
    auto advance = [&](int s) {
        if (s > 0) {
            wsync();                            // LDS of step s-1 read
            store_blocks(s & 1, R);
            set_ptrs((s + 1) & 1, xnext);
            wsync();
            load_blocks((s + 1) & 1, R);
            xnext = load_idx(s + 2);
            if constexpr (kDefer) flush((s - 1) & 1, s - 1);   // outputs of step s-1
        }
    };
    for (int s = 0; s < nsteps; s += 2) {
        advance(s);
        if (!bot && s == sp.m) middle(s & 1);
        step(std::integral_constant<int, 0>(), s & 1, s);
        if (s + 1 < nsteps) {
            advance(s + 1);
            if (!bot && s + 1 == sp.m) middle((s + 1) & 1);
            step(std::integral_constant<int, 1>(), (s + 1) & 1, s + 1);
        }
    }
    if constexpr (kDefer) {
        if (nsteps > 0) flush((nsteps - 1) & 1, nsteps - 1);   // the last step's outputs
    }
    if (bot) {
        // Gamma''_{m+1}: the first columns of the non-pivot half of the last
        // step (half 1 after an even step, half 0 after an odd one); zeros
        // when the bottom side is empty (after the last step's outputs have
        // left its slot)
        wsync();
        const int xh = sp.B == 0 ? -1 : (((sp.B - 1) & 1) ? 0 : 1);
        if (xh < 0) {
            for (int t = lane; t < BS; t += 64) sG[t] = 0.0;
        } else if (half == xh) {
#pragma unroll
            for (int i = 0; i < NV; ++i) sG[NV * j + i] = col1[i];
        }
        __syncthreads();
    }
    // empty steps of the wave's windows (top: window q after step m; bottom:
    // the first pad steps of window 0), and each window's alignment padding
    const int e_lo = bot ? 0 : 1, e_hi = bot ? sp.pad : SPW;
    const int wlast = bot ? 0 : sp.q;
    if (half == 0) {
        for (int st = e_lo; st < e_hi; ++st) {
            const uintptr_t base = (uintptr_t)((wbase + (long)wlast * WL::STRIDE) * (long)sizeof(TF));
            const gdst e = (gdst)(TF *)(aEF + base), g = (gdst)(TF *)(aEF + dGB + base);
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                e[WL::index(j, st * NV + i)] = TF(0);
                g[WL::index(j, st * NV + i)] = TF(0);
            }
        }
        if (WL::STRIDE > WL::ELEMS && j < WL::STRIDE - WL::ELEMS) {
            const int nwin = bot ? sp.nb : sp.q + 1;
            for (int w = 0; w < nwin; ++w) {
                const long at = wbase + (long)w * WL::STRIDE + WL::ELEMS + j;
                ((gdst)(TF *)aEF)[at] = TF(0);
                ((gdst)(TF *)(aEF + dGB))[at] = TF(0);
            }
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

// The element and face tables of tbs_structured_mesh's NL x NK grid (same
// numbering: element e = l NK + k creates its in-line face (e+1, e), then its
// cross face (e+NK, e); boundary ids -(b+2) for S, E, N, W boundary entry b),
// one thread per element, from the closed form of the face count before e:
// F(l, k) = l (2 NK - 1) + k (1 + [l < NL-1]).
__global__ void structured_tables_kernel(int NL, int NK, int *__restrict__ lines, int *__restrict__ fc,
                                         int *__restrict__ lineface, int *__restrict__ epoint, int *__restrict__ ef) {
    const long n = (long)NL * NK;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int e = (int)t, l = e / NK, k = e - l * NK;
        auto f_in = [NL, NK](int l2, int k2) {        // face (e+1, e) of element (l2, k2), k2 < NK-1
            return l2 * (2 * NK - 1) + k2 * (1 + (l2 < NL - 1 ? 1 : 0));
        };
        auto f_cross = [NK, &f_in](int l2, int k2) {   // face (e+NK, e), l2 < NL-1
            return f_in(l2, k2) + (k2 < NK - 1 ? 1 : 0);
        };
        auto bnd = [](int b) { return -b - 2; };
        const int bS = 0, bE = NK, bN = NK + NL, bW = NK + NL + NK;
        if (k < NK - 1) {
            const int f = f_in(l, k);
            fc[2 * f] = e + 1;
            fc[2 * f + 1] = e;
        }
        if (l < NL - 1) {
            const int f = f_cross(l, k);
            fc[2 * f] = e + NK;
            fc[2 * f + 1] = e;
        }
        const int w = k > 0 ? f_in(l, k - 1) : bnd(bW + l);
        ef[4 * t + 0] = l > 0 ? f_cross(l - 1, k) : bnd(bS + k);
        ef[4 * t + 1] = k < NK - 1 ? f_in(l, k) : bnd(bE + l);
        ef[4 * t + 2] = l < NL - 1 ? f_cross(l, k) : bnd(bN + k);
        ef[4 * t + 3] = w;
        epoint[t] = 4 * e;
        if (t == n - 1) epoint[n] = 4 * (int)n;
        lines[t] = e;
        lineface[t] = w;
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
// The gather of an initial U, which also sets *nonzero when any value is
// not +-0 (NaN included): a solve from U = 0 then starts with R = rhs.
__global__ void gather_u0_kernel(long total, int nv, const int *__restrict__ lines,
                                 const double *__restrict__ src, double *__restrict__ dst, int *__restrict__ nonzero) {
    int any = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const long m = t / nv;
        const int i = (int)(t % nv);
        const double v = src[(size_t)nv * lines[m] + i];
        dst[t] = v;
        any |= !(v == 0.0);
    }
    if (__syncthreads_or(any) && threadIdx.x == 0) *nonzero = 1;
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


// Launch on st; with a start event, through hipExtLaunchKernelGGL, which
// records the two events at the kernel's own start and end (timing of exactly
// this kernel, without the dispatch or the completion signal around it).
template <typename K, typename... Args>
void launch_k(K kernel, dim3 grid, dim3 block, hipStream_t st, hipEvent_t e0, hipEvent_t e1, Args... args) {
    if (e0) hipExtLaunchKernelGGL(kernel, grid, block, 0, st, e0, e1, 0, args...);
    else hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
}

// Residual workgroup shapes the autotuner picks from: waves (64-element
// groups) per workgroup; index 0 is the default.  Only the headline block
// size has more than one.
template <int NV> struct ResidualV {
    static constexpr int n = 1;
    static constexpr int rw[1] = {4};
};
template <> struct ResidualV<9> {
    static constexpr int n = 4;
    static constexpr int rw[4] = {4, 2, 8, 1};
};

template <int NV, typename TF, int RW>
int launch_residual_rw(hipStream_t st, const void *sys, const int2 *gtab, int ng, int want_R, int force,
                       hipEvent_t e0, hipEvent_t e1) {
    const int nblk = (ng + RW - 1) / RW;
    launch_k(fres_kernel<NV, TF, RW>, dim3(8 * ((nblk + 7) / 8)), dim3(64 * RW), st, e0, e1,
             (const SysDev<TF> *)sys, gtab, ng, want_R, force);
    TBS_LAUNCHED("fres_kernel");
    return TBS_OK;
}

template <int NV, typename TF>
int launch_residual(hipStream_t st, const void *sys, const int2 *gtab, int ng, int want_R, int force,
                    hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, int rvariant = 0) {
    if (ng == 0) return TBS_OK;
    using P = ResidualV<NV>;
    if (rvariant < 0 || rvariant >= P::n) rvariant = 0;
    if constexpr (P::n == 1) {
        return launch_residual_rw<NV, TF, P::rw[0]>(st, sys, gtab, ng, want_R, force, e0, e1);
    } else {
        switch (rvariant) {
            case 1: return launch_residual_rw<NV, TF, P::rw[1]>(st, sys, gtab, ng, want_R, force, e0, e1);
            case 2: return launch_residual_rw<NV, TF, P::rw[2]>(st, sys, gtab, ng, want_R, force, e0, e1);
            case 3: return launch_residual_rw<NV, TF, P::rw[3]>(st, sys, gtab, ng, want_R, force, e0, e1);
            default: return launch_residual_rw<NV, TF, P::rw[0]>(st, sys, gtab, ng, want_R, force, e0, e1);
        }
    }
}

// Thomas variants the autotuner picks from: {prefetch windows PF, ring
// cycles K per loop iteration (the group of windows staged in LDS), lines per
// workgroup}; index 0 is the default.
struct TV { int pf, k, lpw; };
// (blocks above 16x16: one window prefetched, so the ring of NV-value rows
// stays within the registers)
template <int NV, typename TF> struct ThomasV {
    static constexpr int n = 1;
    static constexpr TV v[1] = {{NV > 16 ? 1 : 2, 2, 1}};
};
// (LDS per two-wave line stays <= 32 KiB so five lines fit a CU.)
template <> struct ThomasV<9, double> {
    static constexpr int n = 8;
    static constexpr TV v[8] = {{2, 2, 1}, {3, 1, 1}, {1, 3, 1}, {1, 2, 1},
                                {2, 2, 2}, {3, 1, 2}, {1, 3, 2}, {1, 2, 2}};
};
template <> struct ThomasV<9, float> {
    static constexpr int n = 8;
    static constexpr TV v[8] = {{3, 2, 1}, {2, 2, 1}, {5, 1, 1}, {3, 1, 1},
                                {3, 2, 2}, {2, 2, 2}, {5, 1, 2}, {3, 1, 2}};
};

template <int NV, typename TF, int PF, int K, int LPW>
int launch_thomas_pf(hipStream_t st, const void *sys, const int4 *ltab, int nl, hipEvent_t e0, hipEvent_t e1) {
    launch_k(fthomas_kernel<NV, TF, PF, K, LPW>, dim3((nl + LPW - 1) / LPW), dim3(128 * LPW), st, e0, e1,
             (const SysDev<TF> *)sys, ltab, nl);
    TBS_LAUNCHED("fthomas_kernel");
    return TBS_OK;
}

template <int NV, typename TF, int I>
int launch_thomas_v(hipStream_t st, const void *sys, const int4 *ltab, int nl, hipEvent_t e0, hipEvent_t e1) {
    using P = ThomasV<NV, TF>;
    constexpr TV t = P::v[I < P::n ? I : 0];
    return launch_thomas_pf<NV, TF, t.pf, t.k, t.lpw>(st, sys, ltab, nl, e0, e1);
}

template <int NV, typename TF>
int launch_thomas(hipStream_t st, const void *sys, const int4 *ltab, int nl, int variant, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr) {
    if (nl == 0) return TBS_OK;
    using P = ThomasV<NV, TF>;
    if (variant < 0 || variant >= P::n) variant = 0;
    if constexpr (P::n == 1) {
        return launch_thomas_v<NV, TF, 0>(st, sys, ltab, nl, e0, e1);
    } else {
        switch (variant) {
            case 1: return launch_thomas_v<NV, TF, 1>(st, sys, ltab, nl, e0, e1);
            case 2: return launch_thomas_v<NV, TF, 2>(st, sys, ltab, nl, e0, e1);
            case 3: return launch_thomas_v<NV, TF, 3>(st, sys, ltab, nl, e0, e1);
            case 4: return launch_thomas_v<NV, TF, 4>(st, sys, ltab, nl, e0, e1);
            case 5: return launch_thomas_v<NV, TF, 5>(st, sys, ltab, nl, e0, e1);
            case 6: return launch_thomas_v<NV, TF, 6>(st, sys, ltab, nl, e0, e1);
            case 7: return launch_thomas_v<NV, TF, 7>(st, sys, ltab, nl, e0, e1);
            default: return launch_thomas_v<NV, TF, 0>(st, sys, ltab, nl, e0, e1);
        }
    }
}

template <typename TF>
int launch_final(hipStream_t st, const void *sys, int nsys, int max_sweeps, double rtol, int force,
                 hipEvent_t e0, hipEvent_t e1) {
    launch_k(ffinal_kernel<TF>, dim3(nsys), dim3(kFinalThreads), st, e0, e1, (const SysDev<TF> *)sys, nsys,
             max_sweeps, rtol, force);
    TBS_LAUNCHED("ffinal_kernel");
    return TBS_OK;
}

// One lane per column of [S | C | A | I] up to NV = 16, two columns per lane above.
template <int NV, typename TF>
void launch_factor_nv(hipStream_t st, const FactorRec &one, int nline, const FactorRec *recs, const int2 *ltab) {
    if constexpr (4 * NV <= 64)
        hipLaunchKernelGGL((ffactor_kernel<NV, TF>), dim3(nline), dim3(128), 0, st, one, nline, recs, ltab);
    else
        hipLaunchKernelGGL((ffactor_wide_kernel<NV, TF>), dim3(nline), dim3(128), 0, st, one, nline, recs, ltab);
}

// The LDS packing holds 16 whole blocks per workgroup (32 KiB at NV = 16);
// larger blocks take the direct packing.
template <int NV>
void launch_pack_lds(hipStream_t pst, long nws, int S, const int2 *src, const double *jacD, const double *jacO1,
                     const double *jacO2, double *blk) {
    if constexpr (NV <= 16) {
        const long units = (long)(kGroup / kPackLanes) * nws;
        hipLaunchKernelGGL(pack_lds_kernel<NV>, dim3((unsigned)std::min<long>(units, 1L << 24)), dim3(64), 0, pst, units,
                           S, src, jacD, jacO1, jacO2, blk);
    }
}

}  // namespace
