// ref_kernels.hip -- drop-in HIP kernels for the reference's active OKL
// kernels, operating on the reference's own layouts (column-major NV x NV
// blocks by element / face id, vectors by element id).
//
// Mapping to gfx950: the reference launches 9-lane workgroups and runs its
// triangular solves in one lane.  Here one wave64 owns a line; lane i < NV
// holds block row i in registers and the values the other rows need are
// broadcast with v_readlane (uniform SGPR operands), so no LDS round trip
// sits on the per-step dependency chain.  Every block entry is computed with
// the reference's sequence of operations (same accumulation order, multiply
// by the reciprocal pivot in LU, divide by the U diagonal in solveLU); the
// compiler contracts a*b+c into FMA as the reference's HIP backend does.
#include "tbs_common.hpp"
#include "lane_ops.hpp"

#include <algorithm>
#include <climits>
#include <mutex>

namespace {

using tbs::rdlane;

// Row-parallel Doolittle LU (src/okl/lineLU_v2.okl:17-27): lane i holds row
// i of the block in r[]; on exit r[] holds row i of the packed L\U factors.
template <int NV>
__device__ __forceinline__ void lu_rows(double (&r)[NV], int i) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double rowj[NV];
#pragma unroll
        for (int k = j; k < NV; ++k) rowj[k] = rdlane(r[k], j);
        const double pivot = 1.0 / rowj[j];
        if (i > j) {
            r[j] = r[j] * pivot;
#pragma unroll
            for (int k = j + 1; k < NV; ++k) r[k] -= r[j] * rowj[k];
        }
    }
}

// Row-parallel solveLU (src/okl/linesmoothLU_TriBlock_v3.okl:23-45): lane i
// holds row i of the L\U factors in a[] and b_i; returns x_i.  Sums run in
// the reference's ascending-j order.
template <int NV>
__device__ __forceinline__ double solve_rows(const double (&a)[NV], double b, int i) {
    double y[NV];
    double tot = 0.0;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        y[j] = rdlane(b - tot, j);           // lane j: b_j - sum_{j'<j} L(j,j') y_j'
        if (i > j) tot += a[j] * y[j];
    }
    double diag = a[0];
#pragma unroll
    for (int j = 1; j < NV; ++j) diag = (i == j) ? a[j] : diag;
    double x[NV];
    double xi = 0.0;
#pragma unroll
    for (int j = NV - 1; j >= 0; --j) {
        double t = 0.0;
#pragma unroll
        for (int jj = j + 1; jj < NV; ++jj) t += a[jj] * x[jj];
        const double v = (y[j] - t) / diag;
        x[j] = rdlane(v, j);
        xi = (i == j) ? x[j] : xi;
    }
    return xi;
}

// ---------------------------------------------------------------- lineLU --
// src/okl/lineLU_v2.okl:58-182.  One wave per line.
template <int NV>
__global__ __launch_bounds__(64) void lineLU_kernel(int nlines, const int *__restrict__ fc,
                                                    const int *__restrict__ linesize,
                                                    const int *__restrict__ linepoint,
                                                    const int *__restrict__ lines,
                                                    const int *__restrict__ lineface,
                                                    double *__restrict__ Dia,
                                                    const double *__restrict__ Of1,
                                                    const double *__restrict__ Of2,
                                                    double *__restrict__ DinvC) {
    constexpr int BS = NV * NV;
    __shared__ double sDl[BS];   // D'_{k-1} (packed L\U)
    __shared__ double sC[BS];
    __shared__ double sA[BS];
    __shared__ double sG[BS];    // Gamma = D'_{k-1}^{-1} C
    __shared__ double sDe[BS];   // raw D_k
    __shared__ int sIe[64], sIf[64], sIo[64];   // 64 steps' element, face, A from Of2
    const int l = blockIdx.x;
    if (l >= nlines) return;
    const int t = threadIdx.x;
    const int i = t;             // block row for t < NV
    const int n = linesize[l];
    const int m0 = linepoint[l];
    if (n <= 0) return;

    // Software pipeline: step k+1's blocks are loaded into registers while
    // step k computes, issued before step k's stores, and the line's indices
    // come 64 steps at a time through LDS.  The wave's loads and stores
    // complete in order on one memory counter, so a load issued after a
    // step's stores would make the next step wait for that store traffic as
    // well as for its own load latency.
    auto idx_chunk = [&](int c) {
        const int kk = c * 64 + t;
        if (kk < n) {
            const int e = lines[m0 + kk];
            const int f = kk > 0 ? lineface[m0 + kk] : 0;   // (step 0 has no face before it)
            sIe[t] = e;
            sIf[t] = f;
            sIo[t] = kk > 0 && e == fc[2 * f];             // :116-117
        }
        __syncthreads();
    };
    constexpr int NQ = (BS + 63) / 64;   // block entries per lane
    double pC[NQ], pA[NQ], pD[NQ];
    auto blocks = [&](int e, int f, bool o2) {
        const double *Ap = o2 ? Of2 : Of1;                           // :116
        const double *Cp = o2 ? Of1 : Of2;                           // :117
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
            const int q = t + 64 * u;
            if (q < BS) {
                pC[u] = Cp[(size_t)BS * f + q];
                pA[u] = Ap[(size_t)BS * f + q];
                pD[u] = Dia[(size_t)BS * e + q];
            }
        }
    };
    idx_chunk(0);
    int ne = 0;                  // element of the step whose blocks are in pC, pA, pD
    int elast = sIe[0];
    if (n > 1) {
        ne = sIe[1];
        blocks(ne, sIf[1], sIo[1] != 0);
    }

    double r[NV];
    {
        const int e0 = elast;
#pragma unroll
        for (int k = 0; k < NV; ++k) r[k] = (i < NV) ? Dia[(size_t)BS * e0 + i + NV * k] : 1.0;
        lu_rows<NV>(r, i);                                           // :91
        if (i < NV) {
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                Dia[(size_t)BS * e0 + i + NV * k] = r[k];
                sDl[i + NV * k] = r[k];
            }
        }
    }
    __syncthreads();
    for (int k = 1; k < n; ++k) {
        const int e = ne;
#pragma unroll
        for (int u = 0; u < NQ; ++u) {                               // :119-139
            const int q = t + 64 * u;
            if (q < BS) {
                sC[q] = pC[u];
                sA[q] = pA[u];
                sDe[q] = pD[u];
            }
        }
        if (k + 1 < n) {                                             // step k+1's blocks
            if ((k + 1) % 64 == 0) idx_chunk((k + 1) / 64);
            const int kn = (k + 1) % 64;
            ne = sIe[kn];
            blocks(ne, sIf[kn], sIo[kn] != 0);
        }
        __syncthreads();
        if (t < NV) {                                                // :142-144
            const int c = t;
            double y[NV], x[NV];
#pragma unroll
            for (int ii = 0; ii < NV; ++ii) {
                double tot = 0.0;
#pragma unroll
                for (int j = 0; j < ii; ++j) tot += sDl[ii + NV * j] * y[j];
                y[ii] = sC[ii + NV * c] - tot;
            }
#pragma unroll
            for (int ii = NV - 1; ii >= 0; --ii) {
                double tot = 0.0;
#pragma unroll
                for (int j = ii + 1; j < NV; ++j) tot += sDl[ii + NV * j] * x[j];
                x[ii] = (y[ii] - tot) / sDl[ii + NV * ii];
            }
#pragma unroll
            for (int ii = 0; ii < NV; ++ii) sG[ii + NV * c] = x[ii];
        }
        __syncthreads();
        // Alpha = A Gamma, D - Alpha (:147-162), then LU (:165)
#pragma unroll
        for (int kk = 0; kk < NV; ++kk) {
            double tot = 0.0;
            if (i < NV) {
#pragma unroll
                for (int j = 0; j < NV; ++j) tot += sA[i + NV * j] * sG[j + NV * kk];
                r[kk] = sDe[i + NV * kk] - tot;
            } else {
                r[kk] = 1.0;
            }
        }
        lu_rows<NV>(r, i);
        __syncthreads();                                             // sDl reads done
        if (i < NV) {                                                // :169-178
#pragma unroll
            for (int kk = 0; kk < NV; ++kk) {
                Dia[(size_t)BS * e + i + NV * kk] = r[kk];
                sDl[i + NV * kk] = r[kk];
            }
        }
        for (int q = t; q < BS; q += 64) DinvC[(size_t)BS * elast + q] = sG[q];
        elast = e;
        __syncthreads();
    }
}

// -------------------------------------------------------------- solveDU --
// src/okl/linesmoothLU_TriBlock_v3.okl:92-226: forward dU_k = D'_k^-1 (-R_k -
// A_k dU_{k-1}) with D'_k's packed LU factors, then dU_k -= DinvC_k dU_{k+1}.
// Three waves per line, in chunks of G line steps:
//   - two producer waves take the chunks in turn (producer p the chunks
//     c with c % 2 == p, in chunk buffer p).  A producer keeps its next
//     chunk's loads in flight while the other producer's chunk is worked on,
//     so two chunks per line are in flight.  Forward it stages D', A and R
//     (64 / G lanes per step: a load reads one array, so a lane's address is
//     its step's block plus a constant) and solves lane-parallel, in place,
//     for every step s and every column c: D'_s w = A_s(:,c) (c < NV: E_s =
//     D'^-1 A) or D'_s w = -R_s (c = NV: g_s), with the reference's forward
//     and back substitution on the packed factors (the back substitution
//     multiplies by the reciprocal of the U pivot where the reference
//     divides).  Backward it stages DinvC and the forward dU;
//   - the consumer wave runs the chain one chunk behind: lane i = row i,
//     forward dU_k = g_k - E_k dU_{k-1}, backward dU_k -= DinvC_k dU_{k+1}
//     (ascending j, as the reference), each with the previous step's vector
//     broadcast from the lanes that produced it.
// One workgroup barrier per chunk hands a chunk to the consumer.  E_s is
// column-major like the reference's blocks.  The forward step is the
// reference's with its two solves distributed over the right-hand sides; the
// rounding differs by O(eps * cond(D')).  The backward step is the
// reference's arithmetic.
template <int NV>
struct DUChunk {
    static constexpr int BS = NV * NV;
    static constexpr int SF = 2 * BS + NV;   // per step, forward: D', A then E, R then g
    static constexpr int SX = BS + NV;       // per step, backward: DinvC, dU
    // steps per chunk (a power of two up to 8): the two chunk buffers and
    // the reciprocal pivots within about 32 KiB, so five lines fit a CU
    static constexpr int G0 = 4096 / (2 * SF + NV);
    static constexpr int G = G0 >= 8 ? 8 : (G0 >= 4 ? 4 : (G0 >= 2 ? 2 : 1));
    static constexpr int LPS = 64 / G;               // producer lanes per step
    static constexpr int NM = (BS + LPS - 1) / LPS;  // loads per lane of a step's block
    static constexpr int NVL = (NV + LPS - 1) / LPS; // of a step's vector
    static constexpr int KC = (NV + 1 + LPS - 1) / LPS;   // right-hand sides per producer lane
};

template <int NV>
__global__ __launch_bounds__(192) void solveDU_kernel(int nlines, const int *__restrict__ linesize,
                                                     const int *__restrict__ linepoint,
                                                     const int *__restrict__ lines,
                                                     const double *__restrict__ Dia,
                                                     const double *__restrict__ DinvC,
                                                     const double *__restrict__ A,
                                                     double *__restrict__ dU,
                                                     const double *__restrict__ R) {
    using C = DUChunk<NV>;
    constexpr int BS = NV * NV, SF = C::SF, SX = C::SX, G = C::G, LPS = C::LPS, KC = C::KC;
    __shared__ double sbuf[2][G * SF];     // chunk buffers, one per producer
    __shared__ double srd[G * NV];         // 1 / U(i,i) of each step's D' (one producer at a time)
    __shared__ int pse[2][G];              // producer p: element indices of its chunk in flight
    __shared__ int se[2][G];               // element indices of the chunk in each buffer
    const int l = blockIdx.x;
    if (l >= nlines) return;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // wave-uniform: the step loops must be scalar loops (the chain
    // broadcasts read lanes that every step writes)
    const int n = __builtin_amdgcn_readfirstlane(linesize[l]);
    const int m0 = __builtin_amdgcn_readfirstlane(linepoint[l]);
    if (n <= 0) return;
    auto wsync = []() { __builtin_amdgcn_wave_barrier(); };
    const int ncf = (n + G - 1) / G;
    const int nb = n - 1;                               // backward steps: n-2 .. 0
    const int ncb = (nb + G - 1) / G;
    auto bwd_k = [&](int c, int s) { return nb - 1 - (c * G + s); };

    if (wave > 0) {
        // ================================================ producers ==
        const int p = wave - 1;
        double *const buf = sbuf[p];
        // element index of line step k, clamped into the line: every load is
        // unconditional and in bounds (a select after a load would make the
        // loop wait for the loads in flight); steps past a chunk's end are
        // never used
        auto elem = [&](int k) -> int { return lines[m0 + min(max(k, 0), n - 1)]; };
        const int gl = min(lane, G - 1);
        // global-space loads count only against vmcnt, so LDS waits do not
        // wait for the loads in flight
        using gptr = const __attribute__((address_space(1))) double *;
        const gptr gD = (gptr)Dia, gA = (gptr)A, gR = (gptr)R, gC = (gptr)DinvC, gU = (gptr)dU;
        // lane (ls, lq) stages entries lq, lq + LPS, ... of step ls's block
        // and vector (the last of them clamped into the block) and solves
        // for its right-hand sides lq, lq + LPS, ... (the last clamped to NV)
        const int ls = lane / LPS, lq = lane - ls * LPS;
        auto mo = [&](int r) { return (r + 1) * LPS <= BS ? r * LPS : min(r * LPS, BS - 1 - lq); };
        auto vo = [&](int r) { return (r + 1) * LPS <= NV ? r * LPS : min(r * LPS, NV - 1 - lq); };
        double Rm[C::NM], Rm2[C::NM], Rv[C::NVL];

        // ---- forward: chunk c (c % 2 == p) is staged and solved in the
        // interval before barrier c, while chunk c+2's loads are issued
        auto fwd_load = [&]() {
            const size_t e = (unsigned)pse[p][ls];
            const gptr pD = gD + e * BS + lq, pA = gA + e * BS + lq, pR = gR + e * NV + lq;
#pragma unroll
            for (int r = 0; r < C::NM; ++r) { Rm[r] = pD[mo(r)]; Rm2[r] = pA[mo(r)]; }
#pragma unroll
            for (int r = 0; r < C::NVL; ++r) Rv[r] = pR[vo(r)];
        };
        if (lane < G) pse[p][lane] = elem(p * G + lane);
        wsync();
        fwd_load();
        int ei = elem((p + 2) * G + gl);
        for (int c = 0; c < ncf; ++c) {
            if ((c & 1) == p) {
                {                                                // chunk c into LDS
                    double *sb = &buf[ls * SF + lq];
#pragma unroll
                    for (int r = 0; r < C::NM; ++r)
                        if (r * LPS + lq < BS) { sb[r * LPS] = Rm[r]; sb[BS + r * LPS] = Rm2[r]; }
#pragma unroll
                    for (int r = 0; r < C::NVL; ++r)
                        if (r * LPS + lq < NV) sb[2 * BS + r * LPS] = Rv[r];
                }
                if (lane < G) se[p][lane] = pse[p][lane];
                wsync();
                if (lane < G) pse[p][lane] = ei;                 // chunk c+2's indices
                wsync();
                ei = elem((c + 4) * G + gl);                     // ahead of the data loads
                const int kend = min(G, n - c * G);
                for (int t = lane; t < kend * NV; t += 64) {
                    const int s = t / NV, ii = t - s * NV;
                    srd[t] = 1.0 / buf[s * SF + ii + NV * ii];
                }
                wsync();
                if (ls < kend) {
                    const int s = ls;
                    const double *LU = &buf[s * SF];
                    // right-hand side k: column cc of A_s, or R_s (cc = NV)
                    // right after A_s; overwritten by the solution (column
                    // cc of E_s, or g_s).  The factor entries are read once
                    // for all of the lane's right-hand sides.
                    double *X[KC];
                    double w[KC][NV];
#pragma unroll
                    for (int k = 0; k < KC; ++k) {
                        const int cc = min(lq + k * LPS, NV);
                        X[k] = &buf[s * SF + BS + NV * cc];
                        const double sgn = cc < NV ? 1.0 : -1.0;
#pragma unroll
                        for (int ii = 0; ii < NV; ++ii) w[k][ii] = sgn * X[k][ii];
                    }
                    // forward: y_i = b_i - sum_{j<i} L(i,j) y_j (ascending j)
#pragma unroll
                    for (int ii = 1; ii < NV; ++ii) {
                        double tot[KC];
#pragma unroll
                        for (int k = 0; k < KC; ++k) tot[k] = 0.0;
#pragma unroll
                        for (int j = 0; j < ii; ++j) {
                            const double lij = LU[ii + NV * j];
#pragma unroll
                            for (int k = 0; k < KC; ++k) tot[k] += lij * w[k][j];
                        }
#pragma unroll
                        for (int k = 0; k < KC; ++k) w[k][ii] -= tot[k];
                        // a row at a time: the factor reads are not all
                        // hoisted into registers
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    // back: x_i = (y_i - sum_{j>i} U(i,j) x_j) / U(i,i)
#pragma unroll
                    for (int ii = NV - 1; ii >= 0; --ii) {
                        double tot[KC];
#pragma unroll
                        for (int k = 0; k < KC; ++k) tot[k] = 0.0;
#pragma unroll
                        for (int j = ii + 1; j < NV; ++j) {
                            const double uij = LU[ii + NV * j];
#pragma unroll
                            for (int k = 0; k < KC; ++k) tot[k] += uij * w[k][j];
                        }
                        const double rd = srd[s * NV + ii];
#pragma unroll
                        for (int k = 0; k < KC; ++k) w[k][ii] = (w[k][ii] - tot[k]) * rd;
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    // (lanes whose last right-hand side is clamped to g_s
                    // store the same g_s as the lane that owns it: the same
                    // operations on the same data; an unconditional store
                    // keeps the solve one straight block)
#pragma unroll
                    for (int k = 0; k < KC; ++k) {
#pragma unroll
                        for (int ii = 0; ii < NV; ++ii) X[k][ii] = w[k][ii];
                    }
                }
                // chunk c+2 into flight, after the solve so that its
                // registers are free during it (past the line's end the
                // clamped indices load the last step again: unconditional
                // loads keep the waits for the loads in flight exact)
                fwd_load();
            }
            __syncthreads();                                     // chunk c to the consumer
        }
        if (nb <= 0) return;
        __syncthreads();                                         // the forward dU, all written

        // ---- backward: chunk c holds steps n-2-cG .. n-2-cG-G+1; each step
        // hands over DinvC(e) and the forward dU(e)
        auto bwd_load = [&]() {
            const size_t e = (unsigned)pse[p][ls];
            const gptr pC = gC + e * BS + lq, pU = gU + e * NV + lq;
#pragma unroll
            for (int r = 0; r < C::NM; ++r) Rm[r] = pC[mo(r)];
#pragma unroll
            for (int r = 0; r < C::NVL; ++r) Rv[r] = pU[vo(r)];
        };
        if (lane < G) pse[p][lane] = elem(bwd_k(p, lane));
        wsync();
        bwd_load();
        ei = elem(bwd_k(p + 2, gl));
        for (int c = 0; c < ncb; ++c) {
            if ((c & 1) == p) {
                {
                    double *sb = &buf[ls * SX + lq];
#pragma unroll
                    for (int r = 0; r < C::NM; ++r)
                        if (r * LPS + lq < BS) sb[r * LPS] = Rm[r];
#pragma unroll
                    for (int r = 0; r < C::NVL; ++r)
                        if (r * LPS + lq < NV) sb[BS + r * LPS] = Rv[r];
                }
                if (lane < G) se[p][lane] = pse[p][lane];
                wsync();
                if (lane < G) pse[p][lane] = ei;
                wsync();
                ei = elem(bwd_k(c + 4, gl));
                bwd_load();                                      // chunk c+2 into flight
            }
            __syncthreads();
        }
        return;
    }

    // ===================================================== consumer ==
    const int i = lane < NV ? lane : 0;    // lanes >= NV shadow row 0 (results unused)
    double prev[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) prev[j] = 0.0;
    // Two register sets take turns: a step's row of its block, its vector
    // entry and its element index are read from LDS while the step before
    // runs, so the chain waits only on itself.  The reads are unconditional
    // (the step index clamped into the chunk), which keeps the waits for
    // them exact.
    double xa[NV + 1], xb[NV + 1];
    int ea, eb;
    auto ld = [&](double (&v)[NV + 1], int &ev, const double *X, int stride, const int *ix, int s) {
#pragma unroll
        for (int j = 0; j <= NV; ++j) v[j] = X[s * stride + NV * j];
        ev = ix[s];
    };
    // forward: dU_k = g_k - E_k dU_{k-1}, three partial sums over j
    auto fstep = [&](const double (&X)[NV + 1], int ev, bool first) {
        const size_t e = (size_t)__builtin_amdgcn_readfirstlane(ev);
        double x = X[NV];
        if (!first) {
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
            for (int j = 0; j < NV; j += 3) {
                t0 += X[j] * prev[j];
                if (j + 1 < NV) t1 += X[j + 1] * prev[j + 1];
                if (j + 2 < NV) t2 += X[j + 2] * prev[j + 2];
            }
            x -= (t0 + t1) + t2;
        }
        if (lane < NV) dU[NV * e + lane] = x;                    // :157, :194
        tbs::bcast_rows<NV>(x, prev);
    };
    // backward: dU_k -= DinvC_k dU_{k+1} (ascending j)
    auto bstep = [&](const double (&X)[NV + 1], int ev) {
        const size_t e = (size_t)__builtin_amdgcn_readfirstlane(ev);
        double tot = 0.0;
#pragma unroll
        for (int j = 0; j < NV; ++j) tot += X[j] * prev[j];      // prev = dU(elast)
        const double v = X[NV] - tot;
        if (lane < NV) dU[NV * e + lane] = v;
        tbs::bcast_rows<NV>(v, prev);
    };
    for (int c = 0; c < ncf; ++c) {
        const int cur = c & 1;
        __syncthreads();                                         // chunk c from its producer
        const int kend = min(G, n - c * G), kl = kend - 1;
        const double *Xc = &sbuf[cur][BS + i];                   // E_s row i, then g_s(i)
        ld(xa, ea, Xc, SF, se[cur], 0);
#pragma unroll 1
        for (int s = 0; s < kend; s += 2) {
            ld(xb, eb, Xc, SF, se[cur], min(s + 1, kl));
            fstep(xa, ea, c + s == 0);
            if (s + 1 >= kend) break;
            ld(xa, ea, Xc, SF, se[cur], min(s + 2, kl));
            fstep(xb, eb, false);
        }
    }
    if (nb <= 0) return;
    __syncthreads();                                             // the forward dU, all written
    for (int c = 0; c < ncb; ++c) {                              // :199-223
        const int cur = c & 1;
        __syncthreads();
        const int kend = min(G, nb - c * G), kl = kend - 1;
        const double *Xc = &sbuf[cur][i];                        // DinvC_s row i, then dU_s(i)
        ld(xa, ea, Xc, SX, se[cur], 0);
#pragma unroll 1
        for (int s = 0; s < kend; s += 2) {
            ld(xb, eb, Xc, SX, se[cur], min(s + 1, kl));
            bstep(xa, ea);
            if (s + 1 >= kend) break;
            ld(xa, ea, Xc, SX, se[cur], min(s + 2, kl));
            bstep(xb, eb);
        }
    }
}

// -------------------------------------------------------------- lineRes --
// src/okl/linesmoothLU_TriBlock_v3.okl:228-336: R_e += D_e U_e + the sum over
// the element's faces (ef order, boundary entries skipped) of the face block
// times the neighbour's U.  A wave takes W = 64 / NV consecutive positions of
// a line at a time; lane (s, i) = (position, row).  The work of a position is
// a sequence of rounds: round 0 is the diagonal block, round g >= 1 the block
// of face slot g-1.  The W blocks of a round are read whole by the wave (64
// consecutive doubles of the round's blocks per load) and staged through the
// wave's LDS buffer with the round's U vectors; RB rounds have their loads
// issued before the first is used.  The sums run in the reference's order.
template <int NV>
struct ResTile {
    static constexpr int BS = NV * NV;
    static constexpr int W = 64 / NV;                 // line positions per wave
    static constexpr int NB = (W * BS + 63) / 64;     // loads per lane of a round's blocks
    static constexpr int RB0 = 48 / NB;               // about 48 doubles of loads per lane in flight
    static constexpr int RB = RB0 < 1 ? 1 : (RB0 > 5 ? 5 : RB0);
};

template <int NV>
__global__ __launch_bounds__(256) void lineRes_kernel(int nlines, const int *__restrict__ epoint,
                                                      const int *__restrict__ ef, const int *__restrict__ fc,
                                                      const int *__restrict__ linesize,
                                                      const int *__restrict__ linepoint,
                                                      const int *__restrict__ lines,
                                                      const double *__restrict__ Dia,
                                                      const double *__restrict__ Of1,
                                                      const double *__restrict__ Of2,
                                                      const double *__restrict__ U, double *__restrict__ R) {
    using T = ResTile<NV>;
    constexpr int BS = NV * NV, W = T::W, NB = T::NB, RB = T::RB;
    __shared__ double sblk[4][W * BS];    // per wave: a round's blocks
    __shared__ double su[4][W * NV];      // and the U vectors they multiply
    __shared__ uintptr_t sptr[4][RB][W];  // per round and position: block address,
    __shared__ int snb[4][RB][W];         // U owner,
    __shared__ int sval[4][RB][W];        // whether the round contributes
    __shared__ int se[4][W], sep[4][W], snf[4][W];
    using gptr = const __attribute__((address_space(1))) double *;
    auto wsync = []() { __builtin_amdgcn_wave_barrier(); };
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int ps = lane / NV, i = lane - ps * NV;      // the lane's position and row
    const bool act = lane < W * NV;
    for (int l = blockIdx.y; l < nlines; l += gridDim.y) {
        const int n = __builtin_amdgcn_readfirstlane(linesize[l]);
        const int m0 = __builtin_amdgcn_readfirstlane(linepoint[l]);
        const int nch = (n + W - 1) / W;
        for (int c = blockIdx.x * 4 + wave; c < nch; c += gridDim.x * 4) {
            const int nval = min(W, n - c * W);          // positions of this chunk
            if (lane < W) {
                const int e = lines[m0 + c * W + min(lane, nval - 1)];
                const int ep = epoint[e];
                se[wave][lane] = e;
                sep[wave][lane] = ep;
                snf[wave][lane] = lane < nval ? epoint[e + 1] - ep : 0;
            }
            wsync();
            int maxnf = 0;
#pragma unroll
            for (int s = 0; s < W; ++s) maxnf = max(maxnf, snf[wave][s]);
            maxnf = __builtin_amdgcn_readfirstlane(maxnf);
            double acc = 0.0;
            for (int g0 = 0; g0 <= maxnf; g0 += RB) {
                // the rounds' tables (a skipped slot points at the diagonal
                // block, so every load below stays in bounds)
                for (int t = lane; t < RB * W; t += 64) {
                    const int r = t / W, s = t - r * W, g = g0 + r;
                    const int e = se[wave][s];
                    uintptr_t p = (uintptr_t)(Dia + (size_t)BS * e);
                    int nb = e, ok = (g == 0 && s < nval) ? 1 : 0;
                    if (g >= 1 && g - 1 < snf[wave][s]) {
                        const int f = ef[sep[wave][s] + g - 1];                      // :299-328
                        if (f >= 0) {
                            const int e1 = fc[2 * f + 0], e2 = fc[2 * f + 1];
                            p = (uintptr_t)(((e == e1) ? Of2 : Of1) + (size_t)BS * f);
                            nb = (e == e1) ? e2 : e1;
                            ok = 1;
                        }
                    }
                    sptr[wave][r][s] = p;
                    snb[wave][r][s] = nb;
                    sval[wave][r][s] = ok;
                }
                wsync();
                double rb[RB][NB], ru[RB];
#pragma unroll
                for (int r = 0; r < RB; ++r) {
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        const int t = min(lane + 64 * j, W * BS - 1);
                        const int s = t / BS, q = t - s * BS;
                        rb[r][j] = ((gptr)sptr[wave][r][s])[q];
                    }
                    const int t = min(lane, W * NV - 1);
                    const int s = t / NV, q = t - s * NV;
                    ru[r] = ((gptr)U)[(size_t)NV * snb[wave][r][s] + q];
                }
#pragma unroll
                for (int r = 0; r < RB; ++r) {
                    if (g0 + r > maxnf) break;
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        const int t = lane + 64 * j;
                        if (t < W * BS) sblk[wave][t] = rb[r][j];
                    }
                    if (act) su[wave][lane] = ru[r];
                    wsync();
                    if (act && sval[wave][r][ps]) {
                        double t2 = 0.0;                                             // :276-293, :306-327
#pragma unroll
                        for (int j = 0; j < NV; ++j) t2 += sblk[wave][ps * BS + i + NV * j] * su[wave][ps * NV + j];
                        acc = (g0 + r == 0) ? t2 : acc + t2;
                    }
                    wsync();
                }
            }
            if (act && ps < nval) R[(size_t)NV * se[wave][ps] + i] += acc;      // :331
            wsync();
        }
    }
}

// -------------------------------------------------------------- dgemvLU --
// src/okl/linesmoothLU_TriBlock_v3.okl:48-69 (an inline helper of the
// reference, applied here to every element's block): b = L U x with the
// packed factors of LU (unit-lower L strictly below the diagonal, U on and
// above it).  One thread per (element, row); a workgroup holds whole
// elements, so the U x products of an element's rows meet in LDS before the
// L pass.  Both sums run in the reference's order: U from the diagonal
// upward in j, then b_i = y_i plus L(i,j) y_j for ascending j < i.
template <int NV>
__global__ __launch_bounds__(256) void dgemvLU_kernel(long nelem, const double *__restrict__ LU,
                                                      const double *__restrict__ x, double *__restrict__ b) {
    constexpr int BS = NV * NV;
    constexpr int EPB = 256 / NV;                 // elements per workgroup
    __shared__ double sy[EPB * NV];
    const int kl = threadIdx.x / NV, i = threadIdx.x - kl * NV;
    const bool lane_ok = kl < EPB;
    for (long e0 = (long)blockIdx.x * EPB; e0 < nelem; e0 += (long)gridDim.x * EPB) {
        const long e = e0 + kl;
        const bool ok = lane_ok && e < nelem;
        const double *A = LU + (size_t)BS * (ok ? e : 0);
        const double *xe = x + (size_t)NV * (ok ? e : 0);
        double y = 0.0;
        if (ok) {
#pragma unroll
            for (int j = 0; j < NV; ++j)
                if (j >= i) y += A[i + NV * j] * xe[j];
            sy[kl * NV + i] = y;
        }
        __syncthreads();
        if (ok) {
            double bi = y;
#pragma unroll
            for (int j = 0; j < NV; ++j)
                if (j < i) bi += A[i + NV * j] * sy[kl * NV + j];
            b[(size_t)NV * e + i] = bi;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------ assembleTriBlocks --
// src/Jacobian.cxx:76-117 as a device gather; grid.y = line.
template <int NV>
__global__ void assemble_kernel(int nline, const int *__restrict__ linesize,
                                const int *__restrict__ linepoint,
                                const int *__restrict__ lines,
                                const int *__restrict__ lineface, const int *__restrict__ fc,
                                const double *__restrict__ O1, const double *__restrict__ O2,
                                double *__restrict__ A) {
    constexpr int BS = NV * NV;
    for (int l = blockIdx.y; l < nline; l += gridDim.y) {
    const int n = linesize[l];
    const int m0 = linepoint[l];
    const long total = (long)(n - 1) * BS;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const int k = 1 + (int)(idx / BS);
        const int q = (int)(idx % BS);
        const int m = m0 + k;
        const int f = lineface[m];
        const int e = lines[m];
        const double *src = (e == fc[2 * f]) ? O2 : O1;
        A[(size_t)BS * e + q] = src[(size_t)BS * f + q];
    }
    }
}

// -------------------------------------------------------- copy / add --
__global__ void copy_kernel(long n, const double *__restrict__ A, double *__restrict__ B) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        B[i] = A[i];
}
__global__ void add_kernel(long n, const double *__restrict__ A, double *__restrict__ B) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        B[i] += A[i];
}

// -------------------------------------------------------------- norm --
// Deterministic two-level sum of squares: block b reduces the fixed slice
// [b*chunk, (b+1)*chunk) in a fixed order; one block then sums the partials.
// The slices are kNormChunk values long whatever n is, so a vector's norm
// does not depend on where its partials are kept.
constexpr long kNormChunk = 8192;
__global__ __launch_bounds__(256) void sumsq_partial(long n, long chunk, const double *__restrict__ v,
                                                     double *__restrict__ part) {
    __shared__ double s[256];
    const long b0 = (long)blockIdx.x * chunk;
    const long b1 = (b0 + chunk < n) ? b0 + chunk : n;
    double acc = 0.0;
    for (long i = b0 + threadIdx.x; i < b1; i += 256) acc += v[i] * v[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ __launch_bounds__(256) void sum_final(int nb, const double *__restrict__ part,
                                                 double *__restrict__ out) {
    __shared__ double s[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sqrt(s[0]);
}

// ------------------------------------------------------ no-alloc norm --
// Partials and result of tbs_norm2_noalloc in a file-scope device array: one
// copy per GPU, held until the process ends (kNoAllocMaxPartials + 1 doubles,
// 512 KiB + 8 B).  Host code never passes this name to a kernel or a copy (in
// host code it is only a stand-in for the symbol): the device address comes
// from hipGetSymbolAddress, once per device, with that device current.
constexpr long kNoAllocMaxPartials = 65536;   // vectors of up to 2^29 values
__device__ __align__(256) static double scratch[kNoAllocMaxPartials + 1];
constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
double *g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

unsigned grid_for(long n) {
    long g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

// Block sizes built into the drop-in kernels.
#define TBS_REF_NVARS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
    X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)

extern "C" {

int tbs_supported_nvar(int *out, int cap) {
    int n = 0;
#define X(v) if (out && n < cap) out[n] = v; ++n;
    TBS_REF_NVARS(X)
#undef X
    return n;
}

int tbs_copyAtoBjac(tbs_ctx *ctx, int nelem, const double *A, double *B) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(nelem >= 0, TBS_ERR_ARG, "nelem < 0");
    const long n = (long)nelem * ctx->nvar * ctx->nvar;
    if (n == 0) return TBS_OK;
    TBS_REQUIRE(A && B, TBS_ERR_ARG, "null array");
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, n, A, B);
    TBS_LAUNCHED("copyAtoBjac");
    return TBS_OK;
}

int tbs_copyAtoB(tbs_ctx *ctx, int nelem, const double *A, double *B) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(nelem >= 0, TBS_ERR_ARG, "nelem < 0");
    const long n = (long)nelem * ctx->nvar;
    if (n == 0) return TBS_OK;
    TBS_REQUIRE(A && B, TBS_ERR_ARG, "null array");
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, n, A, B);
    TBS_LAUNCHED("copyAtoB");
    return TBS_OK;
}

int tbs_addAtoB(tbs_ctx *ctx, int nelem, const double *A, double *B) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(nelem >= 0, TBS_ERR_ARG, "nelem < 0");
    const long n = (long)nelem * ctx->nvar;
    if (n == 0) return TBS_OK;
    TBS_REQUIRE(A && B, TBS_ERR_ARG, "null array");
    hipLaunchKernelGGL(add_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, n, A, B);
    TBS_LAUNCHED("addAtoB");
    return TBS_OK;
}

int tbs_assembleTriBlocks(tbs_ctx *ctx, int nline, const int *linesize, const int *linepoint,
                          const int *lines, const int *lineface, const int *fc,
                          const double *jacO1, const double *jacO2, double *A) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(nline >= 0, TBS_ERR_ARG, "nline %d < 0", nline);
    if (nline == 0) return TBS_OK;
    TBS_REQUIRE(linesize && linepoint && lines && lineface && fc && jacO1 && jacO2 && A,
                TBS_ERR_ARG, "null array");
    switch (ctx->nvar) {
#define X(v) case v: hipLaunchKernelGGL(assemble_kernel<v>, dim3(16, std::min(nline, 65535)), dim3(256), 0, ctx->stream, \
                   nline, linesize, linepoint, lines, lineface, fc, jacO1, jacO2, A); break;
        TBS_REF_NVARS(X)
#undef X
        default: TBS_REQUIRE(false, TBS_ERR_UNSUPPORTED, "nvar %d", ctx->nvar);
    }
    TBS_LAUNCHED("assembleTriBlocks");
    return TBS_OK;
}

int tbs_lineLU(tbs_ctx *ctx, int nelem, int nintfaces, int nlines, const int *fc,
               const int *linesize, const int *linepoint, const int *lines,
               const int *lineface, double *Dia, const double *Of1, const double *Of2,
               double *DinvC) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(nelem >= 0 && nintfaces >= 0 && nlines >= 0, TBS_ERR_ARG, "negative size");
    if (nlines == 0) return TBS_OK;
    TBS_REQUIRE(linesize && linepoint && lines && lineface && Dia && DinvC, TBS_ERR_ARG,
                "null array");
    TBS_REQUIRE(nintfaces == 0 || (fc && Of1 && Of2), TBS_ERR_ARG, "null face array");
    switch (ctx->nvar) {
#define X(v) case v: hipLaunchKernelGGL(lineLU_kernel<v>, dim3(nlines), dim3(64), 0, ctx->stream, \
                   nlines, fc, linesize, linepoint, lines, lineface, Dia, Of1, Of2, DinvC); break;
        TBS_REF_NVARS(X)
#undef X
        default: TBS_REQUIRE(false, TBS_ERR_UNSUPPORTED, "nvar %d", ctx->nvar);
    }
    TBS_LAUNCHED("lineLU");
    return TBS_OK;
}

int tbs_triblock_solveDU(tbs_ctx *ctx, int nelem, int nintfaces, int eftot, int nlines,
                         int linelemtot, const int *linesize, const int *linepoint,
                         const int *lines, const double *Dia, const double *DinvC,
                         const double *A, double *dU, const double *R) {
    (void)nintfaces; (void)eftot; (void)linelemtot;
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(nelem >= 0 && nlines >= 0, TBS_ERR_ARG, "negative size");
    if (nlines == 0) return TBS_OK;
    TBS_REQUIRE(linesize && linepoint && lines && Dia && DinvC && A && dU && R, TBS_ERR_ARG,
                "null array");
    switch (ctx->nvar) {
#define X(v) case v: hipLaunchKernelGGL(solveDU_kernel<v>, dim3(nlines), dim3(192), 0, ctx->stream, \
                   nlines, linesize, linepoint, lines, Dia, DinvC, A, dU, R); break;
        TBS_REF_NVARS(X)
#undef X
        default: TBS_REQUIRE(false, TBS_ERR_UNSUPPORTED, "nvar %d", ctx->nvar);
    }
    TBS_LAUNCHED("triblock_solveDU");
    return TBS_OK;
}

int tbs_triblock_lineRes(tbs_ctx *ctx, int nelem, int nintfaces, int eftot, int nlines,
                         int linelemtot, const int *epoint, const int *ef, const int *fc,
                         const int *linesize, const int *linepoint, const int *lines,
                         const double *Dia, const double *Of1, const double *Of2,
                         const double *U, double *R) {
    (void)nelem; (void)eftot; (void)linelemtot;
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(nlines >= 0, TBS_ERR_ARG, "nlines %d < 0", nlines);
    if (nlines == 0) return TBS_OK;
    TBS_REQUIRE(epoint && ef && linesize && linepoint && lines && Dia && U && R, TBS_ERR_ARG,
                "null array");
    TBS_REQUIRE(nintfaces == 0 || (fc && Of1 && Of2), TBS_ERR_ARG, "null face array");
    const int nv = ctx->nvar;
    switch (nv) {
#define X(v) case v: hipLaunchKernelGGL(lineRes_kernel<v>, dim3(8, std::min(nlines, 65535)), dim3(256), 0, ctx->stream, \
                   nlines, epoint, ef, fc, linesize, linepoint, lines, Dia, Of1, Of2, U, R); break;
        TBS_REF_NVARS(X)
#undef X
        default: TBS_REQUIRE(false, TBS_ERR_UNSUPPORTED, "nvar %d", nv);
    }
    TBS_LAUNCHED("triblock_lineRes");
    return TBS_OK;
}

int tbs_dgemvLU(tbs_ctx *ctx, int nelem, const double *LU, const double *x, double *b) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(nelem >= 0, TBS_ERR_ARG, "nelem < 0");
    if (nelem == 0) return TBS_OK;
    TBS_REQUIRE(LU && x && b, TBS_ERR_ARG, "null array");
    const int nv = ctx->nvar;
    const long epb = 256 / nv;
    const unsigned grid = (unsigned)std::min<long>((nelem + epb - 1) / epb, 65536L);
    switch (nv) {
#define X(v) case v: hipLaunchKernelGGL(dgemvLU_kernel<v>, dim3(grid), dim3(256), 0, ctx->stream, \
                   (long)nelem, LU, x, b); break;
        TBS_REF_NVARS(X)
#undef X
        default: TBS_REQUIRE(false, TBS_ERR_UNSUPPORTED, "nvar %d", nv);
    }
    TBS_LAUNCHED("dgemvLU");
    return TBS_OK;
}

}  // extern "C"

namespace {

// The norm's kernels on ctx's stream with the partials in part[0, nb) and the
// result in part[nb], then the result to the host (synchronous).
int norm2_with(tbs_ctx *ctx, long n, const double *v, double *part, long nb, double *out_host) {
    hipLaunchKernelGGL(sumsq_partial, dim3((unsigned)nb), dim3(256), 0, ctx->stream, n, kNormChunk, v, part);
    TBS_LAUNCHED("sumsq_partial");
    hipLaunchKernelGGL(sum_final, dim3(1), dim3(256), 0, ctx->stream, (int)nb, part, part + nb);
    TBS_LAUNCHED("sum_final");
    TBS_HIP(hipMemcpyAsync(ctx->red_host, part + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TBS_HIP(hipStreamSynchronize(ctx->stream));
    *out_host = *ctx->red_host;
    return TBS_OK;
}

}  // namespace

extern "C" {

int tbs_norm2(tbs_ctx *ctx, long n, const double *v, double *out_host) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(n >= 0 && out_host, TBS_ERR_ARG, "bad norm arguments");
    if (n == 0) { *out_host = 0.0; return TBS_OK; }
    TBS_REQUIRE(v, TBS_ERR_ARG, "null vector");
    const long nb = (n + kNormChunk - 1) / kNormChunk;
    TBS_REQUIRE(nb < INT_MAX, TBS_ERR_ARG, "vector of %ld values too long", n);
    // the context's partial buffer, grown (after its last use) to nb + 1
    if ((size_t)nb + 1 > ctx->red_cap) {
        if (ctx->red_buf) {
            TBS_HIP(hipStreamSynchronize(ctx->stream));
            TBS_HIP(hipFree(ctx->red_buf));
        }
        ctx->red_buf = nullptr;
        ctx->red_cap = 0;
        const size_t cap = std::max<size_t>((size_t)nb + 1, 4096);
        TBS_HIP(hipMalloc((void **)&ctx->red_buf, cap * sizeof(double)));
        ctx->red_cap = cap;
    }
    return norm2_with(ctx, n, v, ctx->red_buf, nb, out_host);
}

int tbs_norm2_noalloc(tbs_ctx *ctx, long n, const double *v, double *out_host) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(n >= 0 && out_host, TBS_ERR_ARG, "bad norm arguments");
    if (n == 0) { *out_host = 0.0; return TBS_OK; }
    TBS_REQUIRE(v, TBS_ERR_ARG, "null vector");
    const long nb = (n + kNormChunk - 1) / kNormChunk;
    TBS_REQUIRE(nb <= kNoAllocMaxPartials, TBS_ERR_ARG,
                "tbs_norm2_noalloc: %ld values exceed its maximum of %ld (its static scratch holds %ld "
                "partials); use tbs_norm2", n, kNoAllocMaxPartials * kNormChunk, kNoAllocMaxPartials);
    const int dev = ctx->device;
    TBS_REQUIRE(dev >= 0 && dev < kMaxDevices, TBS_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
    // the device's copy of the scratch array: its address, taken once per
    // device with that device current (use_ctx above)
    std::call_once(g_scratch_once[dev], [dev]() {
        g_scratch_err[dev] = hipGetSymbolAddress((void **)&g_scratch_ptr[dev], HIP_SYMBOL(scratch));
    });
    if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
        return tbs::hip_fail(g_scratch_err[dev] != hipSuccess ? g_scratch_err[dev] : hipErrorInvalidSymbol,
                             "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    return norm2_with(ctx, n, v, g_scratch_ptr[dev], nb, out_host);
}

}  // extern "C"
