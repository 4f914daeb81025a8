/*
 * oracle.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain C++ restatement, with OCCA "Serial"/"OpenMP" semantics, of the
 * reference's active line-implicit Jacobi (LIJ) path:
 *   - LU / solveLU / lineLU          src/okl/lineLU_v2.okl:17-27, :30-52, :58-182
 *   - triblock_solveDU               src/okl/linesmoothLU_TriBlock_v3.okl:92-226
 *   - triblock_lineRes               src/okl/linesmoothLU_TriBlock_v3.okl:228-336
 *   - copyAtoB / addAtoB / copyAtoBjac
 *                                    src/okl/linesmoothLU_v5.okl:75-103
 *   - Jacobian::assembleTriBlocks    src/Jacobian.cxx:76-117
 *   - the driver loop                src/main.cxx:180-229
 * OCCA semantics used here: every @outer/@inner loop becomes an ordinary loop,
 * @shared arrays become locals and @barrier is a no-op; the "all lanes run
 * LU()" sections of lineLU run once.  OpenMP parallelises the outermost
 * @outer loop (lines), which is what OCCA's OpenMP mode does.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline.
 *
 * Parity pins (tests/test_oracle_*.py): LU and solveLU against numpy dense
 * factorisation/solve; lineLU+solveDU against a scipy sparse direct solve of
 * the assembled block-tridiagonal line system; lineRes against a scipy BSR
 * assembly of the full Jacobian; the loop against an independent numpy LIJ
 * iteration, the scalar (NVAR=1) textbook Thomas algorithm, point-Jacobi for
 * length-1 lines and the one-sweep exact solve when off-line blocks vanish.
 * Planted-mistake builds (-DORACLE_MUTATION=n) must fail those checks.
 *
 * Block layout (as in the reference, OKL @dim(NVAR,NVAR,n), first index
 * fastest): entry (i,j) of block b sits at i + nv*j + nv*nv*b.
 * Vectors: entry i of element e sits at i + nv*e.
 */
#include <cmath>
#include <cstring>
#include <cstdint>
#include <vector>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORACLE_MUTATION
#define ORACLE_MUTATION 0
#endif

#define MX(A, i, j) (A)[(i) + nv * (j)]

static int set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

extern "C" {

/* Doolittle LU, no pivoting, unit-lower L below the diagonal and U on/above;
 * multiplies by the reciprocal pivot.  src/okl/lineLU_v2.okl:17-27 */
void orc_LU(int nv, double *A) {
    for (int j = 0; j < nv; ++j) {
        double pivot = 1.0 / MX(A, j, j);
        for (int i = j + 1; i < nv; ++i) {
            MX(A, i, j) = MX(A, i, j) * pivot;
            for (int k = j + 1; k < nv; ++k) {
#if ORACLE_MUTATION == 1
                MX(A, i, k) -= MX(A, i, j) * MX(A, k, j); /* planted: transposed U operand */
#else
                MX(A, i, k) -= MX(A, i, j) * MX(A, j, k);
#endif
            }
        }
    }
}

/* Solve (LU) x = b with the factors of orc_LU: forward then back
 * substitution, dividing by the U pivot.  src/okl/lineLU_v2.okl:30-52 and
 * src/okl/linesmoothLU_TriBlock_v3.okl:23-45 (identical bodies). */
void orc_solveLU(int nv, const double *A, const double *b, double *x) {
    double y[64];
    for (int i = 0; i < nv; ++i) {
        double tot = 0.0;
        for (int j = 0; j < i; ++j) tot += MX(A, i, j) * y[j];
        y[i] = b[i] - tot;
    }
    for (int i = nv - 1; i >= 0; --i) {
        double tot = 0.0;
        for (int j = i + 1; j < nv; ++j) tot += MX(A, i, j) * x[j];
        x[i] = (y[i] - tot) / MX(A, i, i);
    }
}

/* b = L U x with the packed factors of orc_LU (unit-lower L below the
 * diagonal, U on and above it), one block per element: U first (row sums
 * from the diagonal, ascending j), then L starting from b_i = y_i and adding
 * L(i,j) y_j for ascending j < i.  src/okl/linesmoothLU_TriBlock_v3.okl:48-69 */
void orc_dgemvLU(int nv, int nelem, const double *LU, const double *x, double *b, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nelem; ++e) {
        const double *A = LU + (size_t)nv * nv * e;
        const double *xe = x + (size_t)nv * e;
        double *be = b + (size_t)nv * e;
        double y[64];
        for (int i = 0; i < nv; ++i) {
            double tot = 0.0;
#if ORACLE_MUTATION == 5
            for (int j = i + 1; j < nv; ++j) tot += MX(A, i, j) * xe[j]; /* planted: U diagonal dropped */
#else
            for (int j = i; j < nv; ++j) tot += MX(A, i, j) * xe[j];
#endif
            y[i] = tot;
        }
        for (int i = 0; i < nv; ++i) {
            be[i] = y[i];
            for (int j = 0; j < i; ++j) be[i] += MX(A, i, j) * y[j];
        }
    }
}

/* B[n] = A[n] over n < nelem*nv.  src/okl/linesmoothLU_v5.okl:75-83 */
void orc_copyAtoB(int nv, int nelem, const double *A, double *B, int nthreads) {
    set_threads(nthreads);
    const long n = (long)nelem * nv;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) B[i] = A[i];
}

/* B[n] = A[n] over n < nelem*nv*nv.  src/okl/linesmoothLU_v5.okl:85-93 */
void orc_copyAtoBjac(int nv, int nelem, const double *A, double *B, int nthreads) {
    set_threads(nthreads);
    const long n = (long)nelem * nv * nv;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) B[i] = A[i];
}

/* B[n] += A[n] over n < nelem*nv.  src/okl/linesmoothLU_v5.okl:95-103 */
void orc_addAtoB(int nv, int nelem, const double *A, double *B, int nthreads) {
    set_threads(nthreads);
    const long n = (long)nelem * nv;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) B[i] += A[i];
}

/* Pack the in-line lower block of every non-first line element, by element
 * id: A[e] = (e == fc[2f]) ? jacO2[f] : jacO1[f], f = lineface[m].
 * src/Jacobian.cxx:76-117 */
void orc_assembleTriBlocks(int nv, int nline, const int *linesize,
                           const int *linepoint, const int *lines,
                           const int *lineface, const int *fc,
                           const double *jacO1, const double *jacO2,
                           double *A) {
    const int bs = nv * nv;
    for (int l = 0; l < nline; ++l) {
        const int n = linesize[l];
        for (int k = 1; k < n; ++k) {
            const int m = linepoint[l] + k;
            const int f = lineface[m];
            const int e = lines[m];
            const int e1 = fc[2 * f];
#if ORACLE_MUTATION == 2
            const double *src = (e == e1) ? &jacO1[(size_t)bs * f] : &jacO2[(size_t)bs * f]; /* planted: swapped */
#else
            const double *src = (e == e1) ? &jacO2[(size_t)bs * f] : &jacO1[(size_t)bs * f];
#endif
            std::memcpy(&A[(size_t)bs * e], src, bs * sizeof(double));
        }
    }
}

/* Block-Thomas factorisation of every line, in place on Dia.
 * src/okl/lineLU_v2.okl:58-182 */
void orc_lineLU(int nv, int nelem, int nintfaces, int nlines, const int *fc,
                const int *linesize, const int *linepoint, const int *lines,
                const int *lineface, double *Dia, const double *Of1,
                const double *Of2, double *DinvC, int nthreads) {
    (void)nelem; (void)nintfaces;
    set_threads(nthreads);
    const int bs = nv * nv;
#pragma omp parallel for schedule(dynamic, 1)
    for (int l = 0; l < nlines; ++l) {
        const int nelem_line = linesize[l];
        double Gamma[64 * 64], Alpha[64 * 64], s_C[64 * 64], s_AT[64 * 64];
        double s_Dia_e[64 * 64], s_Dia_elast[64 * 64];
        const int m0 = linepoint[l];
        orc_LU(nv, &Dia[(size_t)bs * lines[m0]]);                       /* :91 */
        for (int k = 1; k < nelem_line; ++k) {
            const int e = lines[m0 + k];
            const int f = lineface[m0 + k];
            const int e1 = fc[2 * f];
            const int elast = lines[m0 + k - 1];
            const double *A = (e == e1) ? Of2 : Of1;                   /* :116 */
            const double *C = (e == e1) ? Of1 : Of2;                   /* :117 */
            for (int i = 0; i < nv; ++i) {                              /* :119-139 */
                for (int j = 0; j < nv; ++j) MX(s_C, i, j) = C[(size_t)bs * f + i + nv * j];
                for (int j = 0; j < nv; ++j) MX(s_AT, j, i) = A[(size_t)bs * f + i + nv * j];
                for (int j = 0; j < nv; ++j) MX(s_Dia_elast, i, j) = Dia[(size_t)bs * elast + i + nv * j];
                for (int j = 0; j < nv; ++j) MX(s_Dia_e, i, j) = Dia[(size_t)bs * e + i + nv * j];
            }
            for (int i = 0; i < nv; ++i)                                /* :142-144 */
                orc_solveLU(nv, s_Dia_elast, &MX(s_C, 0, i), &MX(Gamma, 0, i));
            for (int i = 0; i < nv; ++i) {                              /* :147-155 */
                for (int kk = 0; kk < nv; ++kk) {
                    double tot = 0.0;
                    for (int j = 0; j < nv; ++j) tot += MX(s_AT, j, i) * MX(Gamma, j, kk);
                    MX(Alpha, i, kk) = tot;
                }
            }
            for (int i = 0; i < nv; ++i)                                /* :158-162 */
                for (int j = 0; j < nv; ++j) MX(s_Dia_e, i, j) -= MX(Alpha, i, j);
            orc_LU(nv, s_Dia_e);                                        /* :165 */
            for (int i = 0; i < nv; ++i) {                              /* :169-178 */
                for (int j = 0; j < nv; ++j) Dia[(size_t)bs * e + i + nv * j] = MX(s_Dia_e, i, j);
                for (int j = 0; j < nv; ++j) DinvC[(size_t)bs * elast + i + nv * j] = MX(Gamma, i, j);
            }
        }
    }
}

/* Block-Thomas forward elimination and back substitution, solving
 * M dU = -R per line.  src/okl/linesmoothLU_TriBlock_v3.okl:92-226 */
void orc_solveDU(int nv, int nelem, int nintfaces, int eftot, int nlines,
                 int linelemtot, const int *linesize, const int *linepoint,
                 const int *lines, const double *Dia, const double *DinvC,
                 const double *A, double *dU, const double *R, int nthreads) {
    (void)nelem; (void)nintfaces; (void)eftot; (void)linelemtot;
    set_threads(nthreads);
    const int bs = nv * nv;
#pragma omp parallel for schedule(dynamic, 1)
    for (int l = 0; l < nlines; ++l) {
        const int n = linesize[l];
        const int m0 = linepoint[l];
        double x[64], S[64], s_dU_e[64], s_AT[64 * 64], s_DinvCT[64 * 64], s_Dia[64 * 64];
        const int e0 = lines[m0];
        for (int i = 0; i < nv; ++i) {                                  /* :142-150 */
            S[i] = -R[(size_t)nv * e0 + i];
            for (int j = 0; j < nv; ++j) MX(s_Dia, i, j) = Dia[(size_t)bs * e0 + i + nv * j];
        }
        orc_solveLU(nv, s_Dia, S, s_dU_e);                              /* :153-155 */
        for (int i = 0; i < nv; ++i) dU[(size_t)nv * e0 + i] = s_dU_e[i];
        for (int k = 1; k < n; ++k) {                                   /* :163-196 */
            const int e = lines[m0 + k];
            for (int i = 0; i < nv; ++i) {
                for (int j = 0; j < nv; ++j) MX(s_AT, j, i) = A[(size_t)bs * e + i + nv * j];
                for (int j = 0; j < nv; ++j) MX(s_Dia, i, j) = Dia[(size_t)bs * e + i + nv * j];
            }
            for (int i = 0; i < nv; ++i) {
                double tot = 0.0;
                for (int j = 0; j < nv; ++j) tot += MX(s_AT, j, i) * s_dU_e[j];
                x[i] = tot;
            }
            for (int i = 0; i < nv; ++i) S[i] = -R[(size_t)nv * e + i] - x[i];
            orc_solveLU(nv, s_Dia, S, s_dU_e);
            for (int i = 0; i < nv; ++i) dU[(size_t)nv * e + i] = s_dU_e[i];
        }
        for (int k = n - 2; k >= 0; --k) {                              /* :199-223 */
            const int e = lines[m0 + k];
            const int elast = lines[m0 + k + 1];
            for (int i = 0; i < nv; ++i) {
                for (int j = 0; j < nv; ++j) MX(s_DinvCT, j, i) = DinvC[(size_t)bs * e + i + nv * j];
                s_dU_e[i] = dU[(size_t)nv * elast + i];
            }
            for (int i = 0; i < nv; ++i) {
                double tot = 0.0;
                for (int j = 0; j < nv; ++j) tot += MX(s_DinvCT, j, i) * s_dU_e[j];
#if ORACLE_MUTATION == 3
                dU[(size_t)nv * e + i] += tot;                          /* planted: wrong sign */
#else
                dU[(size_t)nv * e + i] -= tot;
#endif
            }
        }
    }
}

/* Full-row block residual: R(e) += D(e) U(e) + sum over interior faces of e,
 * in ef order, of Off(f) U(nb), Off = (e == fc[2f]) ? Of2 : Of1.  The caller
 * presets R = rhs.  src/okl/linesmoothLU_TriBlock_v3.okl:228-336 */
void orc_lineRes(int nv, int nelem, int nintfaces, int eftot, int nlines,
                 int linelemtot, const int *epoint, const int *ef,
                 const int *fc, const int *linesize, const int *linepoint,
                 const int *lines, const double *Dia, const double *Of1,
                 const double *Of2, const double *U, double *R, int nthreads) {
    (void)nelem; (void)nintfaces; (void)eftot; (void)linelemtot;
    set_threads(nthreads);
    const int bs = nv * nv;
#pragma omp parallel for schedule(dynamic, 1)
    for (int l = 0; l < nlines; ++l) {
        const int n = linesize[l];
        const int m0 = linepoint[l];
        for (int k = 0; k < n; ++k) {
            const int e = lines[m0 + k];
            double s_R[64];
            for (int i = 0; i < nv; ++i) {                              /* :276-293 */
                double tot = 0.0;
                for (int j = 0; j < nv; ++j) tot += Dia[(size_t)bs * e + i + nv * j] * U[(size_t)nv * e + j];
                s_R[i] = tot;
            }
            for (int k2 = epoint[e]; k2 < epoint[e + 1]; ++k2) {        /* :299-328 */
                const int f = ef[k2];
                if (f >= 0) {
                    const int e1 = fc[2 * f + 0];
                    const int e2 = fc[2 * f + 1];
#if ORACLE_MUTATION == 4
                    const double *off = (e == e1) ? Of1 : Of2;          /* planted: swapped */
#else
                    const double *off = (e == e1) ? Of2 : Of1;
#endif
                    const int nb = (e == e1) ? e2 : e1;
                    for (int i = 0; i < nv; ++i) {
                        double tot = 0.0;
                        for (int j = 0; j < nv; ++j) tot += off[(size_t)bs * f + i + nv * j] * U[(size_t)nv * nb + j];
                        s_R[i] += tot;
                    }
                }
            }
            for (int i = 0; i < nv; ++i) R[(size_t)nv * e + i] += s_R[i]; /* :331 */
        }
    }
}

/* Euclidean norm, sequential summation (the reference computes none; this is
 * the build-defined convergence measure). */
double orc_norm2(long n, const double *v) {
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += v[i] * v[i];
    return std::sqrt(s);
}

/* The reference driver: copyAtoBjac + lineLU once, then R = rhs and `iters`
 * sweeps of solveDU, U += dU, R = rhs, lineRes (src/main.cxx:180-229).  U and
 * dU must be zero on entry, as in the reference.  hist[0..iters] receives
 * ||R||_2 before the first and after every sweep.  jacDLU, DinvC, A, dU and R
 * are caller-provided work/output arrays in the reference layout.
 * Returns wall seconds of the factorisation in t_factor and of the sweep
 * loop in t_sweeps (steady_clock); t_each (may be NULL) gets each sweep's own
 * seconds, without the norm (hist may be NULL: then no norm is taken). */
void orc_lij_run(int nv, int nelem, int nintface, int eftot, int nline,
                 int nlineelem, const int *epoint, const int *ef,
                 const int *fc, const int *linesize, const int *linepoint,
                 const int *lines, const int *lineface, const double *jacD,
                 const double *jacO1, const double *jacO2, const double *rhs,
                 double *jacDLU, double *DinvC, double *A, double *U,
                 double *dU, double *R, int iters, double *hist,
                 int nthreads, double *t_factor, double *t_sweeps, double *t_each) {
    using clk = std::chrono::steady_clock;
    orc_assembleTriBlocks(nv, nline, linesize, linepoint, lines, lineface, fc, jacO1, jacO2, A);
    auto t0 = clk::now();
    orc_copyAtoBjac(nv, nelem, jacD, jacDLU, nthreads);
    orc_lineLU(nv, nelem, nintface, nline, fc, linesize, linepoint, lines, lineface,
               jacDLU, jacO1, jacO2, DinvC, nthreads);
    auto t1 = clk::now();
    orc_copyAtoB(nv, nelem, rhs, R, nthreads);
    if (hist) hist[0] = orc_norm2((long)nv * nelem, R);
    for (int p = 0; p < iters; ++p) {
        auto ts = clk::now();
        orc_solveDU(nv, nelem, nintface, eftot, nline, nlineelem, linesize, linepoint, lines,
                    jacDLU, DinvC, A, dU, R, nthreads);
        orc_addAtoB(nv, nelem, dU, U, nthreads);
        orc_copyAtoB(nv, nelem, rhs, R, nthreads);
        orc_lineRes(nv, nelem, nintface, eftot, nline, nlineelem, epoint, ef, fc, linesize,
                    linepoint, lines, jacD, jacO1, jacO2, U, R, nthreads);
        if (t_each) t_each[p] = std::chrono::duration<double>(clk::now() - ts).count();
        if (hist) hist[p + 1] = orc_norm2((long)nv * nelem, R);
    }
    auto t2 = clk::now();
    if (t_factor) *t_factor = std::chrono::duration<double>(t1 - t0).count();
    if (t_sweeps) *t_sweeps = std::chrono::duration<double>(t2 - t1).count();
}

int orc_max_threads(void) { return set_threads(0); }
int orc_mutation(void) { return ORACLE_MUTATION; }

} /* extern "C" */
