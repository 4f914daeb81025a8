// gmres.hip -- line-preconditioned restarted GMRES on a factored plan.
//
// Solves J U = -rhs (the LIJ convention: the residual R = rhs + J U goes to
// zero) with flexible right-preconditioned GMRES(m) (Saad, "A flexible
// inner-outer preconditioned GMRES algorithm", SIAM J. Sci. Comput. 14, 1993):
// the preconditioner z = P v is `sweeps` LIJ sweeps on J z = v from z = 0,
// the first of which is the block-tridiagonal line solve z = M^-1 v.  This is
// the line-preconditioned GMRES the reference's paper pairs with the LIJ
// smoother (kirby-mavriplis-2025, Table 6).  Operator applications and
// preconditioner sweeps run on the plan's own residual and Thomas kernels; the
// Krylov basis lives in HBM; Gram-Schmidt is classical with one
// re-orthogonalisation (two batched dot/AXPY passes per step); the Hessenberg
// least-squares problem is updated with Givens rotations on the host.
// Reductions use a fixed grid and a fixed summation order, so runs repeat
// bitwise.
#include "tbs_common.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace {

constexpr int kDotBlocks = 1024;
constexpr int kDotThreads = 256;
constexpr int kChunk = 8;           // vectors per dot pass

// partial[b][i] = sum over this block's elements of w * V_i, i < kc
template <int KC>
__global__ __launch_bounds__(kDotThreads) void mdot_kernel(long N, int kc, const double *__restrict__ w,
                                                         const double *__restrict__ V, long ldv,
                                                         double *__restrict__ partial) {
    __shared__ double s[KC][kDotThreads];
    double acc[KC];
#pragma unroll
    for (int i = 0; i < KC; ++i) acc[i] = 0.0;
    for (long t = (long)blockIdx.x * kDotThreads + threadIdx.x; t < N; t += (long)gridDim.x * kDotThreads) {
        const double wt = w[t];
#pragma unroll
        for (int i = 0; i < KC; ++i)
            if (i < kc) acc[i] += wt * V[i * ldv + t];
    }
#pragma unroll
    for (int i = 0; i < KC; ++i) s[i][threadIdx.x] = acc[i];
    __syncthreads();
    for (int h = kDotThreads / 2; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) {
#pragma unroll
            for (int i = 0; i < KC; ++i) s[i][threadIdx.x] += s[i][threadIdx.x + h];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < kc) partial[(long)blockIdx.x * KC + threadIdx.x] = s[threadIdx.x][0];
}

// out[i] = sum over blocks (fixed order) of partial[b][i]; one block per i
template <int KC>
__global__ __launch_bounds__(kDotThreads) void mdot_finish_kernel(int nblk, const double *__restrict__ partial,
                                                                double *__restrict__ out) {
    __shared__ double s[kDotThreads];
    const int i = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kDotThreads) acc += partial[(long)b * KC + i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kDotThreads / 2; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = s[0];
}

// y += sum_i c[i] * V_i   (i < k)
__global__ void maxpy_kernel(long N, int k, const double *__restrict__ c, const double *__restrict__ V, long ldv,
                             double *__restrict__ y) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += (long)gridDim.x * blockDim.x) {
        double acc = y[t];
        for (int i = 0; i < k; ++i) acc += c[i] * V[i * ldv + t];
        y[t] = acc;
    }
}

// dst = alpha * src (src may be null: dst = 0)
__global__ void scale_kernel(long N, double alpha, const double *__restrict__ src, double *__restrict__ dst) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += (long)gridDim.x * blockDim.x)
        dst[t] = src ? alpha * src[t] : 0.0;
}

unsigned grid_for(long n) {
    const long g = (n + 255) / 256;
    return (unsigned)std::max(1L, std::min(g, 16384L));
}

}  // namespace

struct tbs_gmres {
    tbs_plan *plan = nullptr;
    tbs_ctx *ctx = nullptr;
    int m = 0;
    long N = 0;
    double *V = nullptr;      // [m+1][N] orthonormal basis
    double *Z = nullptr;      // [m][N] preconditioned directions
    double *x = nullptr;      // [N] iterate (internal order)
    double *r0 = nullptr;     // [N] rhs (internal order)
    double *partial = nullptr, *dots = nullptr, *coef = nullptr;
    double *h_dots = nullptr; // pinned
    double stats[4] = {};     // operator applications, preconditioner sweeps, restarts, final true residual
};

namespace {

int run_dots(tbs_gmres *g, hipStream_t st, const double *w, const double *V, int k, double *out_host) {
    const int nblk = (int)std::min<long>(kDotBlocks, std::max<long>(1, (g->N + kDotThreads - 1) / kDotThreads));
    for (int c0 = 0; c0 < k; c0 += kChunk) {
        const int kc = std::min(kChunk, k - c0);
        hipLaunchKernelGGL((mdot_kernel<kChunk>), dim3(nblk), dim3(kDotThreads), 0, st, g->N, kc, w,
                           V + (long)c0 * g->N, g->N, g->partial);
        hipLaunchKernelGGL((mdot_finish_kernel<kChunk>), dim3(kc), dim3(kDotThreads), 0, st, nblk, g->partial,
                           g->dots + c0);
        TBS_LAUNCHED("gmres dots");
    }
    TBS_HIP(hipMemcpyAsync(g->h_dots, g->dots, sizeof(double) * k, hipMemcpyDeviceToHost, st));
    TBS_HIP(hipStreamSynchronize(st));
    std::memcpy(out_host, g->h_dots, sizeof(double) * k);
    return TBS_OK;
}

int run_axpy(tbs_gmres *g, hipStream_t st, const double *c_host, int k, const double *V, double *y) {
    TBS_HIP(hipMemcpyAsync(g->coef, c_host, sizeof(double) * k, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(maxpy_kernel, dim3(grid_for(g->N)), dim3(256), 0, st, g->N, k, g->coef, V, g->N, y);
    TBS_LAUNCHED("gmres axpy");
    return TBS_OK;
}

int run_scale(tbs_gmres *g, hipStream_t st, double alpha, const double *src, double *dst) {
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(g->N)), dim3(256), 0, st, g->N, alpha, src, dst);
    TBS_LAUNCHED("gmres scale");
    return TBS_OK;
}

}  // namespace

extern "C" {

int tbs_gmres_create(tbs_plan *plan, int restart, tbs_gmres **out) {
    TBS_REQUIRE(plan && out && restart >= 1 && restart <= 256, TBS_ERR_ARG, "bad GMRES arguments");
    *out = nullptr;
    tbs::PlanVecs pv{};
    TBS_TRY(tbs::plan_vecs(plan, &pv));
    tbs_ctx *ctx = pv.ctx;
    TBS_TRY(tbs::use_ctx(ctx));
    std::unique_ptr<tbs_gmres, int (*)(tbs_gmres *)> g(new tbs_gmres(), tbs_gmres_destroy);
    g->plan = plan;
    g->ctx = ctx;
    g->m = restart;
    g->N = pv.n * pv.nv;
    auto dalloc = [&](double **p, size_t count) -> int {
        hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(double));
        if (e == hipErrorOutOfMemory) {
            hipGetLastError();
            tbs::set_error("GMRES workspace of %zu bytes refused", count * sizeof(double));
            return TBS_ERR_NOMEM;
        }
        TBS_HIP(e);
        return TBS_OK;
    };
    TBS_TRY(dalloc(&g->V, (size_t)(restart + 1) * g->N));
    TBS_TRY(dalloc(&g->Z, (size_t)restart * g->N));
    TBS_TRY(dalloc(&g->x, g->N));
    TBS_TRY(dalloc(&g->r0, g->N));
    TBS_TRY(dalloc(&g->partial, (size_t)kDotBlocks * kChunk));
    TBS_TRY(dalloc(&g->dots, restart + 1));
    TBS_TRY(dalloc(&g->coef, restart + 1));
    TBS_HIP(hipHostMalloc((void **)&g->h_dots, sizeof(double) * (restart + 1), hipHostMallocDefault));
    *out = g.release();
    return TBS_OK;
}

int tbs_gmres_destroy(tbs_gmres *g) {
    if (!g) return TBS_OK;
    if (g->ctx) {
        hipSetDevice(g->ctx->device);
        hipStreamSynchronize(g->ctx->stream);
    }
    for (double *p : {g->V, g->Z, g->x, g->r0, g->partial, g->dots, g->coef})
        if (p) hipFree(p);
    if (g->h_dots) hipHostFree(g->h_dots);
    delete g;
    return TBS_OK;
}

int tbs_gmres_solve(tbs_gmres *g, const double *rhs, double *U, int max_iters, double rtol, int sweeps,
                    double *hist, int hist_cap, int *iters) {
    TBS_REQUIRE(g && rhs && U && max_iters >= 0 && rtol >= 0 && sweeps >= 1, TBS_ERR_ARG, "bad GMRES arguments");
    TBS_TRY(tbs::use_ctx(g->ctx));
    tbs::PlanVecs pv{};
    TBS_TRY(tbs::plan_vecs(g->plan, &pv));
    hipStream_t st = g->ctx->stream;
    tbs_plan *p = g->plan;
    const long N = g->N;
    const int m = g->m;
    TBS_TRY(tbs::plan_gather(p, st, rhs, g->r0));
    TBS_TRY(tbs::plan_gather(p, st, U, g->x));
    TBS_TRY(tbs::plan_activate(p, st));
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m), h(m + 1), h2(m + 1);
    auto Hij = [&](int i, int j) -> double & { return H[(size_t)j * (m + 1) + i]; };
    int it = 0, nh = 0;
    double beta0 = 0.0;
    double ops = 0, psweeps = 0, restarts = 0, true_res = 0;
    auto record = [&](double v) {
        if (hist && nh < hist_cap) hist[nh] = v;
        ++nh;
    };
    for (;;) {
        // R = rhs + J x; the GMRES residual of J x = -rhs is -R
        TBS_TRY(run_scale(g, st, 1.0, g->r0, pv.rhs));
        TBS_TRY(run_scale(g, st, 1.0, g->x, pv.U));
        TBS_TRY(tbs::plan_residual(p, st, 1));
        ops += 1;
        double beta = 0.0;
        TBS_TRY(run_dots(g, st, pv.R, pv.R, 1, &beta));
        beta = std::sqrt(beta);
        true_res = beta;
        if (it == 0) {
            beta0 = beta;
            record(beta);
        }
        if (!(beta == beta) || beta > 1e300) {
            tbs::set_error("non-finite GMRES residual");
            return TBS_ERR_NUMERIC;
        }
        if (beta <= rtol * beta0 || beta == 0.0 || it >= max_iters) break;
        restarts += 1;
        TBS_TRY(run_scale(g, st, -1.0 / beta, pv.R, g->V));
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        int k = 0;
        bool done = false;
        for (int j = 0; j < m && it < max_iters; ++j) {
            double *Vj = g->V + (long)j * N, *Zj = g->Z + (long)j * N, *w = g->V + (long)(j + 1) * N;
            // z = P v_j: LIJ sweeps on J z = v_j from z = 0 (residual convention rhs = -v_j)
            TBS_TRY(run_scale(g, st, -1.0, Vj, pv.rhs));
            TBS_TRY(run_scale(g, st, 0.0, nullptr, pv.U));
            TBS_TRY(tbs::plan_activate(p, st));
            for (int s = 0; s < sweeps; ++s) {
                TBS_TRY(tbs::plan_residual(p, st, 0, s == 0));   // first: U = 0, g = D'^-1 v_j
                TBS_TRY(tbs::plan_thomas(p, st));
            }
            psweeps += sweeps;
            TBS_TRY(run_scale(g, st, 1.0, pv.U, Zj));
            // w = J z (rhs = 0, U = z)
            TBS_TRY(run_scale(g, st, 0.0, nullptr, pv.rhs));
            TBS_TRY(tbs::plan_residual(p, st, 1));
            ops += 1;
            TBS_TRY(run_scale(g, st, 1.0, pv.R, w));
            // classical Gram-Schmidt, twice
            TBS_TRY(run_dots(g, st, w, g->V, j + 1, h.data()));
            for (int i = 0; i <= j; ++i) h2[i] = -h[i];
            TBS_TRY(run_axpy(g, st, h2.data(), j + 1, g->V, w));
            TBS_TRY(run_dots(g, st, w, g->V, j + 1, h2.data()));
            for (int i = 0; i <= j; ++i) {
                h[i] += h2[i];
                h2[i] = -h2[i];
            }
            TBS_TRY(run_axpy(g, st, h2.data(), j + 1, g->V, w));
            double hn = 0.0;
            TBS_TRY(run_dots(g, st, w, w, 1, &hn));
            hn = std::sqrt(hn);
            if (hn > 0.0) TBS_TRY(run_scale(g, st, 1.0 / hn, w, w));
            // Hessenberg column, previous rotations, new rotation
            for (int i = 0; i <= j; ++i) Hij(i, j) = h[i];
            Hij(j + 1, j) = hn;
            for (int i = 0; i < j; ++i) {
                const double a = Hij(i, j), b = Hij(i + 1, j);
                Hij(i, j) = cs[i] * a + sn[i] * b;
                Hij(i + 1, j) = -sn[i] * a + cs[i] * b;
            }
            const double a = Hij(j, j), b = Hij(j + 1, j);
            const double r = std::hypot(a, b);
            cs[j] = r > 0 ? a / r : 1.0;
            sn[j] = r > 0 ? b / r : 0.0;
            Hij(j, j) = r;
            Hij(j + 1, j) = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            ++it;
            k = j + 1;
            const double est = std::fabs(gv[j + 1]);
            record(est);
            if (est <= rtol * beta0 || hn == 0.0) {
                done = true;
                break;
            }
        }
        // y = H^-1 g (upper triangular), x += Z y
        for (int i = k - 1; i >= 0; --i) {
            double t = gv[i];
            for (int l = i + 1; l < k; ++l) t -= Hij(i, l) * y[l];
            y[i] = t / Hij(i, i);
        }
        if (k > 0) TBS_TRY(run_axpy(g, st, y.data(), k, g->Z, g->x));
        (void)done;
    }
    TBS_TRY(tbs::plan_scatter(p, st, g->x, U));
    TBS_HIP(hipStreamSynchronize(st));
    g->stats[0] = ops;
    g->stats[1] = psweeps;
    g->stats[2] = restarts;
    g->stats[3] = true_res;
    if (iters) *iters = it;
    return TBS_OK;
}

int tbs_gmres_stats(tbs_gmres *g, double *stats, int n) {
    TBS_REQUIRE(g && stats, TBS_ERR_ARG, "null argument");
    for (int i = 0; i < n && i < 4; ++i) stats[i] = g->stats[i];
    return TBS_OK;
}

}  // extern "C"
