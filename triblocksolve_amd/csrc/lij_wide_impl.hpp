// lij_wide_impl.hpp -- body of lij_wide_lo.hip / lij_wide_hi.hip: the
// dispatch of lij_wide.hpp over the block sizes TBS_WIDE_NVARS, in namespace
// tbs::TBS_WIDE_NS.
#include "lij_kernels.hpp"
#include "lij_wide.hpp"

namespace tbs {
namespace TBS_WIDE_NS {

int residual(int nv, int prec, hipStream_t st, const void *sys, const int2 *gtab, int ng, int want_R, int force,
             hipEvent_t e0, hipEvent_t e1) {
    switch (nv) {
#define X(v) case v: return prec == TBS_FP32 ? launch_residual<v, float>(st, sys, gtab, ng, want_R, force, e0, e1) \
                                         : launch_residual<v, double>(st, sys, gtab, ng, want_R, force, e0, e1);
        TBS_WIDE_NVARS(X)
#undef X
    }
    set_error("nvar %d not built for the fused path", nv);
    return TBS_ERR_UNSUPPORTED;
}

int thomas(int nv, int prec, hipStream_t st, const void *sys, const int4 *ltab, int nl, int variant, hipEvent_t e0,
           hipEvent_t e1) {
    switch (nv) {
#define X(v) case v: return prec == TBS_FP32 ? launch_thomas<v, float>(st, sys, ltab, nl, variant, e0, e1) \
                                         : launch_thomas<v, double>(st, sys, ltab, nl, variant, e0, e1);
        TBS_WIDE_NVARS(X)
#undef X
    }
    set_error("nvar %d not built for the fused path", nv);
    return TBS_ERR_UNSUPPORTED;
}

int factor(int nv, int prec, hipStream_t st, const FactorRec &one, int nline, const FactorRec *recs,
           const int2 *ltab) {
    switch (nv) {
#define X(v) case v: \
        if (prec == TBS_FP32) launch_factor_nv<v, float>(st, one, nline, recs, ltab); \
        else launch_factor_nv<v, double>(st, one, nline, recs, ltab); \
        return TBS_OK;
        TBS_WIDE_NVARS(X)
#undef X
    }
    set_error("nvar %d not built for the fused path", nv);
    return TBS_ERR_UNSUPPORTED;
}

int pack(int nv, hipStream_t st, long total, int S, const int2 *src, const double *D, const double *O1,
         const double *O2, double *blk) {
    switch (nv) {
#define X(v) case v: \
        hipLaunchKernelGGL(pack_blocks_kernel<v>, dim3(gs_grid(total)), dim3(256), 0, st, total, S, src, D, O1, O2, \
                           blk); \
        return TBS_OK;
        TBS_WIDE_NVARS(X)
#undef X
    }
    set_error("nvar %d not built for the fused path", nv);
    return TBS_ERR_UNSUPPORTED;
}

}  // namespace TBS_WIDE_NS
}  // namespace tbs
