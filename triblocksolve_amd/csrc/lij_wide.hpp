// lij_wide.hpp -- the fused path's kernel launches for block sizes 17..32,
// instantiated in their own translation units (lij_wide_lo.hip: 17..24,
// lij_wide_hi.hip: 25..32) so that the library's sources compile in
// parallel.  Each returns TBS_ERR_UNSUPPORTED for a block size outside its
// range; the arguments are those of the launchers in lij_kernels.hpp.
#pragma once
#include "tbs_common.hpp"

namespace tbs {
struct FactorRec;
#define TBS_WIDE_DECLS                                                                                      \
    int residual(int nv, int prec, hipStream_t st, const void *sys, const int2 *gtab, int ng, int want_R,   \
                 int force, hipEvent_t e0, hipEvent_t e1);                                                 \
    int thomas(int nv, int prec, hipStream_t st, const void *sys, const int4 *ltab, int nl, int variant,    \
               hipEvent_t e0, hipEvent_t e1);                                                              \
    int factor(int nv, int prec, hipStream_t st, const FactorRec &one, int nline, const FactorRec *recs,    \
               const int2 *ltab);                                                                          \
    int pack(int nv, hipStream_t st, long total, int S, const int2 *src, const double *D, const double *O1, \
             const double *O2, double *blk);
namespace wide_lo { TBS_WIDE_DECLS }   // block sizes 17..24
namespace wide_hi { TBS_WIDE_DECLS }   // block sizes 25..32
#undef TBS_WIDE_DECLS
}  // namespace tbs
