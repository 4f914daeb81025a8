// lij_wide_lo.hip -- fused-path kernel instances for block sizes 17..24.
#define TBS_WIDE_NVARS(X) X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24)
#define TBS_WIDE_NS wide_lo
#include "lij_wide_impl.hpp"
