// lij_wide_hi.hip -- fused-path kernel instances for block sizes 25..32.
#define TBS_WIDE_NVARS(X) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)
#define TBS_WIDE_NS wide_hi
#include "lij_wide_impl.hpp"
