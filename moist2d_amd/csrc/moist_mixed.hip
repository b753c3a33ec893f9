// moist_mixed.hip — the float (mixed-form) instantiations of the step kernels,
// compiled with FP contraction on: every a * b + c of these kernels may become
// one fused multiply-add (one rounding).  The FP64 form's kernels, compiled in
// moist_api.hip, keep separately rounded operations like the reference's
// Float64 arithmetic.  The pragma precedes the kernel headers, so it covers
// the template definitions these instantiations come from.
#pragma clang fp contract(fast)

#include "moist_kernels.h"

namespace mo {
MO_MIXED_KERNELS()
}  // namespace mo
