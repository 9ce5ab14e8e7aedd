// Split-precision DGEMM (filled in after the plain path).
#include "common.hpp"
#include "ops.hpp"
#include "runtime.hpp"

namespace zg {
int ozaki_dgemm(bool, bool, int, int, int, double, const double*, int64_t, const double*, int64_t, double, double*,
                int64_t, hipStream_t, int, int) {
    return ZYNUM_GPU_DECLINE_UNSUPPORTED;
}
}  // namespace zg
