// DTRSM / DSYRK on the GEMM core (filled in below the GEMM core).
#include "common.hpp"
#include "gemm.hpp"
#include "ops.hpp"
#include "runtime.hpp"

namespace zg {

template <typename T>
int trsm_device(char, char, bool, char, int, int, T, const T*, int64_t, T*, int64_t, hipStream_t, int, int) {
    return ZYNUM_GPU_DECLINE_UNSUPPORTED;
}
template <typename T>
int syrk_device(char, bool, int, int, T, const T*, int64_t, T, T*, int64_t, hipStream_t, int, int) {
    return ZYNUM_GPU_DECLINE_UNSUPPORTED;
}
template int trsm_device<double>(char, char, bool, char, int, int, double, const double*, int64_t, double*, int64_t, hipStream_t, int, int);
template int trsm_device<float>(char, char, bool, char, int, int, float, const float*, int64_t, float*, int64_t, hipStream_t, int, int);
template int syrk_device<double>(char, bool, int, int, double, const double*, int64_t, double, double*, int64_t, hipStream_t, int, int);
template int syrk_device<float>(char, bool, int, int, float, const float*, int64_t, float, float*, int64_t, hipStream_t, int, int);

}  // namespace zg
