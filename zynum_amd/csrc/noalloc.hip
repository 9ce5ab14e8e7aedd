// A GEMM request entry point that allocates nothing, for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops).
//
// It runs the plain FP64/FP32 MFMA GEMM of zynum_gpu_gemm on device operands.
// The only device memory beyond the caller's A, B and C is the split-K
// partial buffer (shapes with few output tiles and a long K), which here is a
// fixed file-scope device array instead of the staging-buffer cache.  The
// split-K plan bounds that buffer: fewer than 128 output tiles of 64 x 64 and
// at most ceil(512 / tiles) K ranges give parts * m * n <= (512 + 127) * 4096
// elements, 20.9 MB in FP64, so every plan fits the 24 MiB array.
//
// Every call on a device shares that device's array, which keeps the last
// call's partials; nothing needs clearing, because a split call writes every
// partial it reads.  Calls that overlap on one device race on the array.
#include <mutex>

#include "abi_common.hpp"
#include "gemm.hpp"
#include "runtime.hpp"

namespace zg {
namespace {

constexpr size_t kScratchBytes = size_t(24) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::once_flag g_scratch_once[kMaxDevices];
int g_scratch_status[kMaxDevices] = {};

// The device address of `scratch` on the current device (each device holds
// its own copy of the array): looked up once per device with that device
// current.  The array's name is a host-side stand-in and is never passed on.
int scratch_ptr(void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return ZYNUM_GPU_DECLINE_NO_DEVICE;
    }
    std::call_once(g_scratch_once[dev], [dev] {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            g_scratch_status[dev] = ZYNUM_GPU_ERR_LAUNCH;
            return;
        }
        g_scratch_ptr[dev] = p;
    });
    if (g_scratch_status[dev] != ZYNUM_GPU_OK) return g_scratch_status[dev];
    *out = g_scratch_ptr[dev];
    return ZYNUM_GPU_OK;
}

template <typename T>
int gemm_noalloc(const zynum_gpu_gemm_request* r, hipStream_t s) {
    const bool ta = r->transa != ZYNUM_GPU_NO_TRANS, tb = r->transb != ZYNUM_GPU_NO_TRANS;
    const T alpha = scalar_from_bits<T>(r->alpha_re), beta = scalar_from_bits<T>(r->beta_re);
    const int m = r->m, n = r->n, k = r->k;
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;  // gemm.zig:1393
    if (take_forced_decline()) return ZYNUM_GPU_DECLINE_FORCED;
    if (!device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    const bool reads_ab = (k > 0) && (alpha != T(0));
    // device operands only: staging host memory would need buffers
    if (!is_device_pointer(r->c) || (reads_ab && (!is_device_pointer(r->a) || !is_device_pointer(r->b))))
        return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    GemmDesc<T> d;
    d.ta = ta; d.tb = tb; d.m = m; d.n = n; d.k = k;
    d.lda = r->lda; d.ldb = r->ldb; d.ldc = r->ldc;
    d.alpha = alpha; d.beta = beta;
    d.a = static_cast<const T*>(r->a); d.b = static_cast<const T*>(r->b); d.c = static_cast<T*>(r->c);
    const int tile = r->tile == 0 ? kTileAuto : r->tile - 1;
    int kchunk = 0;
    const int parts = (reads_ab && tile == kTileAuto) ? gemm_splitk_plan(m, n, k, &kchunk) : 1;
    hipError_t e;
    if (parts > 1) {
        if ((size_t)parts * (size_t)m * (size_t)n * sizeof(T) > kScratchBytes) return ZYNUM_GPU_DECLINE_SCRATCH;
        void* part = nullptr;
        const int st = scratch_ptr(&part);
        if (st != ZYNUM_GPU_OK) return st;
        e = gemm_splitk_launch<T>(d, s, static_cast<T*>(part), parts, kchunk);
    } else {
        e = gemm_launch<T>(d, s, tile);
    }
    if (e == hipSuccess && !(r->flags & ZYNUM_GPU_FLAG_ASYNC)) e = hipStreamSynchronize(s);
    return status_from_hip(e);
}

}  // namespace
}  // namespace zg

extern "C" int zynum_gpu_gemm_noalloc(const zynum_gpu_gemm_request* r) {
    if (r == nullptr) return ZYNUM_GPU_DECLINE_ARGUMENT;
    const char ta = zg::request_trans_char(r->transa, false), tb = zg::request_trans_char(r->transb, false);
    if (zg::gemm_error(ta, tb, r->m, r->n, r->k, r->lda, r->ldb, r->ldc) != 0) return ZYNUM_GPU_DECLINE_ARGUMENT;
    if (r->scalar != ZYNUM_GPU_F64 && r->scalar != ZYNUM_GPU_F32) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    if (r->mode != ZYNUM_GPU_MODE_PLAIN || r->tile > zg::kTileConfigs) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    // the request's stream, else the one set by zynum_gpu_set_stream, else the null stream
    // (the library's own stream is created on first use, so it is not taken here)
    hipStream_t s = r->stream ? reinterpret_cast<hipStream_t>(r->stream) : zg::caller_stream();
    return r->scalar == ZYNUM_GPU_F64 ? zg::gemm_noalloc<double>(r, s) : zg::gemm_noalloc<float>(r, s);
}
