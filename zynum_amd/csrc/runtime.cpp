// Host runtime of the GPU tier.  The reference runs every call on the
// caller's CPU threads (thread_pool.zig); here a call is a stream-ordered
// sequence of kernel launches on the current HIP device.
#include "runtime.hpp"

#include <atomic>
#include <mutex>
#include <string.h>

namespace zg {

namespace {
std::once_flag g_probe_once;
bool g_ready = false;
std::atomic<hipStream_t> g_blas_stream{nullptr};
std::atomic<int> g_forced{0};

std::mutex g_ws_mutex;
void* g_ws_ptr = nullptr;
size_t g_ws_bytes = 0;
}  // namespace

bool device_ready() {
    std::call_once(g_probe_once, [] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
            (void)hipGetLastError();
            return;
        }
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
        // The code objects are built for gfx950 only.
        g_ready = strncmp(prop.gcnArchName, "gfx950", 6) == 0;
    });
    return g_ready;
}

hipStream_t blas_stream() { return g_blas_stream.load(); }

bool take_forced_decline() {
    int v = g_forced.load();
    while (v > 0) {
        if (g_forced.compare_exchange_weak(v, v - 1)) return true;
    }
    return false;
}

bool is_device_pointer(const void* p) {
    if (p == nullptr) return false;
    hipPointerAttribute_t attr;
    memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

int status_from_hip(hipError_t e) {
    if (e == hipSuccess) return ZYNUM_GPU_OK;
    (void)hipGetLastError();
    return ZYNUM_GPU_ERR_LAUNCH;
}

int Operand::bind_in(const void* p, int64_t r, int64_t c, int64_t ldx, size_t es, hipStream_t s) {
    rows = r; cols = c; ld = ldx; elem = es; host = p;
    if (r <= 0 || c <= 0) { dev = const_cast<void*>(p); return ZYNUM_GPU_OK; }
    if (is_device_pointer(p)) { dev = const_cast<void*>(p); return ZYNUM_GPU_OK; }
    staged = true;
    size_t bytes = matrix_extent(r, c, ldx) * es;
    if (hipMalloc(&owned, bytes) != hipSuccess) {
        (void)hipGetLastError();
        owned = nullptr;
        return ZYNUM_GPU_DECLINE_ALLOC;
    }
    dev = owned;
    hipError_t e = hipMemcpy2DAsync(owned, (size_t)ldx * es, p, (size_t)ldx * es, (size_t)r * es, (size_t)c,
                                    hipMemcpyHostToDevice, s);
    if (e != hipSuccess) { (void)hipGetLastError(); return ZYNUM_GPU_DECLINE_ALLOC; }
    return ZYNUM_GPU_OK;
}

int Operand::bind_out(void* p, int64_t r, int64_t c, int64_t ldx, size_t es, bool copy_in, hipStream_t s) {
    rows = r; cols = c; ld = ldx; elem = es; host_out = p;
    if (r <= 0 || c <= 0) { dev = p; return ZYNUM_GPU_OK; }
    if (is_device_pointer(p)) { dev = p; return ZYNUM_GPU_OK; }
    staged = true;
    size_t bytes = matrix_extent(r, c, ldx) * es;
    if (hipMalloc(&owned, bytes) != hipSuccess) {
        (void)hipGetLastError();
        owned = nullptr;
        return ZYNUM_GPU_DECLINE_ALLOC;
    }
    dev = owned;
    if (copy_in) {
        hipError_t e = hipMemcpy2DAsync(owned, (size_t)ldx * es, p, (size_t)ldx * es, (size_t)r * es, (size_t)c,
                                        hipMemcpyHostToDevice, s);
        if (e != hipSuccess) { (void)hipGetLastError(); return ZYNUM_GPU_DECLINE_ALLOC; }
    }
    return ZYNUM_GPU_OK;
}

int Operand::finish(hipStream_t s) {
    if (!staged || host_out == nullptr) return ZYNUM_GPU_OK;
    hipError_t e = hipMemcpy2DAsync(host_out, (size_t)ld * elem, owned, (size_t)ld * elem, (size_t)rows * elem,
                                    (size_t)cols, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return status_from_hip(e);
}

void Operand::release() {
    if (owned) {
        (void)hipFree(owned);
        owned = nullptr;
    }
}

void* workspace_acquire(size_t bytes) {
    g_ws_mutex.lock();
    if (bytes <= g_ws_bytes) return g_ws_ptr;
    if (g_ws_ptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(g_ws_ptr);
        g_ws_ptr = nullptr;
        g_ws_bytes = 0;
    }
    if (hipMalloc(&g_ws_ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        g_ws_ptr = nullptr;
        g_ws_mutex.unlock();
        return nullptr;
    }
    g_ws_bytes = bytes;
    return g_ws_ptr;
}

void workspace_release() { g_ws_mutex.unlock(); }

void workspace_free_all() {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if (g_ws_ptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(g_ws_ptr);
    }
    g_ws_ptr = nullptr;
    g_ws_bytes = 0;
}

}  // namespace zg

extern "C" void zynum_gpu_set_stream(void* stream) { zg::g_blas_stream.store(reinterpret_cast<hipStream_t>(stream)); }

extern "C" void zynum_gpu_debug_force_decline(int count) { zg::g_forced.store(count > 0 ? count : 0); }
