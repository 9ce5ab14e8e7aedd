// Host runtime of the GPU tier.  The reference runs every call on the
// caller's CPU threads (thread_pool.zig); here a call is a stream-ordered
// sequence of kernel launches on the current HIP device.
#include "runtime.hpp"

#include <atomic>
#include <mutex>
#include <string.h>
#include <vector>

namespace zg {

namespace {
std::once_flag g_probe_once;
bool g_ready = false;
std::atomic<hipStream_t> g_blas_stream{nullptr};
std::atomic<int> g_forced{0};

std::mutex g_ws_mutex;
void* g_ws_ptr = nullptr;
size_t g_ws_bytes = 0;

// Staging buffers for host operands: device buffers the kernels read and
// pinned host bounce buffers the DMA engine copies to and from in one
// transfer.  A host-pointer call would otherwise pay a hipMalloc / hipFree
// pair (and the device-wide wait of hipFree) per operand, and a 2D copy from
// pageable memory moves one column per transfer.  Idle buffers are kept up to
// a byte cap and handed to any operand that fits in at most twice its size.
class BufCache {
  public:
    BufCache(bool pinned, size_t cap) : pinned_(pinned), cap_(cap) {}

    void* acquire(size_t bytes, size_t* got) {
        {
            std::lock_guard<std::mutex> lock(mu_);
            size_t best = free_.size();
            for (size_t i = 0; i < free_.size(); ++i) {
                const size_t b = free_[i].bytes;
                if (b >= bytes && b <= 2 * bytes + kGranule && (best == free_.size() || b < free_[best].bytes))
                    best = i;
            }
            if (best < free_.size()) {
                Buf buf = free_[best];
                free_.erase(free_.begin() + (ptrdiff_t)best);
                cached_ -= buf.bytes;
                *got = buf.bytes;
                return buf.ptr;
            }
        }
        const size_t want = (bytes + kGranule - 1) / kGranule * kGranule;
        void* p = nullptr;
        const hipError_t e = pinned_ ? hipHostMalloc(&p, want, hipHostMallocDefault) : hipMalloc(&p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        *got = want;
        return p;
    }

    // The caller has synchronised the stream that used the buffer.
    void release(void* p, size_t bytes) {
        {
            std::lock_guard<std::mutex> lock(mu_);
            if (cached_ + bytes <= cap_) {
                free_.push_back({p, bytes});
                cached_ += bytes;
                return;
            }
        }
        drop(p);
    }

    void free_all() {
        std::lock_guard<std::mutex> lock(mu_);
        for (const Buf& b : free_) drop(b.ptr);
        free_.clear();
        cached_ = 0;
    }

  private:
    struct Buf {
        void* ptr;
        size_t bytes;
    };
    static constexpr size_t kGranule = size_t(2) << 20;
    void drop(void* p) { (void)(pinned_ ? hipHostFree(p) : hipFree(p)); }
    bool pinned_;
    size_t cap_;
    std::mutex mu_;
    std::vector<Buf> free_;
    size_t cached_ = 0;
};

BufCache g_stage_dev(false, size_t(8) << 30);
BufCache g_stage_host(true, size_t(2) << 30);

// Copy the logical rows x cols part of a column-major matrix between two
// buffers with the same leading dimension (one memcpy when there is no padding).
void copy_columns(void* dst, const void* src, int64_t rows, int64_t cols, int64_t ld, size_t es) {
    if (rows == ld) {
        memcpy(dst, src, (size_t)(rows * cols) * es);
        return;
    }
    for (int64_t j = 0; j < cols; ++j)
        memcpy(static_cast<char*>(dst) + (size_t)(j * ld) * es, static_cast<const char*>(src) + (size_t)(j * ld) * es,
               (size_t)rows * es);
}
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

int Operand::stage(const void* p, bool copy_in) {
    const size_t bytes = matrix_extent(rows, cols, ld) * elem;
    owned = g_stage_dev.acquire(bytes, &owned_bytes);
    if (owned == nullptr) return ZYNUM_GPU_DECLINE_ALLOC;
    bounce = g_stage_host.acquire(bytes, &bounce_bytes);
    if (bounce == nullptr) return ZYNUM_GPU_DECLINE_ALLOC;
    dev = owned;
    if (copy_in) {
        copy_columns(bounce, p, rows, cols, ld, elem);
        hipError_t e = hipMemcpyAsync(owned, bounce, bytes, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) { (void)hipGetLastError(); return ZYNUM_GPU_DECLINE_ALLOC; }
    }
    return ZYNUM_GPU_OK;
}

int Operand::bind_in(const void* p, int64_t r, int64_t c, int64_t ldx, size_t es, hipStream_t s) {
    rows = r; cols = c; ld = ldx; elem = es; host = p;
    if (r <= 0 || c <= 0) { dev = const_cast<void*>(p); return ZYNUM_GPU_OK; }
    if (is_device_pointer(p)) { dev = const_cast<void*>(p); return ZYNUM_GPU_OK; }
    staged = true;
    stream = s;
    return stage(p, true);
}

int Operand::bind_out(void* p, int64_t r, int64_t c, int64_t ldx, size_t es, bool copy_in, hipStream_t s) {
    rows = r; cols = c; ld = ldx; elem = es; host_out = p;
    if (r <= 0 || c <= 0) { dev = p; return ZYNUM_GPU_OK; }
    if (is_device_pointer(p)) { dev = p; return ZYNUM_GPU_OK; }
    staged = true;
    stream = s;
    return stage(p, copy_in);
}

int Operand::finish(hipStream_t s) {
    if (!staged || host_out == nullptr) return ZYNUM_GPU_OK;
    hipError_t e = hipMemcpyAsync(bounce, owned, matrix_extent(rows, cols, ld) * elem, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) copy_columns(host_out, bounce, rows, cols, ld, elem);  // the logical part only
    return status_from_hip(e);
}

void Operand::release() {
    if (owned || bounce) {
        // idle already after a completed op; this also covers error returns
        (void)hipStreamSynchronize(stream);
        if (owned) g_stage_dev.release(owned, owned_bytes);
        if (bounce) g_stage_host.release(bounce, bounce_bytes);
        owned = nullptr;
        bounce = nullptr;
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
    g_stage_dev.free_all();
    g_stage_host.free_all();
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
