// Host runtime of the GPU tier.  The reference runs every call on the
// caller's CPU threads (thread_pool.zig); here a call is a stream-ordered
// sequence of kernel launches on the current HIP device.
#include "runtime.hpp"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
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
hipEvent_t g_ws_done = nullptr;    // end of the last user's work
hipStream_t g_ws_stream = nullptr; // the stream it was recorded on
bool g_ws_pending = false;         // g_ws_done holds a record

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

// Host threads for the copies between caller memory and the pinned bounce
// buffers: one core copies pageable memory at about half the rate of the DMA
// engine, several together at more than it.  The caller takes parts too.
// Workers are created on first use and parked on a condition variable; they
// are detached and end with the process.
class CopyPool {
  public:
    static CopyPool& get() {
        static CopyPool* pool = new CopyPool();
        return *pool;
    }

    // Copy threads in use, the caller included (requested through
    // zynum_gpu_set_copy_threads; 0 selects the default).
    static int default_threads() {
        const unsigned hw = std::thread::hardware_concurrency();
        return hw > 0 ? std::min(8, (int)std::max(1u, hw / 2)) : 8;
    }
    static std::atomic<int>& requested() {
        static std::atomic<int> r{0};
        return r;
    }
    int threads() const {
        const int r = requested().load();
        return r > 0 ? r : default_threads();
    }

    // Run fn(0) .. fn(parts - 1), spread over threads() - 1 workers and the
    // caller.  A forked child has none of the parent's workers: it copies alone.
    void run(int parts, const std::function<void(int)>& fn) {
        const int use = threads();
        if (parts <= 1 || use <= 1 || forked_.load()) {
            for (int i = 0; i < parts; ++i) fn(i);
            return;
        }
        std::lock_guard<std::mutex> serial(run_mu_);  // one job at a time
        grow(use - 1);
        {
            std::lock_guard<std::mutex> lock(mu_);
            fn_ = &fn;
            parts_ = parts;
            active_ = use - 1;
            next_.store(0);
            pending_ = (int)workers_;
            ++generation_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lock(mu_);
        done_cv_.wait(lock, [this] { return pending_ == 0; });
        fn_ = nullptr;
    }

  private:
    CopyPool() {
        pthread_atfork(nullptr, nullptr, [] { CopyPool::get().forked_.store(true); });
    }

    // Start workers up to `count` (held under run_mu_); they never exit.
    void grow(int count) {
        std::lock_guard<std::mutex> lock(mu_);
        while (workers_ < count) {
            const int index = workers_++;
            const uint64_t gen = generation_;  // it takes part from the next job on
            std::thread([this, index, gen] { loop(index, gen); }).detach();
        }
    }

    void work() {
        for (int i = next_.fetch_add(1); i < parts_; i = next_.fetch_add(1)) (*fn_)(i);
    }

    void loop(int index, uint64_t seen) {
        for (;;) {
            bool take = false;
            {
                std::unique_lock<std::mutex> lock(mu_);
                cv_.wait(lock, [&] { return generation_ != seen; });
                seen = generation_;
                take = index < active_;
            }
            if (take) work();
            std::lock_guard<std::mutex> lock(mu_);
            if (--pending_ == 0) done_cv_.notify_one();
        }
    }

    int workers_ = 0;  // started so far
    int active_ = 0;   // workers taking parts of the current job
    std::atomic<bool> forked_{false};
    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)>* fn_ = nullptr;
    int parts_ = 0;
    std::atomic<int> next_{0};
    int pending_ = 0;
    uint64_t generation_ = 0;
};

// Copy the logical rows x cols part of a column-major matrix between two
// buffers with the same leading dimension, over the copy threads when it is
// large (pieces of at least 1 MiB; one memcpy per piece without padding).
void copy_columns(void* dst, const void* src, int64_t rows, int64_t cols, int64_t ld, size_t es) {
    const size_t total = (size_t)(rows * cols) * es;
    char* d = static_cast<char*>(dst);
    const char* s = static_cast<const char*>(src);
    CopyPool& pool = CopyPool::get();
    const int parts = (int)std::min<size_t>((size_t)pool.threads() * 2, std::max<size_t>(1, total >> 20));
    if (rows == ld) {
        const size_t piece = (total + parts - 1) / parts;
        pool.run(parts, [&](int i) {
            const size_t off = (size_t)i * piece;
            if (off < total) memcpy(d + off, s + off, std::min(piece, total - off));
        });
        return;
    }
    const int64_t per = (cols + parts - 1) / parts;
    pool.run(parts, [&](int i) {
        for (int64_t j = (int64_t)i * per; j < std::min(cols, (int64_t)(i + 1) * per); ++j)
            memcpy(d + (size_t)(j * ld) * es, s + (size_t)(j * ld) * es, (size_t)rows * es);
    });
}

// Column ranges of a staged matrix are moved in chunks so that the host copy
// of one chunk overlaps the DMA transfer of the previous one (the bounce
// buffer holds the whole extent, so no chunk is overwritten in flight).
int64_t chunk_columns(int64_t rows, int64_t cols, int64_t ld, size_t es) {
    const size_t bytes = matrix_extent(rows, cols, ld) * es;
    const size_t target = std::min<size_t>(size_t(32) << 20, std::max<size_t>(size_t(4) << 20, bytes / 4));
    const size_t col_bytes = (size_t)ld * es;
    return std::max<int64_t>(1, std::min<int64_t>(cols, (int64_t)(target / col_bytes)));
}

// Bytes of columns [j0, j1) of the extent (the last column ends at row `rows`).
size_t chunk_bytes(int64_t rows, int64_t cols, int64_t ld, size_t es, int64_t j0, int64_t j1) {
    return (size_t)((j1 == cols ? (j1 - j0 - 1) * ld + rows : (j1 - j0) * ld)) * es;
}

// Events marking the end of each device-to-host chunk (grown on demand, kept
// per host thread).
thread_local std::vector<hipEvent_t> t_chunk_events;

hipEvent_t chunk_event(size_t i) {
    while (t_chunk_events.size() <= i) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        t_chunk_events.push_back(e);
    }
    return t_chunk_events[i];
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

namespace {
// The library's own stream per device: non-blocking and at default priority,
// so a call does not serialise against unrelated streams of the caller.  It
// is created on first use and lives until the process ends.
struct LibStream {
    hipStream_t stream = nullptr;
    hipEvent_t after_null = nullptr;  // records the null stream's queue at each call
};
constexpr int kLibDevices = 64;
LibStream g_lib[kLibDevices];
std::mutex g_lib_mutex;

LibStream* lib_stream_for_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kLibDevices) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_lib_mutex);
    LibStream& l = g_lib[dev];
    if (l.stream == nullptr) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            if (ev) (void)hipEventDestroy(ev);
            if (st) (void)hipStreamDestroy(st);
            return nullptr;
        }
        l.after_null = ev;
        l.stream = st;
    }
    return &l;
}
}  // namespace

hipStream_t caller_stream() { return g_blas_stream.load(); }

hipStream_t blas_stream() {
    hipStream_t s = g_blas_stream.load();
    if (s != nullptr) return s;
    // No caller stream: the library stream, ordered after everything queued so
    // far on the device's null stream (where a caller that names no stream has
    // written the operands), as a synchronous BLAS call expects.
    LibStream* l = lib_stream_for_current_device();
    if (l == nullptr) return nullptr;
    if (hipEventRecord(l->after_null, nullptr) != hipSuccess || hipStreamWaitEvent(l->stream, l->after_null, 0) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return l->stream;
}

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
    if (!copy_in) return ZYNUM_GPU_OK;
    const int64_t step = chunk_columns(rows, cols, ld, elem);
    for (int64_t j0 = 0; j0 < cols; j0 += step) {
        const int64_t j1 = std::min(cols, j0 + step);
        const size_t off = (size_t)(j0 * ld) * elem;
        copy_columns(static_cast<char*>(bounce) + off, static_cast<const char*>(p) + off, rows, j1 - j0, ld, elem);
        hipError_t e = hipMemcpyAsync(static_cast<char*>(owned) + off, static_cast<char*>(bounce) + off,
                                      chunk_bytes(rows, cols, ld, elem, j0, j1), hipMemcpyHostToDevice, stream);
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
    // Queue every chunk's transfer with an event behind it, then copy each
    // chunk's logical part to the caller as soon as its transfer is done.
    const int64_t step = chunk_columns(rows, cols, ld, elem);
    hipError_t e = hipSuccess;
    size_t nchunks = 0;
    for (int64_t j0 = 0; j0 < cols && e == hipSuccess; j0 += step, ++nchunks) {
        const int64_t j1 = std::min(cols, j0 + step);
        const size_t off = (size_t)(j0 * ld) * elem;
        e = hipMemcpyAsync(static_cast<char*>(bounce) + off, static_cast<char*>(owned) + off,
                           chunk_bytes(rows, cols, ld, elem, j0, j1), hipMemcpyDeviceToHost, s);
        hipEvent_t ev = chunk_event(nchunks);
        if (e == hipSuccess) e = ev ? hipEventRecord(ev, s) : hipErrorOutOfMemory;
    }
    if (e == hipSuccess) {
        size_t i = 0;
        for (int64_t j0 = 0; j0 < cols && e == hipSuccess; j0 += step, ++i) {
            const int64_t j1 = std::min(cols, j0 + step);
            const size_t off = (size_t)(j0 * ld) * elem;
            e = hipEventSynchronize(t_chunk_events[i]);
            if (e == hipSuccess)  // the logical part only
                copy_columns(static_cast<char*>(host_out) + off, static_cast<char*>(bounce) + off, rows, j1 - j0, ld,
                             elem);
        }
    }
    if (e != hipSuccess) (void)hipStreamSynchronize(s);
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

namespace {
struct AuxStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
constexpr int kAuxDevices = 64;
AuxStream g_aux[kAuxDevices];
std::mutex g_aux_mu;  // creation only: callers order their use of the stream with events

AuxStream* aux_for_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kAuxDevices) {
        (void)hipGetLastError();
        return nullptr;
    }
    AuxStream& a = g_aux[dev];
    std::lock_guard<std::mutex> lock(g_aux_mu);
    if (a.stream == nullptr) {
        hipStream_t st = nullptr;
        hipEvent_t f = nullptr, j = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&f, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&j, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            if (j) (void)hipEventDestroy(j);
            if (f) (void)hipEventDestroy(f);
            if (st) (void)hipStreamDestroy(st);
            return nullptr;
        }
        a.fork = f;
        a.join = j;
        a.stream = st;
    }
    return &a;
}
}  // namespace

bool aux_fork(hipStream_t s, hipStream_t* side) {
    AuxStream* a = aux_for_current_device();
    if (a == nullptr) return false;
    if (hipEventRecord(a->fork, s) != hipSuccess || hipStreamWaitEvent(a->stream, a->fork, 0) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    *side = a->stream;
    return true;
}

hipStream_t aux_stream() {
    AuxStream* a = aux_for_current_device();
    return a ? a->stream : nullptr;
}

hipError_t aux_join(hipStream_t s, hipStream_t side) {
    AuxStream* a = aux_for_current_device();
    if (a == nullptr || a->stream != side) return hipErrorInvalidValue;
    hipError_t e = hipEventRecord(a->join, side);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, a->join, 0);
    return e;
}

void* scratch_acquire(size_t bytes, size_t* got) { return g_stage_dev.acquire(bytes, got); }

void scratch_release(void* p, size_t bytes) { g_stage_dev.release(p, bytes); }

void* workspace_acquire(size_t bytes, hipStream_t s) {
    g_ws_mutex.lock();
    if (bytes <= g_ws_bytes) {
        if (g_ws_pending && g_ws_stream != s && hipStreamWaitEvent(s, g_ws_done, 0) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipEventSynchronize(g_ws_done);
        }
        return g_ws_ptr;
    }
    if (g_ws_ptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(g_ws_ptr);
        g_ws_ptr = nullptr;
        g_ws_bytes = 0;
    }
    g_ws_pending = false;
    if (hipMalloc(&g_ws_ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        g_ws_ptr = nullptr;
        g_ws_mutex.unlock();
        return nullptr;
    }
    g_ws_bytes = bytes;
    return g_ws_ptr;
}

void workspace_release(hipStream_t s) {
    if (g_ws_done == nullptr && hipEventCreateWithFlags(&g_ws_done, hipEventDisableTiming) != hipSuccess)
        g_ws_done = nullptr;
    if (g_ws_done != nullptr && hipEventRecord(g_ws_done, s) == hipSuccess) {
        g_ws_stream = s;
        g_ws_pending = true;
    } else {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s);  // no event: the next user finds the work done
        g_ws_pending = false;
    }
    g_ws_mutex.unlock();
}

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
    g_ws_pending = false;
}

}  // namespace zg

extern "C" void zynum_gpu_set_stream(void* stream) { zg::g_blas_stream.store(reinterpret_cast<hipStream_t>(stream)); }

extern "C" int zynum_gpu_set_copy_threads(int threads) {
    if (threads < 0 || threads > 64) return ZYNUM_GPU_DECLINE_ARGUMENT;
    zg::CopyPool::requested().store(threads);
    return ZYNUM_GPU_OK;
}

extern "C" void zynum_gpu_debug_force_decline(int count) { zg::g_forced.store(count > 0 ? count : 0); }
