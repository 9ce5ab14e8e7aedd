// Host runtime of the GPU tier: device checks, stream choice, operand
// residency, staging of host operands and the shared device workspace.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/zynum_gpu.h"

namespace zg {

// True when a gfx950 device is present and current (probed once per process).
bool device_ready();

// Stream used by the classic BLAS symbols: the one set by
// zynum_gpu_set_stream, else the library's own stream (created on first use).
hipStream_t blas_stream();
// The stream set by zynum_gpu_set_stream (nullptr if none); creates nothing.
hipStream_t caller_stream();

// Consumes one injected decline; true if the call must decline.
bool take_forced_decline();

// Residency of a pointer: true for device or managed memory.
bool is_device_pointer(const void* p);

// Elements spanned by a column-major rows x cols matrix with leading dim ld.
inline size_t matrix_extent(int64_t rows, int64_t cols, int64_t ld) {
    if (rows <= 0 || cols <= 0) return 0;
    return (size_t)(rows + (cols - 1) * ld);
}

// A device view of a caller matrix.  Host matrices are copied to a device
// buffer through a pinned bounce buffer (one DMA transfer of the whole extent;
// only the logical rows x cols part is read from or written to the caller)
// and, for outputs, copied back by finish().  Device matrices are used in place.
struct Operand {
    void* dev = nullptr;       // pointer handed to kernels
    void* owned = nullptr;     // staging buffer (host operand only), from the staging cache
    size_t owned_bytes = 0;
    void* bounce = nullptr;    // pinned host bounce buffer of the same extent (host operand only)
    size_t bounce_bytes = 0;
    hipStream_t stream = nullptr;  // stream of the staging copies
    const void* host = nullptr;
    void* host_out = nullptr;
    int64_t rows = 0, cols = 0, ld = 0;
    size_t elem = 0;
    bool staged = false;

    // Prepare an input (copied in when staged).
    int bind_in(const void* p, int64_t r, int64_t c, int64_t ldx, size_t es, hipStream_t s);
    // Prepare an output; `copy_in` when its old values are read (beta != 0).
    int bind_out(void* p, int64_t r, int64_t c, int64_t ldx, size_t es, bool copy_in, hipStream_t s);
    // Copy a staged output back to the host.
    int finish(hipStream_t s);
    // Wait for the staging stream, then return the buffers to their caches.
    void release();
    int stage(const void* p, bool copy_in);
    ~Operand() { release(); }
};

// Shared grow-only device workspace (one per process, current device).  The
// lock is held by the caller for the whole stream-ordered use.
// workspace_free_all also frees the cached staging buffers.
// The shared device workspace, stream-ordered: acquire(bytes, s) holds the
// lock and returns the buffer (nullptr on allocation failure), ready for work
// queued on s (the previous user's work, if it was queued on another stream,
// is waited for on the device); release(s) records the end of this user's
// work on s and drops the lock.  Work of a user that also ran on other
// streams must be joined into s before release.
void* workspace_acquire(size_t bytes, hipStream_t s);
void workspace_release(hipStream_t s);
void workspace_free_all();

// Device scratch memory from the staging-buffer cache (no lock held, unlike
// the workspace).  Release only after the stream that used it is synchronised.
void* scratch_acquire(size_t bytes, size_t* got);  // nullptr on allocation failure
void scratch_release(void* p, size_t bytes);

// A second stream (non-blocking, default priority) per device for work of a
// call that is independent of what its stream is doing: aux_fork makes it
// wait for everything queued so far on `s`; aux_join makes `s` wait for
// everything queued on it.  The fork/join events are shared, so callers hold
// the workspace lock between the two.  aux_fork returns false (nothing forked)
// when the stream or events cannot be created.
bool aux_fork(hipStream_t s, hipStream_t* side);
hipError_t aux_join(hipStream_t s, hipStream_t side);
// The same second stream, for a caller that orders it with events of its own
// (null when it cannot be created).
hipStream_t aux_stream();

int status_from_hip(hipError_t e);

}  // namespace zg
