// GPU-vs-CPU crossover for host-pointer calls.
//
// The reference's planner decides how many CPU threads a GEMM gets from
// fixed work thresholds (planner.zig:13,88-154: work = m*n*k against 96^3,
// 128^3, 192^3, 256^3, 384^3).  A planner with a GPU tier has one more
// choice: hand the call to this library, which stages host operands through
// pinned bounce buffers, or keep it on the CPU.  These functions predict the
// wall time of a synchronous GPU-tier call from a model measured on the
// target machine (offload_tune.inc, written by zynum_amd/autotune.py) and
// compare it with the CPU tier's time at the CPU rate the caller states.
//
// Model: t = t_device(work) + [any host operand] * t_call
//          + bytes_in * ns_in + bytes_out * ns_out
// where t_device is the measured time of a device-resident synchronous call
// of a square problem of equal work m*n*k (log-log interpolation between the
// measured sizes, linear in work beyond the largest), and the staging terms
// are a least-squares fit of the extra time of host-operand calls over shapes
// that separate input and output bytes.  The model is pure host arithmetic: it
// touches no device and works on machines without a GPU.
#include <math.h>
#include <stdint.h>

#include "runtime.hpp"

namespace zg {
namespace {

struct OffloadPoint {
    int scalar;    // 0 f32, 1 f64, 2 complex f32, 3 complex f64
    double work;   // m*n*k of the measured square problem
    double ns;     // best wall time of one synchronous call, device operands
};

struct OffloadStage {
    double call_ns;      // fixed extra time of a call with host operands
    double in_ns_per_b;  // per byte copied host -> device (pageable source)
    double out_ns_per_b; // per byte copied device -> host
};

#include "offload_tune.inc"

constexpr int kScalars = 4;
constexpr double kElemBytes[kScalars] = {4.0, 8.0, 8.0, 16.0};
constexpr double kFlopsPerMac[kScalars] = {2.0, 2.0, 8.0, 8.0};

bool valid(int scalar, int64_t m, int64_t n, int64_t k, int host_mask) {
    return scalar >= 0 && scalar < kScalars && m >= 0 && n >= 0 && k >= 0 && host_mask >= 0 && host_mask <= 7;
}

double device_ns(int scalar, double work) {
    const OffloadPoint* prev = nullptr;
    for (const OffloadPoint& p : kOffloadDevice) {
        if (p.scalar != scalar) continue;
        if (work <= p.work) {
            if (prev == nullptr) return p.ns;
            const double t = (log(work) - log(prev->work)) / (log(p.work) - log(prev->work));
            return exp(log(prev->ns) + t * (log(p.ns) - log(prev->ns)));
        }
        prev = &p;
    }
    return prev ? prev->ns * work / prev->work : -1.0;
}

}  // namespace
}  // namespace zg

using zg::kOffloadStage;

extern "C" double zynum_gpu_offload_predict_ns(int scalar, int64_t m, int64_t n, int64_t k, int host_mask,
                                               int beta_nonzero) {
    if (!zg::valid(scalar, m, n, k, host_mask)) return -1.0;
    if (m == 0 || n == 0) return 0.0;  // quick return: nothing is launched
    const double md = (double)m, nd = (double)n, kd = (double)k;
    // k = 0 is the beta-scaling pass: m*n element updates
    const double t_dev = zg::device_ns(scalar, md * nd * (k > 0 ? kd : 1.0));
    if (t_dev < 0.0) return -1.0;
    const double es = zg::kElemBytes[scalar];
    double in_bytes = 0.0, out_bytes = 0.0;
    if ((host_mask & 1) && k > 0) in_bytes += md * kd * es;
    if ((host_mask & 2) && k > 0) in_bytes += kd * nd * es;
    if (host_mask & 4) {
        out_bytes += md * nd * es;
        if (beta_nonzero) in_bytes += md * nd * es;
    }
    const double t_stage = host_mask ? kOffloadStage.call_ns + in_bytes * kOffloadStage.in_ns_per_b +
                                           out_bytes * kOffloadStage.out_ns_per_b
                                     : 0.0;
    return t_dev + t_stage;
}

extern "C" int zynum_gpu_offload_choose(int scalar, int64_t m, int64_t n, int64_t k, int host_mask,
                                        int beta_nonzero, double cpu_gflops) {
    if (!zg::valid(scalar, m, n, k, host_mask) || !(cpu_gflops > 0.0) || isinf(cpu_gflops))
        return -ZYNUM_GPU_DECLINE_ARGUMENT;
    if (m == 0 || n == 0 || k == 0) return 0;  // no product to compute
    const double flops = zg::kFlopsPerMac[scalar] * (double)m * (double)n * (double)k;
    const double cpu_ns = flops / cpu_gflops;  // GFLOP/s = FLOP per ns
    return zynum_gpu_offload_predict_ns(scalar, m, n, k, host_mask, beta_nonzero) < cpu_ns ? 1 : 0;
}

extern "C" int64_t zynum_gpu_offload_crossover(int scalar, int host_mask, int beta_nonzero, double cpu_gflops) {
    if (!zg::valid(scalar, 1, 1, 1, host_mask) || !(cpu_gflops > 0.0) || isinf(cpu_gflops))
        return -ZYNUM_GPU_DECLINE_ARGUMENT;
    // Smallest square edge that the GPU tier takes.  The GPU tier's time grows
    // more slowly with the edge than the CPU tier's n^3 once the GPU is
    // latency-bound, so the choice flips once; scan coarsely, then bisect.
    const int64_t kMaxEdge = int64_t(1) << 16;
    int64_t lo = 0, hi = -1;
    for (int64_t e = 1; e <= kMaxEdge; e *= 2) {
        if (zynum_gpu_offload_choose(scalar, e, e, e, host_mask, beta_nonzero, cpu_gflops) == 1) {
            hi = e;
            break;
        }
        lo = e;
    }
    if (hi < 0) return -1;  // the CPU tier is faster at every edge up to 65536
    while (hi - lo > 1) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (zynum_gpu_offload_choose(scalar, mid, mid, mid, host_mask, beta_nonzero, cpu_gflops) == 1)
            hi = mid;
        else
            lo = mid;
    }
    return hi;
}
