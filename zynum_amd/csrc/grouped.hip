// Grouped GEMM of many independent, uneven problems in one persistent launch.
//
// The reference has no batched API; a caller loops over cblas_dgemm, each call
// re-entering the planner (planner.zig:88-211) and thread pool.  Here a plan
// validates the batch once (gemmError rules, fortran.zig:88-103), picks a tile
// configuration per problem from the tuned size-bucket table
// (grouped_tune.inc, written by zynum_amd/autotune.py), expands every problem
// into output tiles, orders the tiles longest-first (LPT: tiles with the most
// K work first) and uploads the list.  One persistent kernel (three workgroups
// per CU) pulls tiles from an atomic counter, so small and large problems
// share the machine without a launch per problem; every tile runs the same
// MFMA tile routine as the plain GEMM, so each problem's result is bitwise the
// result of a single plain GEMM call on it.  The launch pair (counter reset +
// persistent kernel) allocates nothing and can be captured into a HIP graph.
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "gemm_tile.hpp"
#include "runtime.hpp"

namespace zg {

template <typename T>
struct GroupedProblem {
    int m, n, k, ta, tb;
    int64_t lda, ldb, ldc;
    T alpha, beta;
    const T* a;
    const T* b;
    T* c;
};

struct WorkItem {
    int problem;
    int path;  // cfg * 4 + ta * 2 + tb, or kScalePath
    int m0, n0;
};

constexpr int kScalePath = 64;
constexpr int kGroupedCfgs = 2;  // 0: 32x32, 1: 64x64 (2x2 waves, 256 threads)
constexpr int kScaleTile = 64;

template <typename T>
struct GroupedArgs {
    const GroupedProblem<T>* problems;
    const WorkItem* items;
    int nitems;
    int* counter;
};

// K slices of 16 for both scalar types (the per-element MFMA order does not
// depend on the slice depth)
template <typename T>
using GCfg0 = TileCfg<32, 32, 16, 2, 2, 3>;
template <typename T>
using GCfg1 = TileCfg<64, 64, 16, 2, 2, 3>;

__global__ void reset_counter_kernel(int* counter) { *counter = 0; }

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void grouped_gemm_kernel(GroupedArgs<T> g) {
    __shared__ __attribute__((aligned(16))) T lds_a[tile_lds_a<T, GCfg1<T>>()];
    __shared__ __attribute__((aligned(16))) T lds_b[tile_lds_b<T, GCfg1<T>>()];
    __shared__ int s_item;
    for (;;) {
        __syncthreads();  // the previous tile is done with LDS and s_item
        if (threadIdx.x == 0) s_item = atomicAdd(g.counter, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= g.nitems) return;  // every wave of the workgroup sees the same item
        const WorkItem w = g.items[it];
        const GroupedProblem<T> pr = g.problems[w.problem];
        GemmArgs<T> p;
        p.m = pr.m; p.n = pr.n; p.k = pr.k;
        p.lda = pr.lda; p.ldb = pr.ldb; p.ldc = pr.ldc;
        p.alpha = pr.alpha; p.beta = pr.beta;
        p.a = pr.a; p.b = pr.b; p.c = pr.c;
        p.tiles_m = p.tiles_n = p.group_m = 1;
        switch (w.path) {
#define ZG_PATH(ID, CFG, TA, TB) \
    case ID: gemm_tile<T, CFG<T>, TA, TB, 1>(p, w.m0, w.n0, lds_a, lds_b); break;
            ZG_PATH(0, GCfg0, false, false) ZG_PATH(1, GCfg0, false, true)
            ZG_PATH(2, GCfg0, true, false) ZG_PATH(3, GCfg0, true, true)
            ZG_PATH(4, GCfg1, false, false) ZG_PATH(5, GCfg1, false, true)
            ZG_PATH(6, GCfg1, true, false) ZG_PATH(7, GCfg1, true, true)
#undef ZG_PATH
            default: {  // k == 0 or alpha == 0: C = beta * C on a 64 x 64 tile
                for (int e = threadIdx.x; e < kScaleTile * kScaleTile; e += 256) {
                    const int i = w.m0 + e % kScaleTile, j = w.n0 + e / kScaleTile;
                    if (i < p.m && j < p.n) {
                        T* cp = p.c + (int64_t)j * p.ldc + i;
                        *cp = (p.beta == T(0)) ? T(0) : p.beta * *cp;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// tile choice per size bucket
// ---------------------------------------------------------------------------
namespace {
struct GroupedTuneRow {
    int elem_bytes;
    int max_dim;  // largest of m and n in the bucket (inclusive)
    int cfg;
};
#include "grouped_tune.inc"
}  // namespace

int grouped_pick(int m, int n, int k, int elem_bytes) {
    (void)k;
    const int d = std::max(m, n);
    for (const GroupedTuneRow& r : kGroupedTuneTable)
        if (r.elem_bytes == elem_bytes && d <= r.max_dim) return r.cfg;
    return 1;
}

static const int kCfgTile[kGroupedCfgs] = {32, 64};

}  // namespace zg

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct zynum_gpu_grouped_plan {
    int scalar = 0;
    int nitems = 0;
    int nproblems = 0;
    int grid = 0;
    double flops = 0.0;
    void* problems = nullptr;  // device GroupedProblem<T>[]
    zg::WorkItem* items = nullptr;
    int* counter = nullptr;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

inline int32_t max1(int32_t x) { return x > 1 ? x : 1; }

bool problem_valid(const zynum_gpu_gemm_problem& p) {
    if (p.transa > ZYNUM_GPU_CONJ_TRANS || p.transb > ZYNUM_GPU_CONJ_TRANS) return false;
    if (p.m < 0 || p.n < 0 || p.k < 0) return false;
    const int32_t nrowa = p.transa == ZYNUM_GPU_NO_TRANS ? p.m : p.k;
    const int32_t nrowb = p.transb == ZYNUM_GPU_NO_TRANS ? p.k : p.n;
    return p.lda >= max1(nrowa) && p.ldb >= max1(nrowb) && p.ldc >= max1(p.m);
}

template <typename T>
int build_plan(const zynum_gpu_gemm_problem* probs, int count, int force_cfg, zynum_gpu_grouped_plan* plan) {
    std::vector<zg::GroupedProblem<T>> dp;
    std::vector<zg::WorkItem> items;
    std::vector<double> cost;
    dp.reserve(count);
    for (int i = 0; i < count; ++i) {
        const zynum_gpu_gemm_problem& q = probs[i];
        if (!problem_valid(q)) return ZYNUM_GPU_DECLINE_ARGUMENT;
        if (q.m == 0 || q.n == 0) continue;
        zg::GroupedProblem<T> g;
        g.m = q.m; g.n = q.n; g.k = q.k;
        g.ta = q.transa != ZYNUM_GPU_NO_TRANS; g.tb = q.transb != ZYNUM_GPU_NO_TRANS;
        g.lda = q.lda; g.ldb = q.ldb; g.ldc = q.ldc;
        g.alpha = (T)q.alpha; g.beta = (T)q.beta;
        g.a = static_cast<const T*>(q.a); g.b = static_cast<const T*>(q.b); g.c = static_cast<T*>(q.c);
        const int pi = (int)dp.size();
        dp.push_back(g);
        if (q.k == 0 || g.alpha == T(0)) {
            for (int n0 = 0; n0 < q.n; n0 += zg::kScaleTile)
                for (int m0 = 0; m0 < q.m; m0 += zg::kScaleTile) {
                    items.push_back({pi, zg::kScalePath, m0, n0});
                    cost.push_back(1.0 * zg::kScaleTile * zg::kScaleTile);
                }
            continue;
        }
        plan->flops += 2.0 * q.m * q.n * (double)q.k;
        const int cfg = force_cfg >= 0 ? force_cfg : zg::grouped_pick(q.m, q.n, q.k, (int)sizeof(T));
        const int t = zg::kCfgTile[cfg];
        const int path = cfg * 4 + g.ta * 2 + g.tb;
        for (int n0 = 0; n0 < q.n; n0 += t)
            for (int m0 = 0; m0 < q.m; m0 += t) {
                items.push_back({pi, path, m0, n0});
                cost.push_back((double)t * t * (double)q.k + 4096.0);
            }
    }
    // longest-processing-time order
    std::vector<int> order(items.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    std::vector<zg::WorkItem> sorted(items.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = items[order[i]];

    plan->nitems = (int)sorted.size();
    plan->nproblems = (int)dp.size();
    const size_t pb = std::max<size_t>(1, dp.size()) * sizeof(zg::GroupedProblem<T>);
    const size_t ib = std::max<size_t>(1, sorted.size()) * sizeof(zg::WorkItem);
    if (hipMalloc(&plan->problems, pb) != hipSuccess || hipMalloc((void**)&plan->items, ib) != hipSuccess ||
        hipMalloc((void**)&plan->counter, sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        return ZYNUM_GPU_DECLINE_ALLOC;
    }
    if (!dp.empty() && hipMemcpy(plan->problems, dp.data(), dp.size() * sizeof(dp[0]), hipMemcpyHostToDevice) != hipSuccess)
        return ZYNUM_GPU_ERR_LAUNCH;
    if (!sorted.empty() &&
        hipMemcpy(plan->items, sorted.data(), sorted.size() * sizeof(sorted[0]), hipMemcpyHostToDevice) != hipSuccess)
        return ZYNUM_GPU_ERR_LAUNCH;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    }
    (void)hipGetLastError();
    plan->grid = std::max(1, std::min(plan->nitems, 3 * cus));
    return ZYNUM_GPU_OK;
}

template <typename T>
hipError_t launch_plan(const zynum_gpu_grouped_plan* plan, hipStream_t s) {
    zg::reset_counter_kernel<<<1, 1, 0, s>>>(plan->counter);
    if (plan->nitems == 0) return hipGetLastError();
    zg::GroupedArgs<T> g;
    g.problems = static_cast<const zg::GroupedProblem<T>*>(plan->problems);
    g.items = plan->items;
    g.nitems = plan->nitems;
    g.counter = plan->counter;
    zg::grouped_gemm_kernel<T><<<plan->grid, 256, 0, s>>>(g);
    return hipGetLastError();
}

hipError_t launch_any(const zynum_gpu_grouped_plan* plan, hipStream_t s) {
    return plan->scalar == ZYNUM_GPU_F64 ? launch_plan<double>(plan, s) : launch_plan<float>(plan, s);
}

}  // namespace

extern "C" {

int zynum_gpu_grouped_plan_create(int scalar, const zynum_gpu_gemm_problem* problems, int count, int tile,
                                  zynum_gpu_grouped_plan** plan_out) {
    if (plan_out == nullptr) return ZYNUM_GPU_DECLINE_ARGUMENT;
    *plan_out = nullptr;
    if (count < 0 || (count > 0 && problems == nullptr)) return ZYNUM_GPU_DECLINE_ARGUMENT;
    if (scalar != ZYNUM_GPU_F64 && scalar != ZYNUM_GPU_F32) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    if (tile < 0 || tile > zg::kGroupedCfgs) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    if (!zg::device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    zynum_gpu_grouped_plan* plan = new zynum_gpu_grouped_plan;
    plan->scalar = scalar;
    const int st = scalar == ZYNUM_GPU_F64 ? build_plan<double>(problems, count, tile - 1, plan)
                                           : build_plan<float>(problems, count, tile - 1, plan);
    if (st != ZYNUM_GPU_OK) {
        zynum_gpu_grouped_plan_destroy(plan);
        return st;
    }
    *plan_out = plan;
    return ZYNUM_GPU_OK;
}

int zynum_gpu_grouped_run(zynum_gpu_grouped_plan* plan, void* stream) {
    if (plan == nullptr) return ZYNUM_GPU_DECLINE_ARGUMENT;
    return zg::status_from_hip(launch_any(plan, reinterpret_cast<hipStream_t>(stream)));
}

int zynum_gpu_grouped_graph_run(zynum_gpu_grouped_plan* plan, void* stream) {
    if (plan == nullptr) return ZYNUM_GPU_DECLINE_ARGUMENT;
    hipError_t e = hipSuccess;
    if (plan->exec == nullptr) {
        if (plan->capture_stream == nullptr) e = hipStreamCreateWithFlags(&plan->capture_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamBeginCapture(plan->capture_stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t le = launch_any(plan, plan->capture_stream);
            e = hipStreamEndCapture(plan->capture_stream, &plan->graph);
            if (e == hipSuccess) e = le;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) return zg::status_from_hip(e);
    }
    return zg::status_from_hip(hipGraphLaunch(plan->exec, reinterpret_cast<hipStream_t>(stream)));
}

int zynum_gpu_grouped_plan_info(const zynum_gpu_grouped_plan* plan, int64_t* tiles_out, int64_t* workgroups_out,
                                double* flops_out) {
    if (plan == nullptr) return ZYNUM_GPU_DECLINE_ARGUMENT;
    if (tiles_out) *tiles_out = plan->nitems;
    if (workgroups_out) *workgroups_out = plan->grid;
    if (flops_out) *flops_out = plan->flops;
    return ZYNUM_GPU_OK;
}

void zynum_gpu_grouped_plan_destroy(zynum_gpu_grouped_plan* plan) {
    if (plan == nullptr) return;
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    if (plan->capture_stream) (void)hipStreamDestroy(plan->capture_stream);
    if (plan->problems || plan->items || plan->counter) (void)hipDeviceSynchronize();
    if (plan->problems) (void)hipFree(plan->problems);
    if (plan->items) (void)hipFree(plan->items);
    if (plan->counter) (void)hipFree(plan->counter);
    (void)hipGetLastError();
    delete plan;
}

int zynum_gpu_grouped_loop(int scalar, const zynum_gpu_gemm_problem* problems, int count, void* stream) {
    if (count < 0 || (count > 0 && problems == nullptr)) return ZYNUM_GPU_DECLINE_ARGUMENT;
    if (scalar != ZYNUM_GPU_F64 && scalar != ZYNUM_GPU_F32) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    for (int i = 0; i < count; ++i)
        if (!problem_valid(problems[i])) return ZYNUM_GPU_DECLINE_ARGUMENT;
    if (!zg::device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    for (int i = 0; i < count; ++i) {
        const zynum_gpu_gemm_problem& q = problems[i];
        hipError_t e;
        if (scalar == ZYNUM_GPU_F64) {
            zg::GemmDesc<double> d;
            d.ta = q.transa != ZYNUM_GPU_NO_TRANS; d.tb = q.transb != ZYNUM_GPU_NO_TRANS;
            d.m = q.m; d.n = q.n; d.k = q.k; d.lda = q.lda; d.ldb = q.ldb; d.ldc = q.ldc;
            d.alpha = q.alpha; d.beta = q.beta;
            d.a = static_cast<const double*>(q.a); d.b = static_cast<const double*>(q.b);
            d.c = static_cast<double*>(q.c);
            e = zg::gemm_launch<double>(d, s);
        } else {
            zg::GemmDesc<float> d;
            d.ta = q.transa != ZYNUM_GPU_NO_TRANS; d.tb = q.transb != ZYNUM_GPU_NO_TRANS;
            d.m = q.m; d.n = q.n; d.k = q.k; d.lda = q.lda; d.ldb = q.ldb; d.ldc = q.ldc;
            d.alpha = (float)q.alpha; d.beta = (float)q.beta;
            d.a = static_cast<const float*>(q.a); d.b = static_cast<const float*>(q.b);
            d.c = static_cast<float*>(q.c);
            e = zg::gemm_launch<float>(d, s);
        }
        if (e != hipSuccess) return zg::status_from_hip(e);
    }
    return ZYNUM_GPU_OK;
}

}  // extern "C"
