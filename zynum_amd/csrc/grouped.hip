// Grouped GEMM of many independent, uneven problems in one persistent launch.
//
// The reference has no batched API; a caller loops over cblas_dgemm, each call
// re-entering the planner (planner.zig:88-211) and thread pool.  Here a plan
// validates the batch once (gemmError rules, fortran.zig:88-103), picks a tile
// configuration per problem from the tuned size-bucket table
// (grouped_tune.inc, written by zynum_amd/autotune.py), expands every problem
// into output tiles and orders them longest-first (LPT: tiles with the most K
// work first).  Tiles are 128x128, 64x64 or 32x64 / 64x32 by size class, and a
// remainder of at most half a tile is covered by the next smaller shape, so
// little MFMA work is spent on padding.  Each
// tile is uploaded as one self-contained descriptor (problem arguments, tile
// origin, code path).  One persistent kernel (two 8-wave workgroups per CU) walks
// the LPT list round-robin (workgroup w takes tiles w, w + G, w + 2G, ...),
// loading the descriptor of its next tile while the current one runs, so no
// atomic or dependent load sits between tiles.  Every tile runs the same MFMA
// tile routine as the plain GEMM, so each problem's result is bitwise the
// result of a single plain GEMM call on it.  The launch allocates nothing and
// can be captured into a HIP graph.
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "gemm_tile.hpp"
#include "runtime.hpp"

namespace zg {

constexpr int kScalePath = 64;
constexpr int kScaleTile = 64;
constexpr int kGroupedThreads = 512;  // 8 waves: every tile shape below uses all of them
// tuned configurations (size classes): 0 small (32x64 / 64x32 tiles), 1 64x64, 2 128x128
constexpr int kGroupedCfgs = 3;

// one output tile with everything its workgroup needs
template <typename T>
struct alignas(16) TileItem {
    GemmArgs<T> p;  // the problem (tiles_m / tiles_n / group_m unused)
    int m0, n0;
    int path;       // shape * 4 + ta * 2 + tb, or kScalePath (k == 0 or alpha == 0: C = beta C)
    int pad;
};

// Tile shapes (K slices of 16 for both scalar types; the per-element MFMA
// order does not depend on the slice depth).  Shape ids: 0: 32x64, 1: 64x32,
// 2: 64x64, 3: 128x128.
using GShape0 = TileCfg<32, 64, 16, 2, 4, 2>;
using GShape1 = TileCfg<64, 32, 16, 4, 2, 2>;
using GShape2 = TileCfg<64, 64, 16, 4, 2, 2>;
using GShape3 = TileCfg<128, 128, 16, 4, 2, 2>;
constexpr int kShapeRows[4] = {32, 64, 64, 128};
constexpr int kShapeCols[4] = {64, 32, 64, 128};

template <typename T>
__device__ __forceinline__ void run_item(const TileItem<T>& w, T* lds_a, T* lds_b) {
    switch (w.path) {
#define ZG_PATH(SH, CFG)                                                                              \
    case SH * 4 + 0: gemm_tile<T, CFG, false, false, 1>(w.p, w.m0, w.n0, lds_a, lds_b); break;       \
    case SH * 4 + 1: gemm_tile<T, CFG, false, true, 1>(w.p, w.m0, w.n0, lds_a, lds_b); break;        \
    case SH * 4 + 2: gemm_tile<T, CFG, true, false, 1>(w.p, w.m0, w.n0, lds_a, lds_b); break;        \
    case SH * 4 + 3: gemm_tile<T, CFG, true, true, 1>(w.p, w.m0, w.n0, lds_a, lds_b); break;
        ZG_PATH(0, GShape0) ZG_PATH(1, GShape1) ZG_PATH(2, GShape2) ZG_PATH(3, GShape3)
#undef ZG_PATH
        default: {  // k == 0 or alpha == 0: C = beta * C on a 64 x 64 tile
            for (int e = threadIdx.x; e < kScaleTile * kScaleTile; e += kGroupedThreads) {
                const int i = w.m0 + e % kScaleTile, j = w.n0 + e / kScaleTile;
                if (i < w.p.m && j < w.p.n) {
                    T* cp = w.p.c + (int64_t)j * w.p.ldc + i;
                    *cp = (w.p.beta == T(0)) ? T(0) : w.p.beta * *cp;
                }
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kGroupedThreads) __attribute__((amdgpu_waves_per_eu(2)))
void grouped_gemm_kernel(const TileItem<T>* __restrict__ items, int nitems) {
    __shared__ __attribute__((aligned(16))) T lds_a[tile_lds_a<T, GShape3>()];
    __shared__ __attribute__((aligned(16))) T lds_b[tile_lds_b<T, GShape3>()];
    int it = blockIdx.x;
    if (it >= nitems) return;
    // descriptors are uniform: scalar loads, the next one issued before this
    // tile's operand loads so it has landed by the tile's first LDS wait
    TileItem<T> cur = items[it];
    for (;;) {
        const int nx = it + gridDim.x;
        TileItem<T> nxt;
        if (nx < nitems) nxt = items[nx];
        run_item<T>(cur, lds_a, lds_b);
        if (nx >= nitems) return;
        // every wave is done reading this tile's LDS; its C stores may still be in flight
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __builtin_amdgcn_s_barrier();
        cur = nxt;
        it = nx;
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

static const int kCfgTile[kGroupedCfgs] = {32, 64, 128};

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
    void* items = nullptr;  // device TileItem<T>[nitems], LPT order
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

// Split [0, dim) into blocks of t, then (tuned plans) the remainder into
// blocks of t/2 and t/4 down to 32, rounding the last one up: the covered
// extent exceeds dim by less than 32 where whole t blocks could waste up to
// t - 1.
struct Seg {
    int start, size, tile;
};
static std::vector<Seg> split_dim(int dim, int t, bool mixed) {
    std::vector<Seg> out;
    if (!mixed || t == 32) {
        out.push_back({0, (dim + t - 1) / t * t, t});
        return out;
    }
    int pos = 0;
    for (int b = t; b >= 32 && pos < dim; b /= 2) {
        int cnt = (dim - pos) / b;
        const int rem = (dim - pos) - cnt * b;
        // round up at this size when the rest would not fit the smaller sizes with less waste
        if (b == 32 || rem > b - 32) cnt += rem > 0;
        if (cnt > 0) {
            out.push_back({pos, cnt * b, b});
            pos += cnt * b;
        }
    }
    return out;
}

// tile shape for blocks of rt rows x ct columns (rt, ct in {32, 64, 128})
static int shape_for(int rt, int ct) {
    if (rt == 128 && ct == 128) return 3;
    if (rt >= 64 && ct >= 64) return 2;
    if (rt >= 64) return 1;  // ct == 32
    return 0;                // rt == 32
}

template <typename T>
int build_plan(const zynum_gpu_gemm_problem* probs, int count, int force_cfg, zynum_gpu_grouped_plan* plan) {
    std::vector<zg::TileItem<T>> items;
    std::vector<double> cost;
    for (int i = 0; i < count; ++i) {
        const zynum_gpu_gemm_problem& q = probs[i];
        if (!problem_valid(q)) return ZYNUM_GPU_DECLINE_ARGUMENT;
        if (q.m == 0 || q.n == 0) continue;
        zg::TileItem<T> w;
        memset(&w, 0, sizeof(w));
        zg::GemmArgs<T>& g = w.p;
        g.m = q.m; g.n = q.n; g.k = q.k;
        g.lda = q.lda; g.ldb = q.ldb; g.ldc = q.ldc;
        g.alpha = (T)q.alpha; g.beta = (T)q.beta;
        g.a = static_cast<const T*>(q.a); g.b = static_cast<const T*>(q.b); g.c = static_cast<T*>(q.c);
        g.tiles_m = g.tiles_n = g.group_m = 1;
        const int ta = q.transa != ZYNUM_GPU_NO_TRANS, tb = q.transb != ZYNUM_GPU_NO_TRANS;
        auto add = [&](int path, int m0, int n0, double c) {
            w.path = path; w.m0 = m0; w.n0 = n0;
            items.push_back(w);
            cost.push_back(c);
        };
        if (q.k == 0 || g.alpha == T(0)) {
            for (int n0 = 0; n0 < q.n; n0 += zg::kScaleTile)
                for (int m0 = 0; m0 < q.m; m0 += zg::kScaleTile)
                    add(zg::kScalePath, m0, n0, 1.0 * zg::kScaleTile * zg::kScaleTile);
            continue;
        }
        plan->flops += 2.0 * q.m * q.n * (double)q.k;
        const int cfg = force_cfg >= 0 ? force_cfg : zg::grouped_pick(q.m, q.n, q.k, (int)sizeof(T));
        const int t = zg::kCfgTile[cfg];
        // the tuned choice covers a remainder of at most t/2 with half-size
        // tiles; a forced configuration uses its size throughout
        const bool mixed = force_cfg < 0;
        // small class: 32-row blocks along the shorter side, 64 along the other
        const int rt0 = (cfg == 0) ? (q.m <= q.n ? 32 : 64) : t;
        const int ct0 = (cfg == 0) ? (q.m <= q.n ? 64 : 32) : t;
        const std::vector<Seg> rs = split_dim(q.m, rt0, mixed && cfg > 0);
        const std::vector<Seg> cs = split_dim(q.n, ct0, mixed && cfg > 0);
        for (const Seg& c : cs)
            for (const Seg& r : rs) {
                const int sh = shape_for(r.tile, c.tile);
                const int th = zg::kShapeRows[sh], tw = zg::kShapeCols[sh];
                for (int n0 = c.start; n0 < c.start + c.size; n0 += tw)
                    for (int m0 = r.start; m0 < r.start + r.size; m0 += th)
                        add(sh * 4 + ta * 2 + tb, m0, n0, (double)th * tw * (double)q.k + 8192.0);
            }
    }
    // longest-processing-time order
    std::vector<int> order(items.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    std::vector<zg::TileItem<T>> sorted(items.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = items[order[i]];

    plan->nitems = (int)sorted.size();
    plan->nproblems = count;
    const size_t ib = std::max<size_t>(1, sorted.size()) * sizeof(zg::TileItem<T>);
    if (hipMalloc(&plan->items, ib) != hipSuccess) {
        (void)hipGetLastError();
        return ZYNUM_GPU_DECLINE_ALLOC;
    }
    if (!sorted.empty() &&
        hipMemcpy(plan->items, sorted.data(), sorted.size() * sizeof(sorted[0]), hipMemcpyHostToDevice) != hipSuccess)
        return ZYNUM_GPU_ERR_LAUNCH;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    }
    (void)hipGetLastError();
    // as many persistent workgroups as are resident at once (register and LDS
    // limits of this kernel), so the round-robin share of every one is running
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(zg::grouped_gemm_kernel<T>),
                                                     zg::kGroupedThreads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    (void)hipGetLastError();
    plan->grid = std::max(1, std::min(plan->nitems, per_cu * cus));
    return ZYNUM_GPU_OK;
}

template <typename T>
hipError_t launch_plan(const zynum_gpu_grouped_plan* plan, hipStream_t s) {
    if (plan->nitems == 0) return hipSuccess;
    zg::grouped_gemm_kernel<T><<<plan->grid, zg::kGroupedThreads, 0, s>>>(static_cast<const zg::TileItem<T>*>(plan->items),
                                                          plan->nitems);
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
    if (plan->items) {
        (void)hipDeviceSynchronize();
        (void)hipFree(plan->items);
    }
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
