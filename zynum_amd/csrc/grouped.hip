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
// origin, code path).  One persistent kernel per tile path walks that path's
// LPT list round-robin (workgroup w takes tiles w, w + G, w + 2G, ...),
// loading the descriptor and first K slice of its next tile while the current
// one finishes, so no atomic or dependent load sits between tiles.  Every tile runs the same MFMA
// tile routine as the plain GEMM, so each problem's result is bitwise the
// result of a single plain GEMM launch on it that walks K in one tile (the
// single call splits K only for shapes with few output tiles and K >= 1024,
// and then agrees within the error bound).  The launch allocates nothing and
// can be captured into a HIP graph.
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "gemm_tile.hpp"
#include "runtime.hpp"

namespace zg {

constexpr int kScalePath = 64;
constexpr int kScaleTile = 64;
// tuned configurations (size classes): 0 small (32x64 / 64x32 tiles), 1 64x64, 2 128x128
constexpr int kGroupedCfgs = 3;

// one output tile with everything its workgroup needs
template <typename T>
struct alignas(16) TileItem {
    GemmArgs<T> p;  // the problem (tiles_m / tiles_n / group_m unused)
    int m0, n0;
    int path;       // shape * 8 + vec * 4 + ta * 2 + tb, or kScalePath (k == 0 or alpha == 0: C = beta C)
    int pad;
};

// Tile shapes (K slices of 16 for both scalar types; the per-element MFMA
// order does not depend on the slice depth).  Shape ids: 0: 32x64, 1: 64x32,
// 2: 64x64, 3: 128x128.
// The small shapes run four waves in a 2x2 grid, so each wave issues a whole
// (BM/2) x (BN/2) block of MFMAs per K slice: enough work per barrier to cover
// the next slice's load latency with the short K of grouped problems.  The
// 128x128 shape runs sixteen waves (4x4 of 32x32), the large plain GEMM's tile.
// OCC: waves per SIMD the registers are sized for.
using GShape0 = TileCfg<32, 64, 16, 2, 2, 4>;
using GShape1 = TileCfg<64, 32, 16, 2, 2, 4>;
using GShape2 = TileCfg<64, 64, 16, 2, 2, 4>;
using GShape3 = TileCfg<128, 128, 16, 4, 4, 4>;
constexpr int kShapeRows[4] = {32, 64, 64, 128};
constexpr int kShapeCols[4] = {64, 32, 64, 128};

// One persistent kernel per tile path (shape and transposes), so each runs
// with only its own registers and LDS.  A workgroup walks its share of the
// path's LPT list round-robin.  Its K slices form one stream across tile
// boundaries, loaded two slices ahead of the one being multiplied (two
// register sets, alternating with the two LDS buffers): a slice's loads have
// two slices of MFMA work to land, and a tile's first slices are already in
// flight while the previous tile finishes and stores its results.  The
// per-element arithmetic (16x16x4 MFMA steps in increasing k, then the
// alpha/beta epilogue) is the plain GEMM tile's, so results are bitwise those
// of single GEMM calls.
template <typename T, class CF, bool TA, bool TB, int VEC>
__global__ __launch_bounds__(CF::NT) __attribute__((amdgpu_waves_per_eu(CF::OCC))) void grouped_path_kernel(const TileItem<T>* __restrict__ items, int nitems) {
    constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, NT = CF::NT;
    constexpr int PAD = lds_pad<T>();
    constexpr int WTM = BM / CF::WM, WTN = BN / CF::WN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    static_assert(FM >= 1 && FN >= 1, "wave tile smaller than one MFMA block");
    using M = Mfma<T>;
    using Acc = typename M::acc_t;
    __shared__ __attribute__((aligned(16))) T lds_a[tile_lds_a<T, CF>()];
    __shared__ __attribute__((aligned(16))) T lds_b[tile_lds_b<T, CF>()];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % CF::WM) * WTM, wn0 = (wave / CF::WM) * WTN;
    const int lr = lane & 15, lk = lane >> 4;
    const int ar = M::a_row(lr);
    const int G = gridDim.x;

    int cit = blockIdx.x;  // tile being multiplied
    if (cit >= nitems) return;
    // vector width per operand: at most the elements each thread stages
    constexpr int VA = (VEC < BM * BK / NT) ? VEC : BM * BK / NT;
    constexpr int VB = (VEC < BN * BK / NT) ? VEC : BN * BK / NT;
    Stage<T, BM, BK, TA, VA, NT, PAD> sa[2];
    Stage<T, BN, BK, !TB, VB, NT, PAD> sb[2];
    using Set0 = std::integral_constant<int, 0>;
    using Set1 = std::integral_constant<int, 1>;

    // load cursor: the next slice to load is K offset lk0 of tile lit
    // (descriptor lw); lmore is false once the stream is exhausted.
    // Descriptors are uniform (scalar loads).
    int lit = cit, lk0 = 0;
    TileItem<T> lw = items[lit];
    bool lmore = true;
    auto load_next = [&](auto set_c) {
        constexpr int S = decltype(set_c)::value;
        const GemmArgs<T>& p = lw.p;
        const bool full = (lw.m0 + BM <= p.m) && (lw.n0 + BN <= p.n) && (lk0 + BK <= p.k);
        sa[S].load(p.a, p.lda, lw.m0, lk0, p.m, p.k, full, tid);
        sb[S].load(p.b, p.ldb, lw.n0, lk0, p.n, p.k, full, tid);
        lk0 += BK;
        if (lk0 >= p.k) {
            lit += G;
            lk0 = 0;
            lmore = lit < nitems;
            if (lmore) lw = items[lit];
        }
    };
    auto barrier = [&]() {  // LDS reads and writes of this step done; global loads may stay in flight
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");  // no LDS access moves above the barrier
    };

    // compute cursor: slice ct of the cnk slices of tile cit (descriptor cw)
    TileItem<T> cw = lw;
    int ct = 0, cnk = (cw.p.k + BK - 1) / BK;
    Acc acc[FM][FN];
    auto zero = [&]() {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = M::zero();
    };
    auto epilogue_store = [&]() {
        // lane holds rows 4*(lane/16)+r, column lane%16 of each block
        const GemmArgs<T>& p = cw.p;
        const T alpha = p.alpha, beta = p.beta;
#pragma unroll
        for (int i = 0; i < FM; ++i) {
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int col = cw.n0 + wn0 + j * 16 + lr;
                const int row = cw.m0 + wm0 + i * 16 + 4 * lk;
                if (col < p.n) {
                    const gptr<T> cp = to_global(p.c) + (int64_t)col * p.ldc + row;
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (row + r < p.m) cp[r] = epilogue_g<T>(alpha, beta, acc[i][j][r], cp + r);
                }
            }
        }
    };
    auto compute = [&](int cb) {
        const T* __restrict__ la = lds_a + cb * (BK * (BM + PAD));
        const T* __restrict__ lb = lds_b + cb * (BK * (BN + PAD));
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = la[(kk + lk) * (BM + PAD) + wm0 + i * 16 + ar];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = lb[(kk + lk) * (BN + PAD) + wn0 + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = M::mma(af[i], bf[j], acc[i][j]);
        }
    };

    // prologue: stream positions 0 and 1 into sets 0 and 1, position 0 into LDS buffer 0
    load_next(Set0{});
    bool next_loaded = lmore;  // position 1 exists
    if (next_loaded) load_next(Set1{});
    sa[0].store(lds_a, tid);
    sb[0].store(lds_b, tid);
    barrier();
    zero();
    // One stream position with its operands in register set / LDS buffer P;
    // returns false when the stream is done.
    auto step = [&](auto par_c) -> bool {
        constexpr int P = decltype(par_c)::value;
        using Own = std::integral_constant<int, P>;
        const bool loaded2 = lmore;  // position q + 2 exists
        if (loaded2) load_next(Own{});  // set P's operands are in LDS buffer P already
        compute(P);
        if (next_loaded) {
            sa[1 - P].store(lds_a + (1 - P) * (BK * (BM + PAD)), tid);
            sb[1 - P].store(lds_b + (1 - P) * (BK * (BN + PAD)), tid);
        }
        barrier();
        next_loaded = loaded2;
        if (++ct == cnk) {
            epilogue_store();
            cit += G;
            if (cit >= nitems) return false;
            cw = items[cit];
            ct = 0;
            cnk = (cw.p.k + BK - 1) / BK;
            zero();
        }
        return true;
    };
    for (;;) {
        if (!step(Set0{})) break;
        if (!step(Set1{})) break;
    }
}

// k == 0 or alpha == 0: C = beta * C (exact 0 for beta == 0) on 64 x 64 tiles
template <typename T>
__global__ __launch_bounds__(256) void grouped_scale_kernel(const TileItem<T>* __restrict__ items, int nitems) {
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const TileItem<T>& w = items[it];
        for (int e = threadIdx.x; e < kScaleTile * kScaleTile; e += 256) {
            const int i = w.m0 + e % kScaleTile, j = w.n0 + e / kScaleTile;
            if (i < w.p.m && j < w.p.n) {
                const gptr<T> cp = to_global(w.p.c) + (int64_t)j * w.p.ldc + i;
                *cp = (w.p.beta == T(0)) ? T(0) : w.p.beta * *cp;
            }
        }
    }
}

// path id -> kernel: shape * 8 + vec * 4 + ta * 2 + tb (vec: 16-byte operand loads)
template <typename T>
const void* path_kernel(int path) {
    constexpr int V = 16 / sizeof(T);
    switch (path) {
#define ZG_PATH4(SH, CFG, VB, VV)                                                                                   \
    case SH * 8 + VB * 4 + 0: return reinterpret_cast<const void*>(grouped_path_kernel<T, CFG, false, false, VV>); \
    case SH * 8 + VB * 4 + 1: return reinterpret_cast<const void*>(grouped_path_kernel<T, CFG, false, true, VV>);  \
    case SH * 8 + VB * 4 + 2: return reinterpret_cast<const void*>(grouped_path_kernel<T, CFG, true, false, VV>);  \
    case SH * 8 + VB * 4 + 3: return reinterpret_cast<const void*>(grouped_path_kernel<T, CFG, true, true, VV>);
#define ZG_PATH(SH, CFG) ZG_PATH4(SH, CFG, 0, 1) ZG_PATH4(SH, CFG, 1, V)
        ZG_PATH(0, GShape0) ZG_PATH(1, GShape1) ZG_PATH(2, GShape2) ZG_PATH(3, GShape3)
#undef ZG_PATH
#undef ZG_PATH4
        default: return reinterpret_cast<const void*>(grouped_scale_kernel<T>);
    }
}
constexpr int kPathThreads[4] = {GShape0::NT, GShape1::NT, GShape2::NT, GShape3::NT};

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
    void* items = nullptr;  // device TileItem<T>[nitems], grouped by path, LPT order within a path
    struct Launch {
        int path, first, count, grid;
    };
    std::vector<Launch> launches;  // one persistent kernel per non-empty path
    // the launches run on up to kBranches streams forked from the caller's
    // stream and joined back, so one kernel's tail overlaps the next one
    static constexpr int kBranches = 4;
    hipStream_t branch[kBranches] = {};
    hipEvent_t fork = nullptr, join[kBranches] = {};
    int nbranch = 0;
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
// Smallest block a remainder is split into (a build knob: 32 measured faster
// than 64 on the uneven 16..512 batch).
#ifndef ZG_GROUPED_MIN_SPLIT
#define ZG_GROUPED_MIN_SPLIT 32
#endif
constexpr int kMinSplit = ZG_GROUPED_MIN_SPLIT;
static std::vector<Seg> split_dim(int dim, int t, bool mixed) {
    std::vector<Seg> out;
    if (!mixed || t <= kMinSplit) {
        out.push_back({0, (dim + t - 1) / t * t, t});
        return out;
    }
    int pos = 0;
    for (int b = t; b >= kMinSplit && pos < dim; b /= 2) {
        int cnt = (dim - pos) / b;
        const int rem = (dim - pos) - cnt * b;
        // round up at this size when the rest would not fit the smaller sizes with less waste
        if (b == kMinSplit || rem > b - kMinSplit) cnt += rem > 0;
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
    std::vector<int> prob;  // problem index of every tile
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
        // 16-byte operand loads when both operands and leading dimensions allow them
        constexpr int V = 16 / sizeof(T);
        const int vec = ((reinterpret_cast<uintptr_t>(q.a) & 15) == 0 && (reinterpret_cast<uintptr_t>(q.b) & 15) == 0 &&
                         q.lda % V == 0 && q.ldb % V == 0) ? 1 : 0;
        auto add = [&](int path, int m0, int n0, double c) {
            w.path = path; w.m0 = m0; w.n0 = n0;
            items.push_back(w);
            cost.push_back(c);
            prob.push_back(i);
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
                        add(sh * 8 + vec * 4 + ta * 2 + tb, m0, n0, (double)th * tw * (double)q.k + 8192.0);
            }
    }
    // group by path (one persistent kernel each, largest tiles first), each
    // group in longest-processing-time order
    auto group = [](int path) { return path == zg::kScalePath ? -1 : path / 8; };
    std::vector<int> order(items.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const int gx = group(items[x].path), gy = group(items[y].path);
        if (gx != gy) return gx > gy;
        if (items[x].path != items[y].path) return items[x].path < items[y].path;
        return cost[x] > cost[y];
    });
    std::vector<zg::TileItem<T>> grouped_items(items.size());
    std::vector<double> grouped_cost(items.size());
    std::vector<int> grouped_prob(items.size());
    for (size_t i = 0; i < order.size(); ++i) {
        grouped_items[i] = items[order[i]];
        grouped_cost[i] = cost[order[i]];
        grouped_prob[i] = prob[order[i]];
    }
    // XCD-aware placement within each path: workgroup w of a launch runs on
    // XCD w % 8 and takes positions w, w + G, ... (G a multiple of 8), so
    // position j runs on XCD j % 8.  Problems are assigned to XCDs greedily
    // by cost (largest first) and each XCD's tiles fill its positions in LPT
    // order, so all tiles of a problem share one L2; once an XCD's tiles run
    // out its positions take the remaining tiles of the others.
    std::vector<zg::TileItem<T>> sorted;
    sorted.reserve(items.size());
    plan->launches.clear();
    for (size_t i = 0; i < grouped_items.size();) {
        size_t j = i;
        while (j < grouped_items.size() && grouped_items[j].path == grouped_items[i].path) ++j;
        // problem cost within this path
        std::vector<std::pair<double, int>> pc;  // (cost, problem)
        {
            std::vector<int> ids;
            for (size_t x = i; x < j; ++x) ids.push_back(grouped_prob[x]);
            std::sort(ids.begin(), ids.end());
            ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
            std::vector<double> tot(ids.size(), 0.0);
            for (size_t x = i; x < j; ++x)
                tot[std::lower_bound(ids.begin(), ids.end(), grouped_prob[x]) - ids.begin()] += grouped_cost[x];
            for (size_t q = 0; q < ids.size(); ++q) pc.push_back({tot[q], ids[q]});
            std::stable_sort(pc.begin(), pc.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        }
        std::vector<double> load(zg::kXcds, 0.0);
        std::vector<std::pair<int, int>> assign;  // (problem, xcd)
        for (const auto& e : pc) {
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[x] += e.first;
            assign.push_back({e.second, x});
        }
        std::sort(assign.begin(), assign.end());
        std::vector<std::vector<size_t>> queue(zg::kXcds);
        for (size_t x = i; x < j; ++x) {  // LPT order is kept within every queue
            const auto a = std::lower_bound(assign.begin(), assign.end(), std::make_pair(grouped_prob[x], -1));
            queue[a->second].push_back(x);
        }
        std::vector<size_t> head(zg::kXcds, 0);
        const size_t first = sorted.size();
        for (size_t pos = 0; pos < j - i; ++pos) {
            int q = (int)(pos % zg::kXcds);
            if (head[q] >= queue[q].size()) {  // this XCD is done: take the longest remaining tile
                int best = -1;
                for (int r = 0; r < zg::kXcds; ++r)
                    if (head[r] < queue[r].size() &&
                        (best < 0 || grouped_cost[queue[r][head[r]]] > grouped_cost[queue[best][head[best]]]))
                        best = r;
                q = best;
            }
            sorted.push_back(grouped_items[queue[q][head[q]++]]);
        }
        plan->launches.push_back({grouped_items[i].path, (int)first, (int)(j - i), 1});
        i = j;
    }

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
    // per path: as many persistent workgroups as are resident at once
    // (register and LDS limits of that kernel), so every round-robin share runs
    // (twice as many for the 16-wave paths, below)
    plan->grid = 0;
    for (auto& l : plan->launches) {
        const int threads = l.path == zg::kScalePath ? 256 : zg::kPathThreads[l.path / 8];
        int per_cu = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, zg::path_kernel<T>(l.path), threads, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        (void)hipGetLastError();
        // A 16-wave path (one workgroup fills a CU) launches twice the
        // resident count: each workgroup's round-robin share halves, and a CU
        // that finishes its share early starts one of the waiting workgroups
        // (measured on the 4096-problem batch: 3 % in FP64, 1 % in FP32;
        // doubling the other paths measured no gain).
        if (threads >= 1024) per_cu *= 2;
        l.grid = std::max(1, std::min(l.count, per_cu * cus));
        plan->grid = std::max(plan->grid, l.grid);
    }
    return ZYNUM_GPU_OK;
}

template <typename T>
hipError_t launch_plan(const zynum_gpu_grouped_plan* plan, hipStream_t caller) {
    constexpr int V = 16 / sizeof(T);
    const zg::TileItem<T>* base = static_cast<const zg::TileItem<T>*>(plan->items);
    const int nb = plan->nbranch;
    if (nb > 1) {
        hipError_t e = hipEventRecord(plan->fork, caller);
        for (int b = 0; b < nb && e == hipSuccess; ++b) e = hipStreamWaitEvent(plan->branch[b], plan->fork, 0);
        if (e != hipSuccess) return e;
    }
    for (size_t li = 0; li < plan->launches.size(); ++li) {
        const auto& l = plan->launches[li];
        hipStream_t s = nb > 1 ? plan->branch[li % nb] : caller;
        const zg::TileItem<T>* it = base + l.first;
        switch (l.path) {
#define ZG_LAUNCH4(SH, CFG, VB, VV)                                                                       \
    case SH * 8 + VB * 4 + 0:                                                                             \
        zg::grouped_path_kernel<T, CFG, false, false, VV><<<l.grid, CFG::NT, 0, s>>>(it, l.count); break; \
    case SH * 8 + VB * 4 + 1:                                                                             \
        zg::grouped_path_kernel<T, CFG, false, true, VV><<<l.grid, CFG::NT, 0, s>>>(it, l.count); break;  \
    case SH * 8 + VB * 4 + 2:                                                                             \
        zg::grouped_path_kernel<T, CFG, true, false, VV><<<l.grid, CFG::NT, 0, s>>>(it, l.count); break;  \
    case SH * 8 + VB * 4 + 3:                                                                             \
        zg::grouped_path_kernel<T, CFG, true, true, VV><<<l.grid, CFG::NT, 0, s>>>(it, l.count); break;
#define ZG_LAUNCH(SH, CFG) ZG_LAUNCH4(SH, CFG, 0, 1) ZG_LAUNCH4(SH, CFG, 1, V)
            ZG_LAUNCH(0, zg::GShape0) ZG_LAUNCH(1, zg::GShape1) ZG_LAUNCH(2, zg::GShape2) ZG_LAUNCH(3, zg::GShape3)
#undef ZG_LAUNCH
#undef ZG_LAUNCH4
            default: zg::grouped_scale_kernel<T><<<l.grid, 256, 0, s>>>(it, l.count);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    for (int b = 0; b < nb && nb > 1; ++b) {
        hipError_t e = hipEventRecord(plan->join[b], plan->branch[b]);
        if (e == hipSuccess) e = hipStreamWaitEvent(caller, plan->join[b], 0);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
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
    int st = scalar == ZYNUM_GPU_F64 ? build_plan<double>(problems, count, tile - 1, plan)
                                     : build_plan<float>(problems, count, tile - 1, plan);
    if (st == ZYNUM_GPU_OK) {
        plan->nbranch = std::min<int>(zynum_gpu_grouped_plan::kBranches, (int)plan->launches.size());
        if (plan->nbranch > 1) {
            hipError_t e = hipEventCreateWithFlags(&plan->fork, hipEventDisableTiming);
            for (int b = 0; b < plan->nbranch && e == hipSuccess; ++b) {
                e = hipStreamCreateWithFlags(&plan->branch[b], hipStreamNonBlocking);
                if (e == hipSuccess) e = hipEventCreateWithFlags(&plan->join[b], hipEventDisableTiming);
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                st = ZYNUM_GPU_DECLINE_ALLOC;
            }
        }
    }
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
    for (int b = 0; b < zynum_gpu_grouped_plan::kBranches; ++b) {
        if (plan->branch[b]) (void)hipStreamDestroy(plan->branch[b]);
        if (plan->join[b]) (void)hipEventDestroy(plan->join[b]);
    }
    if (plan->fork) (void)hipEventDestroy(plan->fork);
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
