// Tile selection for the plain GEMM from the measured tuning table.  Replaces
// the CPU planner's blocking heuristics (planner.zig:88-211, tuning.zig:68-155).
#include "gemm.hpp"

namespace zg {

namespace {
struct TuneRow {
    int elem_bytes;
    int pairs;  // transpose pairs the row covers: bit 0 NN, 1 NT, 2 TN, 3 TT
    long long min_tiles128;
    int tile;
};
#include "tune_table.inc"
}  // namespace

int pick_tile(int m, int n, int k, int elem_bytes, bool ta, bool tb) {
    (void)k;
    const long long tiles128 = (long long)((m + 127) / 128) * ((n + 127) / 128);
    const int pair = 1 << ((ta ? 2 : 0) + (tb ? 1 : 0));
    for (const TuneRow& r : kPlainTuneTable) {
        if (r.elem_bytes == elem_bytes && (r.pairs & pair) && tiles128 >= r.min_tiles128) return r.tile;
    }
    return kTile64x64;
}

}  // namespace zg
