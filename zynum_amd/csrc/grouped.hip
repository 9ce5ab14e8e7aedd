// Grouped GEMM of many uneven problems (filled in after the plain path).
#include "common.hpp"
#include "runtime.hpp"

extern "C" {
int zynum_gpu_grouped_plan_create(int, const zynum_gpu_gemm_problem*, int, zynum_gpu_grouped_plan** out) {
    if (out) *out = nullptr;
    return ZYNUM_GPU_DECLINE_UNSUPPORTED;
}
int zynum_gpu_grouped_run(zynum_gpu_grouped_plan*, void*) { return ZYNUM_GPU_DECLINE_UNSUPPORTED; }
int zynum_gpu_grouped_graph_run(zynum_gpu_grouped_plan*, void*) { return ZYNUM_GPU_DECLINE_UNSUPPORTED; }
int zynum_gpu_grouped_plan_info(const zynum_gpu_grouped_plan*, int64_t*, int64_t*, double*) {
    return ZYNUM_GPU_DECLINE_UNSUPPORTED;
}
void zynum_gpu_grouped_plan_destroy(zynum_gpu_grouped_plan*) {}
int zynum_gpu_grouped_loop(int, const zynum_gpu_gemm_problem*, int, void*) { return ZYNUM_GPU_DECLINE_UNSUPPORTED; }
}
