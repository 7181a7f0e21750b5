// Batched cycle over many independent 3D C2C grids (no counterpart in the reference).
#include "internal.h"

extern "C" {

int32_t dtfft_hip_cycle_create(int32_t, const int32_t*, int32_t, int32_t, void*, dtfft_hip_cycle_t* cycle) {
    if (cycle) *cycle = nullptr;
    return DTFFT_HIP_ERROR_INVALID_USAGE;
}
int32_t dtfft_hip_cycle_get_layout(dtfft_hip_cycle_t, int64_t*, size_t*) { return DTFFT_HIP_ERROR_PLAN_NOT_CREATED; }
int32_t dtfft_hip_cycle_execute(dtfft_hip_cycle_t, void*, void*, void*, int32_t) {
    return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
}
int32_t dtfft_hip_cycle_execute_graph(dtfft_hip_cycle_t, void*, void*, void*, void*) {
    return DTFFT_HIP_ERROR_PLAN_NOT_CREATED;
}
int32_t dtfft_hip_cycle_destroy(dtfft_hip_cycle_t*) { return DTFFT_HIP_ERROR_PLAN_NOT_CREATED; }

}  // extern "C"
