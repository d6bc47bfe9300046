// sph_probe.hip -- a plain device copy of known size, used only to calibrate HBM counters
// (rocprofv3 FETCH_SIZE / WRITE_SIZE per dispatch against a known byte count). Not part of the step.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/spearhed_sph.h"

namespace {
__global__ __launch_bounds__(256) void k_probe_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}
}  // namespace

extern "C" int sph_probe_copy(int64_t bytes, int32_t reps, double* ms_per_copy)
{
    if (bytes < 16 || reps < 1) return SPH_E_INVALID;
    const int64_t n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return SPH_E_NOMEM;
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        hipFree(a);
        return SPH_E_NOMEM;
    }
    hipMemset(a, 0, n * 16);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const int grid = 256 * 32;
    k_probe_copy<<<grid, 256>>>(a, b, n);  // warm-up
    hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) k_probe_copy<<<grid, 256>>>(a, b, n);
    hipEventRecord(e1, nullptr);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_copy) *ms_per_copy = ms / reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return err == hipSuccess ? SPH_OK : SPH_E_DEVICE;
}
