// Measurement hook: the FMA rate this GPU sustains, for the rooflines the
// benchmark reports beside the published peaks.  Each thread runs CH
// independent FMA chains (x = a x + b), so the issue rate and not the FMA
// latency bounds the kernel; 8 chains per thread, 256-thread workgroups, 32
// workgroups per CU.  Each thread writes its sum to one element of a
// grid-sized output, which keeps the chains live.
#include <vector>

#include "common.hpp"

namespace gbgpu {

template <class T, int CH>
__global__ void __launch_bounds__(256) k_fma_probe(T *__restrict__ out, T a, T b, int iters) {
    T x[CH];
#pragma unroll
    for (int c = 0; c < CH; c++) x[c] = (T)(threadIdx.x + c);
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < CH; c++) x[c] = x[c] * a + b;
    }
    T s = 0;
#pragma unroll
    for (int c = 0; c < CH; c++) s += x[c];
    out[(long long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <class T>
static int fma_probe(double *tflops, hipStream_t st) {
    constexpr int CH = 8, WG = 256, ITERS = 4096, REPS = 10;
    int dev = 0, ncu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return hip_fail(e, "gbgpu_fma_probe(device)");
    const int grid = 32 * (ncu > 0 ? ncu : 256);
    T *out = nullptr;
    e = hipMallocAsync((void **)&out, sizeof(T) * (size_t)grid * WG, st);
    if (e != hipSuccess) return hip_fail(e, "gbgpu_fma_probe(alloc)");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0.0f;
    if (e == hipSuccess) {
        for (int w = 0; w < 3; w++)
            hipLaunchKernelGGL((k_fma_probe<T, CH>), dim3(grid), dim3(WG), 0, st, out, (T)0.999, (T)0.001, ITERS);
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < REPS; r++)
            hipLaunchKernelGGL((k_fma_probe<T, CH>), dim3(grid), dim3(WG), 0, st, out, (T)0.999, (T)0.001, ITERS);
        (void)hipEventRecord(e1, st);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    const hipError_t ef = hipFreeAsync(out, st);
    if (e != hipSuccess) return hip_fail(e, "gbgpu_fma_probe");
    if (ef != hipSuccess) return hip_fail(ef, "gbgpu_fma_probe(free)");
    *tflops = 2.0 * CH * (double)ITERS * grid * WG * REPS / (ms * 1e-3) / 1e12;
    return GBGPU_OK;
}

}  // namespace gbgpu

using namespace gbgpu;

extern "C" int gbgpu_fma_probe(int precision, double *tflops, void *stream) {
    if (!tflops || (precision != GBGPU_FP64 && precision != GBGPU_MIXED)) {
        set_error("gbgpu_fma_probe: bad arguments");
        return GBGPU_E_ARG;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    return precision == GBGPU_FP64 ? fma_probe<double>(tflops, st) : fma_probe<float>(tflops, st);
}
