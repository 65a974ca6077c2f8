// Host-side launch helper shared by the waveform, likelihood and derivative
// kernels: per-binary set-up into a stream-ordered workspace, optional
// start-bin ordering, the caller's launch, then release of the workspace on
// the same stream.
#pragma once

#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "wave_common.hpp"

namespace gbgpu {

// Per-binary set-up into a stream-ordered workspace, then the waveform
// kernel, then the workspace is released on the same stream.
// Batches of at least this many binaries are processed in start-bin order.
constexpr long long SORT_MIN_BATCH = 8192;

template <class P, class Launch>
inline int with_pack(WaveArgs a, int N, hipStream_t st, Launch launch) {
    const size_t B = (size_t)a.B;
    const bool sort = a.B >= SORT_MIN_BATCH;
    size_t sort_bytes = 0;
    if (sort) {
        hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (unsigned *)nullptr,
                                                          (unsigned *)nullptr, (int *)nullptr, (int *)nullptr,
                                                          (int)a.B, 0, SORT_KEY_BITS, st);
        if (e != hipSuccess) return hip_fail(e, "SortPairs(size)");
    }
    auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
    const size_t bytes = al(sizeof(Pack<P>) * B) + (sort ? 2 * al(4 * B) + 2 * al(4 * B) + al(sort_bytes) : 0);
    char *ws = nullptr;
    hipError_t e = hipMallocAsync((void **)&ws, bytes, st);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(per-binary workspace)");
    size_t off = 0;
    auto carve = [&](size_t n) {
        char *p = ws + off;
        off += al(n);
        return p;
    };
    Pack<P> *pk = (Pack<P> *)carve(sizeof(Pack<P>) * B);
    unsigned *keys = nullptr, *keys_out = nullptr;
    int *vals = nullptr, *order = nullptr;
    void *tmp = nullptr;
    if (sort) {
        keys = (unsigned *)carve(4 * B);
        keys_out = (unsigned *)carve(4 * B);
        vals = (int *)carve(4 * B);
        order = (int *)carve(4 * B);
        tmp = carve(sort_bytes);
    }
    a.keys = keys;
    a.vals = vals;
    hipLaunchKernelGGL((k_setup<P>), dim3((unsigned)((a.B + 255) / 256)), dim3(256), 0, st, a, N, pk);
    int rc = GBGPU_OK;
    if (sort) {
        e = hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, keys, keys_out, vals, order, (int)a.B, 0,
                                               SORT_KEY_BITS, st);
        if (e != hipSuccess) rc = hip_fail(e, "SortPairs");
    }
    a.keys = nullptr;
    a.vals = nullptr;
    a.pack = pk;
    a.order = order;
    if (!rc) rc = launch(a);
    e = hipFreeAsync(ws, st);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipFreeAsync(per-binary workspace)");
    return GBGPU_OK;
}

}  // namespace gbgpu
