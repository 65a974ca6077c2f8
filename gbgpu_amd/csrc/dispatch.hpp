// Host-side launch helper shared by the waveform, likelihood and derivative
// kernels: per-binary set-up into a workspace, optional start-bin ordering,
// the caller's launch, then release of the workspace on the same stream.
#pragma once

#include "common.hpp"
#include "sort.hpp"
#include "wave_common.hpp"

namespace gbgpu {

// Batches of at least this many binaries are processed in start-bin order.
constexpr long long SORT_MIN_BATCH = 8192;

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Bytes of the per-binary workspace of with_pack for B binaries: the Pack
// array and, for sorted batches, keys, values and the radix sort's temporary
// storage (for the sort with_pack will run).  The size query of the sort
// allocates nothing.
template <class P>
inline int pack_workspace_bytes(long long B, hipStream_t st, size_t *total, size_t *sort_bytes,
                                bool allow_onesweep) {
    *sort_bytes = 0;
    if (B >= SORT_MIN_BATCH) {
        hipError_t e = sort_pairs<unsigned>(allow_onesweep, nullptr, *sort_bytes, nullptr, nullptr, nullptr, nullptr,
                                            (int)B, SORT_KEY_BITS, st);
        if (e != hipSuccess) return hip_fail(e, "SortPairs(size)");
    }
    *total = align256(sizeof(Pack<P>) * (size_t)B) +
             (B >= SORT_MIN_BATCH ? 4 * align256(4 * (size_t)B) + align256(*sort_bytes) : 0);
    return GBGPU_OK;
}

// Per-binary set-up (k_setup) into a workspace, the start-bin sort for large
// batches, then `launch(args)`.  With ext == nullptr the workspace is
// stream-ordered scratch (hipMallocAsync, freed with hipFreeAsync after the
// launch) and large batches may use the onesweep radix sort; otherwise the
// workspace is carved from the caller's ext_bytes at ext, nothing is
// allocated and the sort is the merge sort, which issues no memset, so the
// call can be captured in a graph.
template <class P, class Launch>
inline int with_pack(WaveArgs a, int N, hipStream_t st, Launch launch, char *ext = nullptr, size_t ext_bytes = 0) {
    const size_t B = (size_t)a.B;
    const bool sort = a.B >= SORT_MIN_BATCH;
    const bool allow_onesweep = ext == nullptr;
    size_t bytes = 0, sort_bytes = 0;
    int rc = pack_workspace_bytes<P>(a.B, st, &bytes, &sort_bytes, allow_onesweep);
    if (rc) return rc;
    char *ws = ext;
    if (ext) {
        if (bytes > ext_bytes) {
            set_error("per-binary workspace exceeds the caller's buffer");
            return GBGPU_E_ARG;
        }
    } else {
        hipError_t e = hipMallocAsync((void **)&ws, bytes, st);
        if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(per-binary workspace)");
    }
    size_t off = 0;
    auto carve = [&](size_t n) {
        char *p = ws + off;
        off += align256(n);
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
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail(e, "k_setup");
    if (!rc && sort) {
        e = sort_pairs<unsigned>(allow_onesweep, tmp, sort_bytes, keys, keys_out, vals, order, (int)a.B, SORT_KEY_BITS,
                                 st);
        if (e != hipSuccess) rc = hip_fail(e, "SortPairs");
    }
    a.keys = nullptr;
    a.vals = nullptr;
    a.pack = pk;
    a.order = order;
    if (!rc) rc = launch(a);
    if (ext) return rc;
    e = hipFreeAsync(ws, st);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipFreeAsync(per-binary workspace)");
    return GBGPU_OK;
}

}  // namespace gbgpu
