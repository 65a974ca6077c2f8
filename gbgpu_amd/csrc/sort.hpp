// Stable key/value radix sorts of the start-bin orderings (rocPRIM).
//
// rocPRIM's default configuration merge-sorts up to 2^20 keys (21 launches
// at 1e6 keys).  Its onesweep radix sort takes five launches for 24-bit keys
// (digit histograms, their scan, three 8-bit passes) plus look-back resets,
// a fixed cost of about 75 us.  Measured on MI355X (random 24-bit keys with
// int values, us per sort, merge / onesweep): 2^13 23 / 77, 2^17 51 / 90,
// 2^18 89 / 92, 393,216 113 / 94, 2^20 160 / 102.  Onesweep therefore runs
// above ONESWEEP_MIN keys, and only where allowed: it clears its look-back
// state with hipMemsetAsync, which a captured graph must not record, so
// callers that may be captured (the caller-workspace paths, at most 65,536
// binaries) keep the memset-free merge sort, which rocPRIM uses up to 2^20
// keys.  Both sorts are stable and give the same order.
#pragma once

#include <rocprim/device/device_radix_sort.hpp>

namespace gbgpu {

constexpr int ONESWEEP_MIN = 1 << 18;

// Onesweep at every size above one block's worth of keys.
using OnesweepSort =
    rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config, rocprim::default_config, 0>;

// Sorts (keys, vals) by the low end_bit bits of keys into (keys_out,
// vals_out), by onesweep when allow_onesweep and n > ONESWEEP_MIN.  With
// tmp == nullptr only sets bytes, the temporary storage the same call needs
// (the query allocates nothing).
template <class K>
inline hipError_t sort_pairs(bool allow_onesweep, void *tmp, size_t &bytes, const K *keys, K *keys_out,
                             const int *vals, int *vals_out, int n, int end_bit, hipStream_t st) {
    if (allow_onesweep && n > ONESWEEP_MIN)
        return rocprim::radix_sort_pairs<OnesweepSort>(tmp, bytes, keys, keys_out, vals, vals_out, n, 0u,
                                                       (unsigned)end_bit, st);
    return rocprim::radix_sort_pairs(tmp, bytes, keys, keys_out, vals, vals_out, n, 0u, (unsigned)end_bit, st);
}

}  // namespace gbgpu
