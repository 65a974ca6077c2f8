// Stable key/value sorts of the start-bin orderings.
//
// Three sorts, chosen by size.  Measured on MI355X (random 24-bit keys with
// int values, us per sort; merge = rocPRIM's default configuration, which
// merge-sorts up to 2^20 keys; onesweep = rocPRIM's onesweep radix sort;
// lsd = lsd_sort below):
//
//   keys        merge  onesweep  lsd
//   2^13           23        77   49
//   2^16           38        88   51
//   2^18           93        94   54
//   1e6           160       103   71
//   4e6           130       130  164
//
// Up to LSD_MIN keys the merge sort runs, up to LSD_MAX lsd_sort, above it
// onesweep where allowed.  Onesweep clears its look-back state with
// hipMemsetAsync, which a captured graph must not record, so callers that
// may be captured (the caller-workspace paths, at most 65,536 binaries)
// pass allow_onesweep = false and get rocPRIM's default above LSD_MAX.
// lsd_sort issues no memset and no global atomic.  All three are stable and
// give the same order.
#pragma once

#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"

namespace gbgpu {

constexpr int LSD_MIN = 1 << 17;
constexpr int LSD_MAX = 1 << 21;

// Onesweep at every size above one block's worth of keys.
using OnesweepSort =
    rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config, rocprim::default_config, 0>;

// ---------------------------------------------------------------- lsd_sort ---
// LSD radix sort with 8-bit digits over tiles of LSD_TILE items, three
// launches per digit: per-tile digit counts, their scan, and a stable
// scatter.  Counts are tile-major, cnt[t * 256 + d]; the scan rewrites them
// with the exclusive prefix of digit d over tiles before t and writes the
// digit totals to cnt[256 T + d].
constexpr int LSD_TPB = 256, LSD_IPT = 16, LSD_TILE = LSD_TPB * LSD_IPT, LSD_NW = LSD_TPB / 64;

template <class K>
__global__ void __launch_bounds__(LSD_TPB) k_lsd_count(const K *__restrict__ kin, int n, int shift,
                                                       unsigned *__restrict__ cnt) {
    __shared__ unsigned c[256];
    c[threadIdx.x] = 0;
    __syncthreads();
    const long long t0 = (long long)blockIdx.x * LSD_TILE;
#pragma unroll
    for (int s = 0; s < LSD_IPT; s++) {
        const long long g = t0 + threadIdx.x + (long long)s * LSD_TPB;
        if (g < n) atomicAdd(&c[(unsigned)(kin[g] >> shift) & 255u], 1u);
    }
    __syncthreads();
    cnt[(long long)blockIdx.x * 256 + threadIdx.x] = c[threadIdx.x];
}

// One workgroup of 256 threads per digit d (grid 256).
template <class K>
__global__ void __launch_bounds__(256) k_lsd_scan(unsigned *__restrict__ cnt, int T) {
    __shared__ unsigned wsum[4];
    const int d = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned carry = 0;
    for (int c0 = 0; c0 < T; c0 += 256) {
        const int t = c0 + (int)threadIdx.x;
        const unsigned x = t < T ? cnt[(long long)t * 256 + d] : 0u;
        unsigned incl = x;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned u = __shfl_up(incl, off, 64);
            if (lane >= off) incl += u;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        unsigned pre = 0;
        for (int j = 0; j < w; j++) pre += wsum[j];
        const unsigned total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (t < T) cnt[(long long)t * 256 + d] = carry + pre + incl - x;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[256LL * T + d] = carry;
}

// A tile's items are taken in (wavefront, step, lane) order, item
// w 64 LSD_IPT + s 64 + lane, which is their input order.  Ranks within a
// wavefront come from the match mask of the lanes holding the same digit; a
// per-wavefront running count in LDS carries them across steps, and the
// wavefronts' counts are prefixed in order, so the scatter is stable.
template <class K>
__global__ void __launch_bounds__(LSD_TPB) k_lsd_scatter(const K *__restrict__ kin, const int *__restrict__ vin,
                                                         K *__restrict__ kout, int *__restrict__ vout, int n,
                                                         int shift, const unsigned *__restrict__ cnt, int T) {
    __shared__ unsigned run[LSD_NW][256];
    __shared__ unsigned base[LSD_NW][256];
    __shared__ unsigned wtot[LSD_NW];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < LSD_NW; j++) run[j][threadIdx.x] = 0;
    __syncthreads();
    const long long g0 = (long long)blockIdx.x * LSD_TILE + (long long)w * 64 * LSD_IPT + lane;
    K k[LSD_IPT];
    int v[LSD_IPT];
    unsigned r[LSD_IPT];
#pragma unroll
    for (int s = 0; s < LSD_IPT; s++) {
        const long long g = g0 + 64 * s;
        const bool valid = g < n;
        k[s] = valid ? kin[g] : (K)0;
        v[s] = valid ? vin[g] : 0;
    }
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int s = 0; s < LSD_IPT; s++) {
        const bool valid = g0 + 64 * s < n;
        const unsigned d = (unsigned)(k[s] >> shift) & 255u;
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const unsigned before = run[w][d];
        r[s] = before + (unsigned)__popcll(m & below);
        // the group's highest lane adds the group's size
        if (valid && (m >> lane) == 1ull) run[w][d] = before + (unsigned)__popcll(m);
    }
    {
        // where each digit starts: exclusive scan of the digit totals
        const unsigned tot = cnt[256LL * T + threadIdx.x];
        unsigned incl = tot;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned u = __shfl_up(incl, off, 64);
            if (lane >= off) incl += u;
        }
        if (lane == 63) wtot[w] = incl;
        __syncthreads();  // also completes every wavefront's running counts
        unsigned acc = incl - tot + cnt[(long long)blockIdx.x * 256 + threadIdx.x];
        for (int j = 0; j < w; j++) acc += wtot[j];
#pragma unroll
        for (int j = 0; j < LSD_NW; j++) {
            base[j][threadIdx.x] = acc;
            acc += run[j][threadIdx.x];
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < LSD_IPT; s++) {
        if (g0 + 64 * s < n) {
            const unsigned dst = base[w][(unsigned)(k[s] >> shift) & 255u] + r[s];
            if (dst < (unsigned)n) {
                kout[dst] = k[s];
                vout[dst] = v[s];
            }
        }
    }
}

inline size_t lsd_align(size_t x) { return (x + 255) & ~(size_t)255; }

// Temporary storage of lsd_sort: a key and a value buffer of n each, and
// 256 (T + 1) counters.
template <class K>
inline size_t lsd_sort_bytes(int n) {
    const size_t T = ((size_t)n + LSD_TILE - 1) / LSD_TILE;
    return lsd_align(sizeof(K) * (size_t)n) + lsd_align(4 * (size_t)n) + lsd_align(4 * 256 * (T + 1));
}

// Sorts (kin, vin) by the low end_bit bits of the keys into (kout, vout):
// ceil(end_bit / 8) passes, alternating with tmp's buffers so that the last
// one writes kout / vout.
template <class K>
inline hipError_t lsd_sort(void *tmp, const K *kin, K *kout, const int *vin, int *vout, int n, int end_bit,
                           hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int T = (n + LSD_TILE - 1) / LSD_TILE, P = (end_bit + 7) / 8;
    K *ktmp = (K *)tmp;
    int *vtmp = (int *)((char *)tmp + lsd_align(sizeof(K) * (size_t)n));
    unsigned *cnt = (unsigned *)((char *)vtmp + lsd_align(4 * (size_t)n));
    const K *sk = kin;
    const int *sv = vin;
    for (int p = 0; p < P; p++) {
        const bool to_out = ((P - 1 - p) % 2) == 0;
        K *dk = to_out ? kout : ktmp;
        int *dv = to_out ? vout : vtmp;
        hipLaunchKernelGGL(k_lsd_count<K>, dim3(T), dim3(LSD_TPB), 0, st, sk, n, 8 * p, cnt);
        hipLaunchKernelGGL(k_lsd_scan<K>, dim3(256), dim3(256), 0, st, cnt, T);
        hipLaunchKernelGGL(k_lsd_scatter<K>, dim3(T), dim3(LSD_TPB), 0, st, sk, sv, dk, dv, n, 8 * p, cnt, T);
        sk = dk;
        sv = dv;
    }
    return hipGetLastError();
}

// Sorts (keys, vals) by the low end_bit bits of keys into (keys_out,
// vals_out), choosing the sort by n as above.  With tmp == nullptr only
// sets bytes, the temporary storage the same call needs (the query
// allocates nothing).
template <class K>
inline hipError_t sort_pairs(bool allow_onesweep, void *tmp, size_t &bytes, const K *keys, K *keys_out,
                             const int *vals, int *vals_out, int n, int end_bit, hipStream_t st) {
    if (n > LSD_MIN && n <= LSD_MAX) {
        if (!tmp) {
            bytes = lsd_sort_bytes<K>(n);
            return hipSuccess;
        }
        return lsd_sort<K>(tmp, keys, keys_out, vals, vals_out, n, end_bit, st);
    }
    if (allow_onesweep && n > LSD_MAX)
        return rocprim::radix_sort_pairs<OnesweepSort>(tmp, bytes, keys, keys_out, vals, vals_out, n, 0u,
                                                       (unsigned)end_bit, st);
    return rocprim::radix_sort_pairs(tmp, bytes, keys, keys_out, vals, vals_out, n, 0u, (unsigned)end_bit, st);
}

}  // namespace gbgpu
