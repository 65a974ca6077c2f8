// Drop-in likelihood and global-template kernels (gfx950).
//
// get_ll: templates arrive (M, num_bin)-major as in the reference
// (src/gbgpu/cutils/gbgpu_utils.cu:125-185).  A wavefront owns 8 binaries;
// lane -> (binary = lane & 7, row = lane >> 3), so every load instruction
// touches 8 template rows x 8 adjacent binaries (8 full 128-B lines) and
// 8 binaries x 8 consecutive data/PSD bins.  Each binary's partial sums are
// reduced with xor shuffles over the 8 lanes that share it.
//
// fill_global: deterministic replacement of the reference's atomic scatter
// (gbgpu_utils.cu:55-101).  Binaries are stably radix-sorted by
// key = group * (L + M) + (start + M); a workgroup then owns a 64-bin output
// tile, finds the contiguous key range of binaries overlapping it by binary
// search, splits it over its wavefronts and sums their samples in sorted
// order -- one lane per output bin, fixed summation order, so results are
// bit-identical run to run.

#include "common.hpp"
#include "fastmath.hpp"
#include "sort.hpp"

namespace gbgpu {

constexpr unsigned long long KEY_NONE = ~0ULL;
// rows per lane per iteration of k_get_ll: 1 measured fastest (0.88 of HBM
// peak at N = 256), 2 and 4 were 3% and 7% slower
constexpr int GET_LL_UNROLL = 1;

__global__ void __launch_bounds__(256) k_get_ll(double2 *__restrict__ d_h, double2 *__restrict__ h_h,
                                                const double2 *__restrict__ At, const double2 *__restrict__ Et,
                                                const double2 *__restrict__ Ad, const double2 *__restrict__ Ed,
                                                const double *__restrict__ SA, const double *__restrict__ SE,
                                                double df4, const int *__restrict__ start, int M, long long B,
                                                const int *__restrict__ di, const int *__restrict__ ni,
                                                long long L) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int bsub = lane & 7, row = lane >> 3;
    const long long b = ((long long)blockIdx.x * 4 + wave) * 8 + bsub;
    const bool valid = b < B;
    double dhr = 0.0, dhi = 0.0, hh = 0.0;
    if (valid) {
        const long long s = start[b];
        const long long doff = (di ? (long long)di[b] : 0LL) * L;
        const long long noff = (ni ? (long long)ni[b] : 0LL) * L;
        // branch-free rows: a row past M or a bin outside [0, L) reads a clamped
        // in-range element and adds nothing (a branch around the loads makes
        // the compiler wait for each row's loads before issuing the next)
        for (int i0 = row; i0 < M; i0 += 8 * GET_LL_UNROLL) {
            double2 hA[GET_LL_UNROLL], hE[GET_LL_UNROLL], dA[GET_LL_UNROLL], dE[GET_LL_UNROLL];
            double sA[GET_LL_UNROLL], sE[GET_LL_UNROLL];
            bool ok[GET_LL_UNROLL];
#pragma unroll
            for (int u = 0; u < GET_LL_UNROLL; u++) {
                const int i = i0 + 8 * u;
                const long long j = s + i;
                ok[u] = i < M && j >= 0 && j < L;
                const int ic = i < M ? i : M - 1;
                const long long jc = s + ic < 0 ? 0 : (s + ic >= L ? L - 1 : s + ic);
                const long long t = (long long)ic * B + b;
                hA[u] = At[t];
                hE[u] = Et[t];
                dA[u] = Ad[doff + jc];
                dE[u] = Ed[doff + jc];
                sA[u] = SA[noff + jc];
                sE[u] = SE[noff + jc];
            }
#pragma unroll
            for (int u = 0; u < GET_LL_UNROLL; u++) {
                const double iA = rcp_f64(sA[u]), iE = rcp_f64(sE[u]);
                const double cr = (dA[u].x * hA[u].x + dA[u].y * hA[u].y) * iA +
                                  (dE[u].x * hE[u].x + dE[u].y * hE[u].y) * iE;
                const double ci = (dA[u].x * hA[u].y - dA[u].y * hA[u].x) * iA +
                                  (dE[u].x * hE[u].y - dE[u].y * hE[u].x) * iE;
                const double cx = (hA[u].x * hA[u].x + hA[u].y * hA[u].y) * iA +
                                  (hE[u].x * hE[u].x + hE[u].y * hE[u].y) * iE;
                dhr += ok[u] ? cr : 0.0;
                dhi += ok[u] ? ci : 0.0;
                hh += ok[u] ? cx : 0.0;
            }
        }
    }
#pragma unroll
    for (int off = 8; off < 64; off <<= 1) {
        dhr += __shfl_xor(dhr, off, 64);
        dhi += __shfl_xor(dhi, off, 64);
        hh += __shfl_xor(hh, off, 64);
    }
    if (valid && row == 0) {
        d_h[b] = make_double2(df4 * dhr, df4 * dhi);
        h_h[b] = make_double2(df4 * hh, 0.0);
    }
}

// ------------------------------------------------------------ fill_global ---
__global__ void k_fill_keys(unsigned long long *__restrict__ keys, int *__restrict__ vals,
                            const int *__restrict__ start, const long long *__restrict__ start64,
                            const int *__restrict__ group, long long B, int M, long long L, int G,
                            long long start_shift) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const long long s = (start ? (long long)start[b] : start64[b]) - start_shift;
    const long long g = group ? group[b] : 0;
    unsigned long long key = KEY_NONE;
    if (g >= 0 && g < G && s > -(long long)M && s < L)
        key = (unsigned long long)g * (unsigned long long)(L + M) + (unsigned long long)(s + M);
    keys[b] = key;
    vals[b] = (int)b;
}

// (M, B)-major templates -> binary-major rows (B, M), both channels; 64x64
// tiles through LDS so both the reads and the writes are coalesced.
__global__ void __launch_bounds__(256) k_transpose(double2 *__restrict__ rowsA, double2 *__restrict__ rowsE,
                                                   const double2 *__restrict__ At,
                                                   const double2 *__restrict__ Et, int M, long long B) {
    __shared__ double2 tA[64][65], tE[64][65];
    const long long b0 = (long long)blockIdx.x * 64;
    const int i0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    for (int r = ty; r < 64; r += 4) {
        const int i = i0 + r;
        const long long b = b0 + tx;
        if (i < M && b < B) {
            tA[r][tx] = At[(long long)i * B + b];
            tE[r][tx] = Et[(long long)i * B + b];
        }
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const long long b = b0 + r;
        const int i = i0 + tx;
        if (i < M && b < B) {
            rowsA[b * M + i] = tA[tx][r];
            rowsE[b * M + i] = tE[tx][r];
        }
    }
}

__device__ __forceinline__ long long lower_bound_u64(const unsigned long long *__restrict__ a, long long n,
                                                     unsigned long long x) {
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// One workgroup per (group, 64-bin output tile).  rows: binary-major, row b
// starts at rows + b * row_stride; sample i of binary b is
// rows[b * row_stride + i].  The rows overlapping the tile (a contiguous range
// of the sorted keys) are split into GATHER_WAVES contiguous parts, one per
// wavefront; each wavefront adds its part in key order with GATHER_UNROLL
// row loads in flight, and the partial sums are combined in wavefront order.
// Every output bin's sum therefore has a fixed order: results are
// bit-identical from run to run.
constexpr int GATHER_WAVES = 4;
constexpr int GATHER_UNROLL = 8;

__global__ void __launch_bounds__(64 * GATHER_WAVES) k_gather_sum(
    double2 *__restrict__ Aglob, double2 *__restrict__ Eglob, const double2 *__restrict__ rowsA,
    const double2 *__restrict__ rowsE, long long row_stride, const unsigned long long *__restrict__ keys,
    const int *__restrict__ perm, long long B, int M, long long L, long long tiles_per_group) {
    __shared__ double4 part[GATHER_WAVES][64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const long long tile = blockIdx.x;
    const long long g = tile / tiles_per_group;
    const long long j0 = (tile - g * tiles_per_group) * 64;
    const unsigned long long KS = (unsigned long long)(L + M);
    const unsigned long long base = (unsigned long long)g * KS;
    // binaries with start in [j0 - M + 1, j0 + 63], i.e. key in [base + j0 + 1, base + j0 + 63 + M]
    const unsigned long long key_lo = base + (unsigned long long)j0 + 1ULL;
    const unsigned long long key_hi = base + (unsigned long long)(j0 + 64 + M);
    // tiles outside the sorted key range (most of them when the binaries
    // cover a narrow band) leave before the binary searches
    if (keys[0] >= key_hi || keys[B - 1] < key_lo) return;  // uniform over the workgroup
    const long long lo = lower_bound_u64(keys, B, key_lo);
    const long long hi = lower_bound_u64(keys, B, key_hi);
    if (hi <= lo) return;  // uniform over the workgroup
    const long long per = (hi - lo + GATHER_WAVES - 1) / GATHER_WAVES;
    // the wavefront's row range is wave-uniform: the sorted keys and the
    // permutation are read with scalar loads, which do not wait behind the
    // template rows in flight
    const long long r0 = lo + (long long)__builtin_amdgcn_readfirstlane(w) * per;
    const long long r1 = r0 + per < hi ? r0 + per : hi;
    const long long j = j0 + lane;
    double ar = 0.0, ai = 0.0, er = 0.0, ei = 0.0;
    // row metadata of one group of GATHER_UNROLL rows: first template bin
    // relative to the tile origin, and the template row; rows past r1 get an
    // empty window
    long long st[GATHER_UNROLL];
    long long rowoff[GATHER_UNROLL];
    auto meta = [&](long long r) {
#pragma unroll
        for (int u = 0; u < GATHER_UNROLL; u++) {
            const bool in = r + u < r1;
            st[u] = in ? (long long)(keys[r + u] - base) - M : j0 + 64;
            rowoff[u] = in ? (long long)perm[r + u] * row_stride : 0;
        }
    };
    if (r0 < r1) meta(r0);
    for (long long r = r0; r < r1; r += GATHER_UNROLL) {
        double2 va[GATHER_UNROLL], ve[GATHER_UNROLL];
        bool ok[GATHER_UNROLL];
        // branch-free loads (a divergent branch around each load makes the
        // compiler drain the loads in flight): lanes outside the row's window
        // read its clamped end element, one cache line, and add nothing
#pragma unroll
        for (int u = 0; u < GATHER_UNROLL; u++) {
            const long long i = j - st[u];
            ok[u] = i >= 0 && i < M;
            const long long ic = i < 0 ? 0 : (i >= M ? M - 1 : i);
            va[u] = rowsA[rowoff[u] + ic];
            ve[u] = rowsE[rowoff[u] + ic];
        }
        // the next group's metadata while this group's rows are in flight
        if (r + GATHER_UNROLL < r1) meta(r + GATHER_UNROLL);
#pragma unroll
        for (int u = 0; u < GATHER_UNROLL; u++) {
            ar += ok[u] ? va[u].x : 0.0;
            ai += ok[u] ? va[u].y : 0.0;
            er += ok[u] ? ve[u].x : 0.0;
            ei += ok[u] ? ve[u].y : 0.0;
        }
    }
    part[w][lane] = make_double4(ar, ai, er, ei);
    __syncthreads();
    if (w == 0 && j < L) {
        double4 t = part[0][lane];
#pragma unroll
        for (int k = 1; k < GATHER_WAVES; k++) {
            const double4 p = part[k][lane];
            t.x += p.x;
            t.y += p.y;
            t.z += p.z;
            t.w += p.w;
        }
        const long long o = g * L + j;
        double2 A = Aglob[o], E = Eglob[o];
        A.x += t.x;
        A.y += t.y;
        E.x += t.z;
        E.y += t.w;
        Aglob[o] = A;
        Eglob[o] = E;
    }
}

// Sort binaries by (group, start) and run the gather-sum over binary-major rows.
static int fill_rows_impl(double2 *Aglob, double2 *Eglob, const double2 *rowsA, const double2 *rowsE,
                          long long row_stride, const int *start, const long long *start64, long long start_shift,
                          int M, long long B, const int *group, int G, long long L, hipStream_t st) {
    if (B == 0 || G == 0) return GBGPU_OK;
    unsigned long long *keys = nullptr, *keys_sorted = nullptr;
    int *vals = nullptr, *perm = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    hipError_t e;
    // Keys are < G (L + M), or KEY_NONE, whose low bits are all ones.  Sorting
    // the low end_bit bits with 2^end_bit > G (L + M) therefore keeps the full
    // 64-bit order (KEY_NONE last) and the sort stable, in ceil(end_bit / 8)
    // radix passes (3 for one 4-yr stream) instead of 8.
    const unsigned long long span = (unsigned long long)G * (unsigned long long)(L + M);
    int end_bit = 1;
    while (end_bit < 64 && (1ULL << end_bit) <= span) end_bit++;
    e = sort_pairs<unsigned long long>(true, nullptr, tmp_bytes, keys, keys_sorted, vals, perm, (int)B, end_bit, st);
    if (e != hipSuccess) return hip_fail(e, "SortPairs(size)");
    const size_t kb = sizeof(unsigned long long) * (size_t)B, vb = sizeof(int) * (size_t)B;
    char *ws = nullptr;
    const size_t total = 2 * kb + 2 * vb + tmp_bytes + 256 * 4;
    e = hipMallocAsync((void **)&ws, total, st);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(fill workspace)");
    size_t off = 0;
    auto carve = [&](size_t n) {
        char *p = ws + off;
        off += (n + 255) & ~(size_t)255;
        return p;
    };
    keys = (unsigned long long *)carve(kb);
    keys_sorted = (unsigned long long *)carve(kb);
    vals = (int *)carve(vb);
    perm = (int *)carve(vb);
    tmp = carve(tmp_bytes);
    hipLaunchKernelGGL(k_fill_keys, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st, keys, vals, start, start64,
                       group, B, M, L, G, start_shift);
    e = sort_pairs<unsigned long long>(true, tmp, tmp_bytes, keys, keys_sorted, vals, perm, (int)B, end_bit, st);
    if (e != hipSuccess) {
        (void)hipFreeAsync(ws, st);
        return hip_fail(e, "SortPairs");
    }
    const long long tpg = (L + 63) / 64;
    const long long nt = tpg * G;
    if (nt > 0x7fffffffLL) {
        (void)hipFreeAsync(ws, st);
        set_error("fill: num_groups * ceil(data_length / 64) exceeds the launch limit");
        return GBGPU_E_ARG;
    }
    {
        KernelTimer kt(st);
        hipLaunchKernelGGL(k_gather_sum, dim3((unsigned)nt), dim3(64 * GATHER_WAVES), 0, st, Aglob, Eglob, rowsA,
                           rowsE, row_stride, keys_sorted, perm, B, M, L, tpg);
    }
    e = hipGetLastError();
    hipError_t e2 = hipFreeAsync(ws, st);
    if (e != hipSuccess) return hip_fail(e, "k_gather_sum");
    if (e2 != hipSuccess) return hip_fail(e2, "hipFreeAsync");
    return GBGPU_OK;
}

}  // namespace gbgpu

using namespace gbgpu;

extern "C" int gbgpu_get_ll(gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const gbgpu_cmplx *A_template,
                            const gbgpu_cmplx *E_template, const gbgpu_cmplx *A_data, const gbgpu_cmplx *E_data,
                            const double *A_psd, const double *E_psd, double df, const int *start_ind, int M,
                            int num_bin, const int *data_index, const int *noise_index, int data_length,
                            void *stream) {
    if (num_bin < 0 || M < 0 || data_length <= 0) {
        set_error("gbgpu_get_ll: negative num_bin/M or data_length <= 0");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    if (!d_h || !h_h || !A_template || !E_template || !A_data || !E_data || !A_psd || !E_psd || !start_ind) {
        set_error("gbgpu_get_ll: null pointer argument");
        return GBGPU_E_ARG;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const long long waves = ((long long)num_bin + 7) / 8;
    KernelTimer kt(st);
    hipLaunchKernelGGL(k_get_ll, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, (double2 *)d_h, (double2 *)h_h,
                       (const double2 *)A_template, (const double2 *)E_template, (const double2 *)A_data,
                       (const double2 *)E_data, A_psd, E_psd, 4.0 * df, start_ind, M, (long long)num_bin,
                       data_index, noise_index, (long long)data_length);
    GBGPU_CHECK_LAUNCH("k_get_ll");
    return GBGPU_OK;
}

extern "C" int gbgpu_fill_global(gbgpu_cmplx *A_glob, gbgpu_cmplx *E_glob, const gbgpu_cmplx *A_template,
                                 const gbgpu_cmplx *E_template, const int *start_ind, int M, int num_bin,
                                 const int *group_index, int data_length, int num_groups, void *stream) {
    if (num_bin < 0 || M < 0 || data_length <= 0 || num_groups < 0) {
        set_error("gbgpu_fill_global: negative size argument");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0 || M == 0 || num_groups == 0) return GBGPU_OK;
    if (!A_glob || !E_glob || !A_template || !E_template || !start_ind || !group_index) {
        set_error("gbgpu_fill_global: null pointer argument");
        return GBGPU_E_ARG;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t rows_bytes = sizeof(double2) * (size_t)num_bin * (size_t)M;
    double2 *rows = nullptr;
    hipError_t e = hipMallocAsync((void **)&rows, 2 * rows_bytes, st);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(rows)");
    dim3 grid((unsigned)(((long long)num_bin + 63) / 64), (unsigned)((M + 63) / 64));
    hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, st, rows, rows + (size_t)num_bin * M,
                       (const double2 *)A_template, (const double2 *)E_template, M, (long long)num_bin);
    e = hipGetLastError();
    if (e != hipSuccess) {
        (void)hipFreeAsync(rows, st);
        return hip_fail(e, "k_transpose");
    }
    int rc = fill_rows_impl((double2 *)A_glob, (double2 *)E_glob, rows, rows + (size_t)num_bin * M, M, start_ind,
                            nullptr, 0, M, num_bin, group_index, num_groups, data_length, st);
    e = hipFreeAsync(rows, st);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipFreeAsync(rows)");
    return GBGPU_OK;
}

extern "C" int gbgpu_fill_global_rows(gbgpu_cmplx *A_glob, gbgpu_cmplx *E_glob, const gbgpu_cmplx *A_rows,
                                      const gbgpu_cmplx *E_rows, long long row_stride, const long long *start_inds,
                                      long long start_freq_ind, int M, int num_bin, const int *group_index,
                                      int data_length, int num_groups, void *stream) {
    if (num_bin < 0 || M < 0 || data_length <= 0 || num_groups < 0 || row_stride < M) {
        set_error("gbgpu_fill_global_rows: bad size argument");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0 || M == 0 || num_groups == 0) return GBGPU_OK;
    if (!A_glob || !E_glob || !A_rows || !E_rows || !start_inds) {
        set_error("gbgpu_fill_global_rows: null pointer argument");
        return GBGPU_E_ARG;
    }
    return fill_rows_impl((double2 *)A_glob, (double2 *)E_glob, (const double2 *)A_rows, (const double2 *)E_rows,
                          row_stride, nullptr, start_inds, start_freq_ind, M, num_bin, group_index, num_groups,
                          data_length, static_cast<hipStream_t>(stream));
}
