// Uneven-N batches (gfx950).
//
// The reference evaluates one N per call (src/gbgpu/gbgpu.py:205-207 takes
// N.max(); the native get_ll takes a scalar M, gbgpu_utils.cu:126), so an
// uneven batch costs one call per N.  Here every binary carries its own N
// (a power of two in [32, 8192]) and one call does, on the device:
//   k_batch_reset    clear the control block (a fill kernel, no memset)
//   k_batch_setup    per-binary constants (Pack) and a sort key (segment:
//                    N bucket and route, then log-spaced start-bin cell);
//                    with the mixed accuracy guard, binaries whose
//                    estimated <h|h> already fails it are routed to an FP64
//                    segment
//   lsd_sort         the binaries into one list ordered by that key
//                    (neighbouring binaries read neighbouring data windows,
//                    which stay in L2); stable, so the list is deterministic
//   k_batch_bounds   each segment's first and last list position
//   k_batch_table    per-segment counts and the persistent work table
// and then either
//   strategy 0: one launch of the uniform fused kernel per N bucket, each a
//               grid-stride loop over its segment of the list, reading the
//               segment's size and offset from the control block; or
//   strategy 1: one persistent launch whose workgroups pull work items (BPW
//               binaries of one bucket, largest N first) from an atomic queue.
// All scratch lives in a caller-provided workspace and nothing depends on a
// host-side count, so the sequence can be captured in a hipGraph and
// replayed (the per-N orbit tables and twiddles must have been created by an
// earlier, uncaptured call).  Which workgroup runs a binary depends on
// scheduling; each binary's result does not, so results are deterministic.
#include <atomic>

#include "common.hpp"
#include "sort.hpp"
#include "wave_common.hpp"

namespace gbgpu {

int check_data(const gbgpu_data *data);
void set_data(WaveArgs &a, const gbgpu_data *data);

int launch_ll_segment(int N, int precision, int pack_precision, int variant, bool third, const WaveArgs &a,
                      hipStream_t st);

constexpr int NBUCKET = 9;  // N = 32 << k, k < 9
// The persistent strategy (switched off) covers N <= 2048 only.
constexpr int NBUCKET_PERSISTENT = 7;
// List segments: segment k holds bucket k evaluated in the batch's
// precision, segment NBUCKET + k the binaries of bucket k routed to FP64.
constexpr int NSEG = 2 * NBUCKET;
constexpr int NCB = 4096;  // start-bin cells per segment (log-spaced)
constexpr int NCELL = NSEG * NCB;
enum { CTL_COUNT = 0, CTL_LISTOFF = 32, CTL_ITEMOFF = 64, CTL_ITEMS = 80, CTL_TOTAL = 96, CTL_WORK = 97,
       CTL_REDO = 104, CTL_WORDS = 128 };
static_assert(CTL_LISTOFF - CTL_COUNT >= NSEG && CTL_ITEMOFF - CTL_LISTOFF >= NSEG &&
                  CTL_ITEMS - CTL_ITEMOFF >= NBUCKET && CTL_TOTAL - CTL_ITEMS >= NBUCKET &&
                  CTL_WORDS - CTL_REDO >= NBUCKET,
              "control block layout");
// Sort keys: segment * NCB + cell, below 2^BATCH_KEY_BITS; binaries that are
// not evaluated get BATCH_KEY_NONE, which sorts last and lies in no segment.
constexpr int BATCH_KEY_BITS = 17;
constexpr unsigned BATCH_KEY_NONE = (1u << BATCH_KEY_BITS) - 1u;
static_assert(NCELL <= (int)BATCH_KEY_NONE && BATCH_KEY_NONE / NCB >= (unsigned)NSEG, "sort key range");

__host__ __device__ constexpr int bucket_bpw(int k) {
    return WG_SIZE / (((32 << k) / 4 < WG_SIZE) ? (32 << k) / 4 : WG_SIZE);
}

__device__ __forceinline__ int bucket_of(int N) {
    if (N < 32 || N > (32 << (NBUCKET - 1)) || (N & (N - 1))) return -1;
    return 26 - __clz(N);  // log2(N) - 5 for 32-bit N
}

struct BatchArgs {
    WaveArgs a;
    const double *tabs[NBUCKET];
    const void *tws[NBUCKET];
    const int *Nb;
    int *list;
    unsigned *key;  // sort key per binary (BATCH_KEY_NONE: not evaluated)
    int *idx;       // binary indices 0 .. B-1, the values the sort carries
    int *redo;  // accuracy-guard list (same per-bucket offsets as `list`)
    int *ctl;
    long long cap[NBUCKET];  // host-side bound on each bucket's size (0: none)
    double guard_tol;
    double route_tol;  // > 0: route binaries whose estimated error bound exceeds it to FP64
    int transform;     // GBGPU_TRANSFORM_* requested for the mixed N = 256 bucket
    bool third;        // 14 parameter rows (third-body binaries)
};

__global__ void k_batch_reset(int *ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < CTL_WORDS) ctl[i] = 0;
}

// Estimated <h|h> of a binary from its amplitude, frequency and the A-channel
// PSD of its noise stream at its centre bin: 1.1 A^2 T (4 wL sin wL)^2
// |fac_TDI2|^2 / S_A(f0), wL = f0/f* (fac_TDI2 = 1 for TDI1).  The factor 1.1
// is the median, over sky positions and orientations, of the exact <h|h>
// over the rest of the expression (0.27 to 3.84 across 2400 sources at
// 0.1-20 mHz, 4 yr); the estimate only routes work, and binaries it misses
// are still caught by the guard after the mixed pass.
__device__ __forceinline__ double hh_estimate(const Src &s, long long b, int N, const WaveArgs &a) {
    long long j = s.start + N / 2 - a.start_freq_ind;
    j = j < 0 ? 0 : (j >= a.L ? a.L - 1 : j);
    const long long noff = (a.ni ? (long long)a.ni[b] : 0LL) * a.L;
    const double wL = s.f0 * a.fstar_inv;
    const double r = 4.0 * wL * sin(wL);
    return 1.1 * s.amp * s.amp * a.T * r * r * (s.tdr * s.tdr + s.tdi * s.tdi) / a.SA[noff + j];
}

// Sort key of a binary: segment k, then its start bin within the data on a
// logarithmic grid of NCB cells.  Galactic binaries crowd at low frequency;
// log-spaced cells resolve the crowded band finely, and neighbouring
// binaries land in neighbouring cells.
__device__ __forceinline__ int batch_cell(int k, long long start, const WaveArgs &a) {
    long long j = start - a.start_freq_ind;
    j = j < 0 ? 0 : (j >= a.L ? a.L - 1 : j);
    const float x = __log2f(1.0f + (float)j) / __log2f(1.0f + (float)a.L);
    int c = (int)(x * (float)NCB);
    c = c < 0 ? 0 : (c >= NCB ? NCB - 1 : c);
    return k * NCB + c;
}

template <class P>
__global__ void __launch_bounds__(256) k_batch_setup(BatchArgs ba, Pack<P> *__restrict__ pack) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ba.a.B) return;
    const int N = ba.Nb[b];
    const int k = bucket_of(N);
    if (k < 0 || ba.tabs[k] == nullptr) {
        // unsupported N or no orbit table for it: flag the result, skip the binary
        const double nan = __builtin_nan("");
        ba.a.loglike[b] = nan;
        if (ba.a.d_h) ba.a.d_h[b] = make_double2(nan, nan);
        if (ba.a.h_h) ba.a.h_h[b] = make_double2(nan, nan);
        ba.key[b] = BATCH_KEY_NONE;
        ba.idx[b] = (int)b;
        return;
    }
    Pack<P> pk;
    pk.s = make_src(ba.a.params, ba.a.num_params, ba.a.B, b, N, ba.a.T, ba.a.tdi2, ba.a.armlength);
    pk.aux = P::make_aux(pk.s);
    pack[b] = pk;
    const bool routed = ba.route_tol > 0.0 && GUARD_REL * hh_estimate(pk.s, b, N, ba.a) > ba.route_tol;
    const int cell = batch_cell(routed ? NBUCKET + k : k, pk.s.start, ba.a);
    ba.key[b] = (unsigned)cell;
    ba.idx[b] = (int)b;
}

// Segment boundaries in the sorted keys: the first list position of
// segment s goes to ctl[CTL_LISTOFF + s] and one past its last to
// ctl[CTL_COUNT + s] (k_batch_table turns it into a count); both stay zero
// for an empty segment.
__global__ void k_batch_bounds(const unsigned *__restrict__ key, long long B, int *__restrict__ ctl) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const unsigned s = key[i] / NCB;
    if (s >= (unsigned)NSEG) return;
    if (i == 0 || key[i - 1] / NCB != s) ctl[CTL_LISTOFF + s] = (int)i;
    if (i == B - 1 || key[i + 1] / NCB != s) ctl[CTL_COUNT + s] = (int)(i + 1);
}

// Per-segment counts, and the persistent strategy's work-item table
// (descending bucket).  One thread.
__global__ void k_batch_table(int *ctl) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int k = 0; k < NSEG; k++) ctl[CTL_COUNT + k] -= ctl[CTL_LISTOFF + k];
    int item_run = 0;
    for (int k = NBUCKET_PERSISTENT - 1; k >= 0; k--) {
        const int c = ctl[CTL_COUNT + k];
        const int bpw = bucket_bpw(k);
        ctl[CTL_ITEMOFF + k] = item_run;
        ctl[CTL_ITEMS + k] = (c + bpw - 1) / bpw;
        item_run += (c + bpw - 1) / bpw;
    }
    ctl[CTL_TOTAL] = item_run;
}

template <int N, class P>
struct BatchLds {
    using G = Geo<N>;
    static constexpr size_t pack_bytes = (sizeof(Pack<P>) * G::BPW + 15) & ~(size_t)15;
    static constexpr size_t tot_bytes = (sizeof(double) * G::BPW * G::WPB + 15) & ~(size_t)15;
    static constexpr size_t buf_bytes = sizeof(typename P::CT) * G::BPW * 2 * N;
    static constexpr size_t total = pack_bytes + tot_bytes + buf_bytes;
};

template <class P>
constexpr size_t batch_lds_bytes() {
    size_t m = BatchLds<32, P>::total;
    m = m > BatchLds<64, P>::total ? m : BatchLds<64, P>::total;
    m = m > BatchLds<128, P>::total ? m : BatchLds<128, P>::total;
    m = m > BatchLds<256, P>::total ? m : BatchLds<256, P>::total;
    m = m > BatchLds<512, P>::total ? m : BatchLds<512, P>::total;
    m = m > BatchLds<1024, P>::total ? m : BatchLds<1024, P>::total;
    m = m > BatchLds<2048, P>::total ? m : BatchLds<2048, P>::total;
    return m;
}

// One work item: up to BPW binaries of bucket k (N = 32 << k).
template <int N, class P>
__device__ __noinline__ void batch_item(const BatchArgs &ba, int k, int local, int count, int list_off,
                                           char *lds) {
    using G = Geo<N>;
    using CT = typename P::CT;
    using L = BatchLds<N, P>;
    constexpr int TPB = G::TPB, BPW = G::BPW, WPB = G::WPB;
    Pack<P> *spack = reinterpret_cast<Pack<P> *>(lds);
    double *stot = reinterpret_cast<double *>(lds + L::pack_bytes);
    CT *sbuf = reinterpret_cast<CT *>(lds + L::pack_bytes + L::tot_bytes);
    const int tid = threadIdx.x;
    const int slot = tid / TPB;
    const int lane = tid % TPB;
    const int idx = local * BPW + slot;
    const bool valid = idx < count;
    const long long b = valid ? (long long)ba.list[list_off + idx] : 0;
    if (valid) load_pack<P, TPB>(ba.a, b, lane, spack + slot);
    __syncthreads();
    const Src &s = spack[slot].s;  // read in place from LDS
    CT *buf = sbuf + slot * 2 * N;
    double *tot = stot + slot * WPB;
    ThirdScratch ts{nullptr};
    binary_spectrum<N, 2, false, P>(s, spack[slot].aux, valid, lane, ba.tabs[k], ba.a.fstar_inv, 0.5 * ba.a.T / N,
                                    buf, static_cast<const CT *>(ba.tws[k]), ts, tot);
    const double2 fac = P::fac(s);
    double dhr = 0.0, dhi = 0.0, hh = 0.0;
    if (valid) {
        const long long doff = (ba.a.di ? (long long)ba.a.di[b] : 0LL) * ba.a.L;
        const long long noff = (ba.a.ni ? (long long)ba.a.ni[b] : 0LL) * ba.a.L;
        ll_partial<N, P>(buf, fac, lane, s.start - ba.a.start_freq_ind, doff, noff, ba.a, dhr, dhi, hh);
    }
    dhr = group_total<TPB>(dhr, lane, tot);
    dhi = group_total<TPB>(dhi, lane, tot);
    hh = group_total<TPB>(hh, lane, tot);
    if (valid && lane == 0) {
        dhr *= ba.a.df4;
        dhi *= ba.a.df4;
        hh *= ba.a.df4;
        if (ba.a.d_h) ba.a.d_h[b] = make_double2(dhr, dhi);
        if (ba.a.h_h) ba.a.h_h[b] = make_double2(hh, 0.0);
        const double dh_use = ba.a.phase_marg ? sqrt(dhr * dhr + dhi * dhi) : dhr;
        ba.a.loglike[b] = -0.5 * (ba.a.d_d + hh - 2.0 * dh_use);
    }
    __syncthreads();  // LDS is reused by the next item
}

template <class P>
__device__ __forceinline__ void batch_loop(const BatchArgs &ba, char *lds) {
    __shared__ int s_item;
    const int total = ba.ctl[CTL_TOTAL];
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&ba.ctl[CTL_WORK], 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= total) break;  // uniform: every thread read the same item
        int k = NBUCKET_PERSISTENT - 1;
        while (k > 0 && item >= ba.ctl[CTL_ITEMOFF + k] + ba.ctl[CTL_ITEMS + k]) k--;
        const int local = item - ba.ctl[CTL_ITEMOFF + k];
        const int count = ba.ctl[CTL_COUNT + k];
        const int loff = ba.ctl[CTL_LISTOFF + k];
        switch (k) {
            case 0: batch_item<32, P>(ba, k, local, count, loff, lds); break;
            case 1: batch_item<64, P>(ba, k, local, count, loff, lds); break;
            case 2: batch_item<128, P>(ba, k, local, count, loff, lds); break;
            case 3: batch_item<256, P>(ba, k, local, count, loff, lds); break;
            case 4: batch_item<512, P>(ba, k, local, count, loff, lds); break;
            case 5: batch_item<1024, P>(ba, k, local, count, loff, lds); break;
            default: batch_item<2048, P>(ba, k, local, count, loff, lds); break;
        }
    }
}

// The per-N item bodies are separate (non-inlined) functions so each gets
// its own register allocation; they read the arguments in place from the
// kernel-argument segment instead of a private copy.  Occupancy targets:
// 3 waves/SIMD for FP64 (~150 VGPRs needed), 4 for mixed.
__global__ void __launch_bounds__(WG_SIZE) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_batch_ll_f64(BatchArgs ba) {
    (void)ba;
    extern __shared__ __align__(16) char lds[];
    batch_loop<F64>(*(const BatchArgs *)__builtin_amdgcn_kernarg_segment_ptr(), lds);
}

__global__ void __launch_bounds__(WG_SIZE) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_batch_ll_mixed(BatchArgs ba) {
    (void)ba;
    extern __shared__ __align__(16) char lds[];
    batch_loop<Mixed>(*(const BatchArgs *)__builtin_amdgcn_kernarg_segment_ptr(), lds);
}

template <class P>
struct BatchKernel;
template <>
struct BatchKernel<F64> {
    static void *fn() { return (void *)k_batch_ll_f64; }
};
template <>
struct BatchKernel<Mixed> {
    static void *fn() { return (void *)k_batch_ll_mixed; }
};

// Byte offset of the control block in a batch workspace of B binaries whose
// packs take pack_bytes each: packs, then five int arrays of B (list, keys,
// redo list, indices, sorted keys), each piece 256-byte aligned.
inline size_t batch_ctl_offset(size_t pack_bytes, long long B) {
    auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
    return al(pack_bytes * (size_t)B) + 5 * al(sizeof(int) * (size_t)B);
}

template <class P>
size_t batch_ws_bytes(long long B) {
    auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
    return batch_ctl_offset(sizeof(Pack<P>), B) + al(sizeof(int) * CTL_WORDS) + lsd_sort_bytes<unsigned>((int)B);
}

template <class P>
static int launch_persistent(BatchArgs &ba, hipStream_t st) {
    // persistent grid: as many workgroups as can be resident at once
    // (initialised once per policy, thread-safe)
    const size_t lds = batch_lds_bytes<P>();
    static const long long grid = [lds] {
        int dev = 0, num_cu = 0, nb = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipFuncSetAttribute(BatchKernel<P>::fn(), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, BatchKernel<P>::fn(), WG_SIZE, lds) != hipSuccess ||
            nb < 1)
            nb = 1;
        return (long long)nb * (num_cu > 0 ? num_cu : 256);
    }();
    void *kargs[] = {&ba};
    hipError_t e;
    {
        KernelTimer kt(st);
        e = hipLaunchKernel(BatchKernel<P>::fn(), dim3((unsigned)grid), dim3(WG_SIZE), kargs, lds, st);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "k_batch_ll");
    return GBGPU_OK;
}

// Variant of the per-bucket kernel, per (precision, bucket): 0 = N/4 threads
// per binary, 1 = half as many; for the mixed N = 256 bucket also
// VARIANT_DFT0 + t - 1 = matrix-core transform t (GBGPU_TRANSFORM_*); set by
// gbgpu_batch_autotune or gbgpu_batch_set_variants.
static std::atomic<int> g_variant[2][NBUCKET] = {};

// One launch of the uniform fused kernel per populated bucket.  With the
// accuracy guard (mixed precision, guard_tol > 0), each mixed launch lists
// the binaries whose error bound exceeds guard_tol and an FP64 launch over
// that list (reading the same mixed per-binary constants) overwrites their
// results.
// variants: tile shape per bucket (NULL: the tuned table); bucket_events:
// when non-NULL, a pair of HIP events is recorded around each bucket's
// launches (2 * NBUCKET entries, created by the caller).
template <class P>
static int launch_buckets(const BatchArgs &ba, int precision, hipStream_t st, const int *variants,
                          hipEvent_t *bucket_events) {
    const bool guard = precision == GBGPU_MIXED && ba.guard_tol > 0.0;
    const bool route = precision == GBGPU_MIXED && ba.route_tol > 0.0;
    for (int k = NBUCKET - 1; k >= 0; k--) {
        if (!ba.tabs[k]) continue;
        const int N = 32 << k;
        int var = variants ? variants[k] : g_variant[precision == GBGPU_MIXED ? 1 : 0][k].load();
        // a requested matrix-core transform of the mixed N = 256 bucket wins;
        // third-body binaries take the LDS transform on the same tile shape
        if (precision == GBGPU_MIXED && N == 256 && ba.transform != GBGPU_TRANSFORM_LDS)
            var = VARIANT_DFT0 + ba.transform - 1;
        if (ba.third && var >= VARIANT_DFT0) var = 0;
        const int var64 = g_variant[0][k].load();
        if (bucket_events) (void)hipEventRecord(bucket_events[2 * k], st);
        WaveArgs a = ba.a;
        a.tab = ba.tabs[k];
        a.tw = ba.tws[k];
        a.scale = 0.5 * a.T / N;
        a.order = ba.list;
        a.seg_count = ba.ctl + CTL_COUNT + k;
        a.seg_off = ba.ctl + CTL_LISTOFF + k;
        a.seg_cap = ba.cap[k];  // also bounds the guard's and the routed FP64 passes
        if (guard) {
            a.redo_list = ba.redo;
            a.redo_count = ba.ctl + CTL_REDO + k;
            a.redo_tol = ba.guard_tol;
        }
        int rc = launch_ll_segment(N, precision, precision, var, ba.third, a, st);
        if (rc) return rc;
        if (guard) {
            WaveArgs r = a;
            r.redo_list = nullptr;
            r.redo_count = nullptr;
            r.order = ba.redo;
            r.seg_count = ba.ctl + CTL_REDO + k;
            r.tw = twiddles_f64(N);
            if (!r.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
            // the FP64 pass takes the FP64 tile of the bucket (tuned or default),
            // as the plain FP64 batch does: same speed, same results
            rc = launch_ll_segment(N, GBGPU_FP64, GBGPU_MIXED, var64, ba.third, r, st);
            if (rc) return rc;
        }
        if (route) {
            // the binaries routed to FP64 at set-up: FP64 tile, mixed constants
            WaveArgs r = a;
            r.redo_list = nullptr;
            r.redo_count = nullptr;
            r.seg_count = ba.ctl + CTL_COUNT + NBUCKET + k;
            r.seg_off = ba.ctl + CTL_LISTOFF + NBUCKET + k;
            r.tw = twiddles_f64(N);
            if (!r.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
            rc = launch_ll_segment(N, GBGPU_FP64, GBGPU_MIXED, var64, ba.third, r, st);
            if (rc) return rc;
        }
        if (bucket_events) (void)hipEventRecord(bucket_events[2 * k + 1], st);
    }
    return GBGPU_OK;
}

template <class P>
static int launch_batch(BatchArgs ba, int precision, int strategy, void *workspace, hipStream_t st,
                        const int *variants = nullptr, hipEvent_t *bucket_events = nullptr) {
    const long long B = ba.a.B;
    char *ws = static_cast<char *>(workspace);
    bool own = false;
    const size_t bytes = batch_ws_bytes<P>(B);
    if (!ws) {
        hipError_t e = hipMallocAsync((void **)&ws, bytes, st);
        if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(batch workspace)");
        own = true;
    }
    auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
    // layout (each piece 256-byte aligned): packs, list, keys, redo list,
    // indices, sorted keys, control block, sort storage
    const size_t ib = al(sizeof(int) * (size_t)B);
    char *p = ws;
    Pack<P> *pack = reinterpret_cast<Pack<P> *>(p);
    p += al(sizeof(Pack<P>) * (size_t)B);
    ba.list = reinterpret_cast<int *>(p);
    ba.key = reinterpret_cast<unsigned *>(p + ib);
    ba.redo = reinterpret_cast<int *>(p + 2 * ib);
    ba.idx = reinterpret_cast<int *>(p + 3 * ib);
    unsigned *key_sorted = reinterpret_cast<unsigned *>(p + 4 * ib);
    ba.ctl = reinterpret_cast<int *>(ws + batch_ctl_offset(sizeof(Pack<P>), B));
    void *sort_tmp = ws + batch_ctl_offset(sizeof(Pack<P>), B) + al(sizeof(int) * CTL_WORDS);
    ba.a.pack = pack;
    const unsigned gb = (unsigned)((B + 255) / 256);
    hipLaunchKernelGGL(k_batch_reset, dim3((CTL_WORDS + 255) / 256), dim3(256), 0, st, ba.ctl);
    hipLaunchKernelGGL((k_batch_setup<P>), dim3(gb), dim3(256), 0, st, ba, pack);
    // lsd_sort for every size: no memset, so the sequence stays capturable
    hipError_t e = lsd_sort<unsigned>(sort_tmp, ba.key, key_sorted, ba.idx, ba.list, (int)B, BATCH_KEY_BITS, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_batch_bounds, dim3(gb), dim3(256), 0, st, key_sorted, B, ba.ctl);
        hipLaunchKernelGGL(k_batch_table, dim3(1), dim3(64), 0, st, ba.ctl);
        e = hipGetLastError();
    }
    int rc = e == hipSuccess ? GBGPU_OK : hip_fail(e, "batch set-up");
    if (!rc) rc = strategy == GBGPU_BATCH_PERSISTENT ? launch_persistent<P>(ba, st)
                                                     : launch_buckets<P>(ba, precision, st, variants, bucket_events);
    hipError_t e2 = own ? hipFreeAsync(ws, st) : hipSuccess;
    if (rc) return rc;
    if (e2 != hipSuccess) return hip_fail(e2, "hipFreeAsync(batch workspace)");
    return GBGPU_OK;
}

// ----------------------------------------------------- parameter jumps ---
// Gaussian relative jumps out = in * (1 + sigma z), z ~ N(0, 1) from a
// counter-based hash (splitmix64 of seed, call, row, binary) and Box-Muller;
// the same (seed, call) always gives the same draw.  Optionally recomputes N
// per binary from f0 with a step rule (N_levels[i] below f_edges[i]).
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

struct NRule {
    int n_edges;
    double edges[8];
    int levels[9];
};

__global__ void k_perturb(double *out, const double *in, int num_params, long long B, double sigma,
                          unsigned long long seed, unsigned long long call, int *N_out, NRule rule) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    for (int r = 0; r < num_params; r++) {
        const unsigned long long h1 = splitmix64(seed ^ splitmix64(call * 0x100000001b3ULL + (unsigned long long)r) ^
                                                 splitmix64((unsigned long long)b * 0x2545f4914f6cdd1dULL));
        const unsigned long long h2 = splitmix64(h1);
        const double u1 = ((h1 >> 11) + 1) * (1.0 / 9007199254740993.0);  // (0, 1]
        const double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);        // [0, 1)
        const double z = sqrt(-2.0 * log(u1)) * cos(TWO_PI * u2);
        out[r * B + b] = in[r * B + b] * (1.0 + sigma * z);
    }
    if (N_out) {
        const double f0 = fabs(out[1 * B + b]);
        int i = 0;
        while (i < rule.n_edges && f0 >= rule.edges[i]) i++;
        N_out[b] = rule.levels[i];
    }
}

}  // namespace gbgpu

using namespace gbgpu;

extern "C" long long gbgpu_batch_workspace_bytes(int num_bin, int precision) {
    if (num_bin < 0) return -1;
    return precision == GBGPU_MIXED ? (long long)batch_ws_bytes<Mixed>(num_bin)
                                    : (long long)batch_ws_bytes<F64>(num_bin);
}

static int make_batch_args(BatchArgs &ba, double *loglike, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h,
                           const double *params, int num_params, int num_bin, const int *N_bin,
                           const gbgpu_batch_cfg *cfg, const gbgpu_data *data, double d_d, int phase_marginalize) {
    if (!cfg || !params || !N_bin || !loglike) {
        set_error("null cfg, params, N_bin or loglike");
        return GBGPU_E_ARG;
    }
    if (num_params != GBGPU_NPARAMS && num_params != GBGPU_NPARAMS_THIRD) {
        set_error("num_params must be 9 or 14");
        return GBGPU_E_ARG;
    }
    if (num_bin < 0 || !(cfg->T > 0) || !(cfg->armlength > 0) ||
        (cfg->precision != GBGPU_FP64 && cfg->precision != GBGPU_MIXED)) {
        set_error("bad num_bin, T, armlength or precision");
        return GBGPU_E_ARG;
    }
    if (cfg->strategy == GBGPU_BATCH_PERSISTENT) {
        // switched off: its kernel took a GPU memory fault in this build whose
        // cause has not been shown; the per-bucket launches cover every batch
        set_error("the persistent batch strategy is disabled in this build; use GBGPU_BATCH_BUCKETS");
        return GBGPU_E_ARG;
    }
    if (cfg->strategy != GBGPU_BATCH_BUCKETS) {
        set_error("unknown batch strategy");
        return GBGPU_E_ARG;
    }
    if (cfg->transform < GBGPU_TRANSFORM_LDS || cfg->transform > GBGPU_TRANSFORM_MFMA_BF16X6 ||
        (cfg->transform != GBGPU_TRANSFORM_LDS && cfg->precision != GBGPU_MIXED)) {
        set_error("transform: unknown, or a matrix-core transform without GBGPU_MIXED");
        return GBGPU_E_ARG;
    }
    int rc = check_data(data);
    if (rc) return rc;
    ba = BatchArgs{};
    WaveArgs &a = ba.a;
    a.params = params;
    a.num_params = num_params;
    a.B = num_bin;
    a.T = cfg->T;
    a.fstar_inv = 1.0 / (C_SI / (cfg->armlength * 2 * PI));
    a.armlength = cfg->armlength;
    a.tdi2 = cfg->tdi2;
    set_data(a, data);
    a.d_d = d_d;
    a.phase_marg = phase_marginalize;
    a.loglike = loglike;
    a.d_h = reinterpret_cast<double2 *>(d_h);
    a.h_h = reinterpret_cast<double2 *>(h_h);
    ba.Nb = N_bin;
    ba.transform = cfg->transform;
    ba.third = num_params == GBGPU_NPARAMS_THIRD;
    if (cfg->precision == GBGPU_MIXED && cfg->orbit_tables[3]) {
        // the N = 256 bucket may run a matrix-core transform (requested or tuned)
        a.dftc = dft256_table();
        if (!a.dftc) return hip_fail(hipErrorOutOfMemory, "transform constants");
    }
    for (int k = 0; k < NBUCKET; k++) {
        if (cfg->bucket_capacity[k] < 0) {
            set_error("bucket_capacity must be >= 0");
            return GBGPU_E_ARG;
        }
        ba.cap[k] = cfg->bucket_capacity[k];
    }
    ba.guard_tol = cfg->guard_tol;
    ba.route_tol = (cfg->precision == GBGPU_MIXED && cfg->guard_tol > 0.0) ? cfg->guard_tol : 0.0;
    for (int k = 0; k < NBUCKET; k++) {
        ba.tabs[k] = cfg->orbit_tables[k];
        ba.tws[k] = nullptr;
        if (ba.tabs[k]) {
            const int N = 32 << k;
            ba.tws[k] = cfg->precision == GBGPU_MIXED ? (const void *)twiddles_f32(N) : (const void *)twiddles_f64(N);
            if (!ba.tws[k]) return hip_fail(hipErrorOutOfMemory, "twiddles");
            // the guard's FP64 pass needs the FP64 table (created here, outside any capture)
            if (cfg->precision == GBGPU_MIXED && cfg->guard_tol > 0.0 && !twiddles_f64(N))
                return hip_fail(hipErrorOutOfMemory, "twiddles");
        }
    }
    return GBGPU_OK;
}

extern "C" int gbgpu_wave_ll_batched(double *loglike, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const double *params,
                                     int num_params, int num_bin, const int *N_bin, const gbgpu_batch_cfg *cfg,
                                     const gbgpu_data *data, double d_d, int phase_marginalize, void *workspace,
                                     void *stream) {
    BatchArgs ba;
    int rc = make_batch_args(ba, loglike, d_h, h_h, params, num_params, num_bin, N_bin, cfg, data, d_d,
                             phase_marginalize);
    if (rc) return rc;
    if (num_bin == 0) return GBGPU_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    return cfg->precision == GBGPU_MIXED ? launch_batch<Mixed>(ba, cfg->precision, cfg->strategy, workspace, st)
                                         : launch_batch<F64>(ba, cfg->precision, cfg->strategy, workspace, st);
}

extern "C" int gbgpu_batch_autotune(double *loglike, const double *params, int num_params, int num_bin,
                                    const int *N_bin, const gbgpu_batch_cfg *cfg, const gbgpu_data *data,
                                    double d_d, void *workspace, int *chosen, void *stream) {
    BatchArgs ba;
    int rc = make_batch_args(ba, loglike, nullptr, nullptr, params, num_params, num_bin, N_bin, cfg, data, d_d, 0);
    if (rc) return rc;
    if (cfg->strategy != GBGPU_BATCH_BUCKETS) {
        set_error("the autotuner tunes the per-bucket launches (GBGPU_BATCH_BUCKETS)");
        return GBGPU_E_ARG;
    }
    const int pi = cfg->precision == GBGPU_MIXED ? 1 : 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t ev[2 * NBUCKET];
    for (int i = 0; i < 2 * NBUCKET; i++)
        if (hipEventCreate(&ev[i]) != hipSuccess) return hip_fail(hipErrorOutOfMemory, "hipEventCreate");
    // candidates: the two tile shapes for every bucket; for the mixed N = 256
    // bucket also the FP32-accurate matrix-core transforms (FP32 and
    // three-piece bf16 products; the two-piece split is less accurate than the
    // LDS transform and is used only on request)
    constexpr int NVAR = 5, REPS = 3;
    auto allowed = [&](int v, int k) {
        return v < VARIANT_DFT0 ||
               (cfg->precision == GBGPU_MIXED && k == 3 && cfg->transform == GBGPU_TRANSFORM_LDS && !ba.third &&
                (v == VARIANT_DFT0 + GBGPU_TRANSFORM_MFMA_F32 - 1 || v == VARIANT_DFT0 + GBGPU_TRANSFORM_MFMA_BF16X6 - 1));
    };
    float best[NVAR][NBUCKET];
    for (int v = 0; v < NVAR; v++)
        for (int k = 0; k < NBUCKET; k++) best[v][k] = 3.0e38f;
    for (int v = 0; v < NVAR && !rc; v++) {
        int var[NBUCKET];
        bool any = false;
        for (int k = 0; k < NBUCKET; k++) {
            var[k] = allowed(v, k) ? v : 0;
            any = any || (v >= VARIANT_DFT0 && allowed(v, k) && ba.tabs[k]);
        }
        if (v >= VARIANT_DFT0 && !any) continue;
        for (int r = 0; r < REPS && !rc; r++) {  // the first repetition warms up
            rc = cfg->precision == GBGPU_MIXED
                     ? launch_batch<Mixed>(ba, cfg->precision, cfg->strategy, workspace, st, var, ev)
                     : launch_batch<F64>(ba, cfg->precision, cfg->strategy, workspace, st, var, ev);
            if (rc) break;
            hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                rc = hip_fail(e, "autotune");
                break;
            }
            for (int k = 0; k < NBUCKET && r > 0; k++) {
                if (!ba.tabs[k] || !allowed(v, k)) continue;
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) == hipSuccess && ms < best[v][k])
                    best[v][k] = ms;
            }
        }
    }
    for (int i = 0; i < 2 * NBUCKET; i++) (void)hipEventDestroy(ev[i]);
    if (rc) return rc;
    for (int k = 0; k < NBUCKET; k++) {
        int pick = -1;
        if (ba.tabs[k]) {
            pick = 0;
            for (int v = 1; v < NVAR; v++)
                if (best[v][k] < best[pick][k]) pick = v;
            g_variant[pi][k].store(pick);
        }
        if (chosen) chosen[k] = pick;
    }
    return GBGPU_OK;
}

extern "C" int gbgpu_batch_set_variants(int precision, const int *variants) {
    if ((precision != GBGPU_FP64 && precision != GBGPU_MIXED) || !variants) {
        set_error("gbgpu_batch_set_variants: bad arguments");
        return GBGPU_E_ARG;
    }
    for (int k = 0; k < NBUCKET; k++) {
        const bool dft = precision == GBGPU_MIXED && k == 3 && variants[k] >= VARIANT_DFT0 &&
                         variants[k] <= VARIANT_DFT0 + GBGPU_TRANSFORM_MFMA_BF16X6 - 1;
        if (variants[k] != 0 && variants[k] != 1 && !dft) {
            set_error("gbgpu_batch_set_variants: variants are 0 or 1, and 2-4 (matrix-core transforms) for the "
                      "mixed N = 256 bucket");
            return GBGPU_E_ARG;
        }
    }
    for (int k = 0; k < NBUCKET; k++) g_variant[precision == GBGPU_MIXED ? 1 : 0][k].store(variants[k]);
    return GBGPU_OK;
}

extern "C" int gbgpu_batch_guard_counts(const void *workspace, int num_bin, int precision, int *counts,
                                        void *stream) {
    if (!workspace || !counts || num_bin < 0 || (precision != GBGPU_FP64 && precision != GBGPU_MIXED)) {
        set_error("gbgpu_batch_guard_counts: bad arguments");
        return GBGPU_E_ARG;
    }
    const size_t pack = precision == GBGPU_MIXED ? sizeof(Pack<Mixed>) : sizeof(Pack<F64>);
    const char *ctl = static_cast<const char *>(workspace) + batch_ctl_offset(pack, num_bin);
    hipStream_t st = static_cast<hipStream_t>(stream);
    // binaries evaluated in FP64: re-evaluated after the mixed pass, plus
    // those routed to FP64 at set-up
    int redo[NBUCKET], routed[NBUCKET];
    hipError_t e = hipMemcpyAsync(redo, ctl + sizeof(int) * CTL_REDO, sizeof(int) * NBUCKET, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(routed, ctl + sizeof(int) * (CTL_COUNT + NBUCKET), sizeof(int) * NBUCKET,
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "gbgpu_batch_guard_counts");
    for (int k = 0; k < NBUCKET; k++) counts[k] = redo[k] + (precision == GBGPU_MIXED ? routed[k] : 0);
    return GBGPU_OK;
}

extern "C" int gbgpu_perturb_params(double *out, const double *in, int num_params, int num_bin, double sigma_rel,
                                    unsigned long long seed, unsigned long long call_index, int *N_out,
                                    int n_edges, const double *f_edges, const int *N_levels, void *stream) {
    if (!out || !in || num_params < 2 || num_bin < 0 || n_edges < 0 || n_edges > 8 ||
        (N_out && (!f_edges || !N_levels) && n_edges > 0)) {
        set_error("gbgpu_perturb_params: bad arguments");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    NRule rule = {};
    rule.n_edges = N_out ? n_edges : 0;
    for (int i = 0; i < rule.n_edges; i++) rule.edges[i] = f_edges[i];
    for (int i = 0; N_out && i <= rule.n_edges; i++) rule.levels[i] = N_levels[i];
    hipLaunchKernelGGL(k_perturb, dim3((unsigned)((num_bin + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), out, in, num_params, (long long)num_bin, sigma_rel, seed,
                       call_index, N_out, rule);
    GBGPU_CHECK_LAUNCH("k_perturb");
    return GBGPU_OK;
}
