// Large-N waveform kernels (gfx950): N = 4096 and 8192, one 1024-thread
// workgroup per binary, channels transformed in passes that fit in LDS.
#include <type_traits>

#include "common.hpp"
#include "dispatch.hpp"
#include "wave_common.hpp"
#include "wave_pass.hpp"

namespace gbgpu {

// ---- N = 4096 and 8192: one 1024-thread workgroup per binary --------------
//
// The spectra of all channels no longer fit in LDS at once (FP64 X, Y, Z at
// N = 8192 would take 384 KB), so the channels go through the transform in
// passes of as many as fit in BIG_LDS_BYTES; each pass recomputes the slow
// part (the per-sample work is shared by the channels, so a pass costs about
// one slow-part evaluation) and consumes its channels before the next:
//   MODE_LL:   each channel's <d|x> and <x|x> terms are added as it leaves
//              the transform;
//   MODE_AE:   each channel's spectrum is written;
//   MODE_WAVE: X and Y are written to the output (XYZf, or AETf as a stash
//              when only AET is wanted) and, in the pass that holds Z, read
//              back by the same thread at the same positions to form
//              X, Y, Z and A, E, T.
// One workgroup fills a CU's LDS budget (1 workgroup = 16 wavefronts per CU).
constexpr size_t BIG_LDS_BYTES = (size_t)128 << 10;

template <int N, int NCH, class CT>
constexpr int big_channels_per_pass() {
    constexpr int fit = (int)(BIG_LDS_BYTES / (sizeof(CT) * (size_t)N));
    return fit >= NCH ? NCH : fit;
}

// LDS of one large-N workgroup.
template <int N, int NCH, bool THIRD, class P>
struct BigLds {
    using G = Geo<N, 0>;
    using CT = typename P::CT;
    static constexpr int CPP = big_channels_per_pass<N, NCH, CT>();
    static_assert(CPP >= 1, "one channel must fit in LDS");
    CT sbuf[CPP * N];
    double stot[G::WPB];
    double s_edge[THIRD ? 6 * G::WPB : 1];
};

// One binary b on the whole workgroup.  PK: layout of the per-binary
// constants (the FP64 passes of a guarded mixed batch read mixed ones).
// `first`: offset of the segment in redo_list (guarded mixed likelihood).
template <int N, int MODE, bool THIRD, class P, class PK>
__device__ __forceinline__ void big_binary(const WaveArgs &a, long long b, long long first,
                                           BigLds<N, MODE == MODE_WAVE ? 3 : 2, THIRD, P> &L) {
    using G = Geo<N, 0>;
    using CT = typename P::CT;
    constexpr int TPB = G::TPB, SPT = G::SPT;
    static_assert(G::BPW == 1 && TPB == WG_MAX, "one 1024-thread workgroup per binary");
    constexpr int NCH = (MODE == MODE_WAVE) ? 3 : 2;
    constexpr int CPP = BigLds<N, NCH, THIRD, P>::CPP;
    CT *sbuf = L.sbuf;

    const int lane = threadIdx.x;
    const Pack<PK> *pk = static_cast<const Pack<PK> *>(a.pack) + b;
    const Src s = to_sgpr(pk->s);
    const typename P::Aux aux = to_sgpr(aux_of<P, PK>(*pk));
    const double2 fac = P::fac(s);
    ThirdScratch ts{L.s_edge};

    // MODE_LL: sums over both channels of conj(d) x / S and |x|^2 / S
    double xr = 0.0, xi = 0.0, xx = 0.0;
    const long long j0 = s.start - a.start_freq_ind;
    const long long doff = (a.di ? (long long)a.di[b] : 0LL) * a.L;
    const long long noff = (a.ni ? (long long)a.ni[b] : 0LL) * a.L;
    double2 *stash = MODE == MODE_WAVE ? (a.XYZf ? a.XYZf + b * 3 * N : (a.AETf ? a.AETf + b * 3 * N : nullptr))
                                       : nullptr;

    auto pass = [&](auto c0_c) {
        constexpr int C0 = decltype(c0_c)::value;
        constexpr int NCP = (NCH - C0) < CPP ? (NCH - C0) : CPP;
        binary_spectrum<N, NCH, THIRD, P, 0, GBGPU_TRANSFORM_LDS, C0, NCP>(
            s, aux, true, lane, a.tab, a.fstar_inv, a.scale, sbuf, static_cast<const CT *>(a.tw), ts, L.stot);
        // value of channel c (in this pass) at fftshifted output position m
        auto spec = [&](int c, int m) { return cmul(P::up(sbuf[(c - C0) * N + ((m + N / 2) & (N - 1))]), fac); };
        if constexpr (MODE == MODE_LL) {
            if constexpr (C0 == 0) ll_channel<N, P, 0, 0>(sbuf, lane, j0, doff, noff, a, xr, xi, xx);
            if constexpr (C0 <= 1 && C0 + NCP > 1)
                ll_channel<N, P, 1, 0>(sbuf + (1 - C0) * N, lane, j0, doff, noff, a, xr, xi, xx);
        } else if constexpr (MODE == MODE_AE) {
            double2 *ae = a.AETf + b * 2 * N;
#pragma unroll
            for (int q = 0; q < SPT; q++) {
                const int m = lane + q * TPB;
#pragma unroll
                for (int c = C0; c < C0 + NCP; c++) ae[c * N + m] = spec(c, m);
            }
        } else {
            if (stash) {
#pragma unroll
                for (int q = 0; q < SPT; q++) {
                    const int m = lane + q * TPB;
                    if constexpr (C0 + NCP < NCH) {
                        // X and/or Y: to the stash, read back in Z's pass
#pragma unroll
                        for (int c = C0; c < C0 + NCP; c++) stash[c * N + m] = spec(c, m);
                    } else {
                        double2 v[3];
#pragma unroll
                        for (int c = 0; c < 3; c++) v[c] = c >= C0 ? spec(c, m) : stash[c * N + m];
                        if (a.XYZf) {
                            double2 *xyz = a.XYZf + b * 3 * N;
#pragma unroll
                            for (int c = C0; c < 3; c++) xyz[c * N + m] = v[c];
                        }
                        if (a.AETf) {
                            double2 *aet = a.AETf + b * 3 * N;
                            aet[m] = cscale(csub(v[2], v[0]), INV_SQRT2);
                            aet[N + m] = cscale(cadd(csub(v[0], cscale(v[1], 2.0)), v[2]), INV_SQRT6);
                            if constexpr (std::is_same<P, Mixed>::value)
                                aet[2 * N + m] = make_double2(__builtin_nan(""), __builtin_nan(""));
                            else
                                aet[2 * N + m] = cscale(cadd(cadd(v[0], v[1]), v[2]), INV_SQRT3);
                        }
                    }
                }
            }
        }
        __syncthreads();  // LDS is reused by the next pass
    };
    pass(std::integral_constant<int, 0>{});
    if constexpr (CPP < NCH) pass(std::integral_constant<int, CPP>{});
    if constexpr (2 * CPP < NCH) pass(std::integral_constant<int, 2 * CPP>{});

    if constexpr (MODE == MODE_LL) {
        double dhr = fac.x * xr - fac.y * xi;
        double dhi = fac.x * xi + fac.y * xr;
        double hh = (fac.x * fac.x + fac.y * fac.y) * xx;
        dhr = group_total<TPB>(dhr, lane, L.stot);
        dhi = group_total<TPB>(dhi, lane, L.stot);
        hh = group_total<TPB>(hh, lane, L.stot);
        if (lane == 0) {
            dhr *= a.df4;
            dhi *= a.df4;
            hh *= a.df4;
            if (a.d_h) a.d_h[b] = make_double2(dhr, dhi);
            if (a.h_h) a.h_h[b] = make_double2(hh, 0.0);
            const double dh_use = a.phase_marg ? sqrt(dhr * dhr + dhi * dhi) : dhr;
            a.loglike[b] = -0.5 * (a.d_d + hh - 2.0 * dh_use);
            if (a.redo_list && GUARD_REL * (fabs(hh) + sqrt(dhr * dhr + dhi * dhi)) > a.redo_tol) {
                const int pos = atomicAdd(a.redo_count, 1);
                a.redo_list[first + pos] = (int)b;
            }
        }
    } else {
        if (lane == 0) a.start_inds[b] = s.start;
    }
}

template <int N, int MODE, bool THIRD, class P>
__global__ void __launch_bounds__(WG_MAX) k_wave_big(WaveArgs a) {
    __shared__ BigLds<N, MODE == MODE_WAVE ? 3 : 2, THIRD, P> lds;
    const long long item = blockIdx.x;
    const long long b = a.order ? (long long)a.order[item] : item;
    big_binary<N, MODE, THIRD, P, P>(a, b, 0, lds);
}

// Fused likelihood over one device-side segment of an uneven-N batch (item
// i < *a.seg_count is binary a.order[*a.seg_off + i]); the grid strides over
// the segment, one binary per workgroup pass.
template <int N, bool THIRD, class P, class PK>
__global__ void __launch_bounds__(WG_MAX) k_wave_big_seg(WaveArgs a) {
    __shared__ BigLds<N, 2, THIRD, P> lds;
    const long long count = *a.seg_count;
    const long long first = *a.seg_off;
    for (long long item = blockIdx.x; item < count; item += gridDim.x) {
        const long long b = a.order[first + item];
        big_binary<N, MODE_LL, THIRD, P, PK>(a, b, first, lds);
    }
}

template <int MODE, bool THIRD, class P>
static int launch_big(int N, const WaveArgs &a, hipStream_t st, char *ws, size_t ws_bytes) {
    auto run = [&](auto nn) -> int {
        constexpr int NN = decltype(nn)::value;
        return with_pack<P>(
            a, NN, st,
            [&](const WaveArgs &aa) -> int {
                KernelTimer kt(st);
                hipLaunchKernelGGL((k_wave_big<NN, MODE, THIRD, P>), dim3((unsigned)aa.B), dim3(WG_MAX), 0, st, aa);
                GBGPU_CHECK_LAUNCH("k_wave_big");
                return GBGPU_OK;
            },
            ws, ws_bytes);
    };
    switch (N) {
        case 4096: return run(std::integral_constant<int, 4096>{});
        case 8192: return run(std::integral_constant<int, 8192>{});
        default:
            set_error("large-N kernels take N = 4096 or 8192");
            return GBGPU_E_ARG;
    }
}

int launch_wave_big(int N, int mode, bool third, int precision, const WaveArgs &a, hipStream_t st, char *ws,
                    size_t ws_bytes) {
#define GBGPU_BIG(MODE)                                                                              \
    if (mode == MODE) {                                                                              \
        if (precision == GBGPU_MIXED)                                                                \
            return third ? launch_big<MODE, true, Mixed>(N, a, st, ws, ws_bytes)                     \
                         : launch_big<MODE, false, Mixed>(N, a, st, ws, ws_bytes);                   \
        return third ? launch_big<MODE, true, F64>(N, a, st, ws, ws_bytes)                           \
                     : launch_big<MODE, false, F64>(N, a, st, ws, ws_bytes);                         \
    }
    GBGPU_BIG(MODE_WAVE)
    GBGPU_BIG(MODE_LL)
    GBGPU_BIG(MODE_AE)
#undef GBGPU_BIG
    set_error("unknown waveform mode");
    return GBGPU_E_ARG;
}

template <bool THIRD, class P, class PK>
static int big_segment(int N, const WaveArgs &a, hipStream_t st) {
    // the segment's size is known only on the device: a few resident
    // workgroups per CU (one fits at a time), striding
    static const int grid = [] {
        int dev = 0, ncu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        return 4 * (ncu > 0 ? ncu : 256);
    }();
    const long long g = a.B < grid ? a.B : grid;
    KernelTimer kt(st);
    if (N == 4096)
        hipLaunchKernelGGL((k_wave_big_seg<4096, THIRD, P, PK>), dim3((unsigned)g), dim3(WG_MAX), 0, st, a);
    else
        hipLaunchKernelGGL((k_wave_big_seg<8192, THIRD, P, PK>), dim3((unsigned)g), dim3(WG_MAX), 0, st, a);
    GBGPU_CHECK_LAUNCH("k_wave_big_seg");
    return GBGPU_OK;
}

int launch_big_segment(int N, int precision, int pack_precision, bool third, const WaveArgs &a, hipStream_t st) {
    if (N != 4096 && N != 8192) {
        set_error("large-N segments take N = 4096 or 8192");
        return GBGPU_E_ARG;
    }
    if (precision == GBGPU_MIXED)
        return third ? big_segment<true, Mixed, Mixed>(N, a, st) : big_segment<false, Mixed, Mixed>(N, a, st);
    if (pack_precision == GBGPU_MIXED)
        return third ? big_segment<true, F64, Mixed>(N, a, st) : big_segment<false, F64, Mixed>(N, a, st);
    return third ? big_segment<true, F64, F64>(N, a, st) : big_segment<false, F64, F64>(N, a, st);
}

}  // namespace gbgpu
