// Batch segment kernels (gfx950): the fused likelihood over one device-side
// segment of an uneven-N batch, for the N-bucket launches of
// gbgpu_wave_ll_batched (batched.hip).  The templates are instantiated for
// circular binaries in wave_seg.hip and for third-body binaries in
// wave_seg3.hip (two translation units that compile in parallel).
#pragma once

#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "wave_common.hpp"
#include "wave_pass.hpp"

namespace gbgpu {

// Fused likelihood over one device-side segment of a batch: item i <
// *a.seg_count is binary a.order[*a.seg_off + i]; the grid strides over the
// workgroup passes blk0, blk0 + 1, ... of the segment.  The orbit-table and
// twiddle reads and the per-thread index arithmetic are the same in every
// pass: opaque copies of the pointers and of the thread index keep the
// compiler from hoisting all of them out of the loop into registers.  The
// loop still keeps more values live than one pass does (146 against 120
// VGPRs for the FP64 N = 256 tile), hence k_wave_seg1 below.
template <int N, bool THIRD, class P, class PK, int M, int DFT = GBGPU_TRANSFORM_LDS>
__global__ void __launch_bounds__((Geo<N, M>::WG)) k_wave_seg(WaveArgs a, long long blk0) {
    constexpr int BPW = Geo<N, M>::BPW;
    __shared__ WaveLds<N, MODE_LL, THIRD, P, PK, M, DFT> lds;
    const long long count = *a.seg_count;
    const long long first = *a.seg_off;
    for (long long blk = blk0 + blockIdx.x; blk * BPW < count; blk += gridDim.x) {
        const double *tab = a.tab;
        const void *twp = a.tw;
        int tid = threadIdx.x;
        asm volatile("" : "+s"(tab), "+s"(twp), "+v"(tid));
        wave_pass<N, MODE_LL, THIRD, P, PK, M, DFT>(a, lds, blk, count, first, tid, tab, twp);
        __syncthreads();  // LDS is reused by the next pass
    }
}

// One workgroup pass per workgroup: pass blockIdx.x of the segment, for a
// grid sized from a host-side bound on the segment (workgroups past the
// device-side count leave at once, about 0.2 ns each).
template <int N, bool THIRD, class P, class PK, int M, int DFT = GBGPU_TRANSFORM_LDS>
__global__ void __launch_bounds__((Geo<N, M>::WG)) k_wave_seg1(WaveArgs a) {
    constexpr int BPW = Geo<N, M>::BPW;
    __shared__ WaveLds<N, MODE_LL, THIRD, P, PK, M, DFT> lds;
    const long long count = *a.seg_count;
    const long long blk = blockIdx.x;
    if (blk * BPW >= count) return;  // uniform over the workgroup
    wave_pass<N, MODE_LL, THIRD, P, PK, M, DFT>(a, lds, blk, count, *a.seg_off, threadIdx.x, a.tab, a.tw);
}

// Likelihood kernel over one device-side segment of a batch (the N-bucket
// launches of gbgpu_wave_ll_batched): a.pack, a.order, a.seg_count and
// a.seg_off are set by the caller, and the segment's size is known only on
// the device.  With a host-side bound a.seg_cap, one k_wave_seg1 pass covers
// the first ceil(seg_cap / BPW) workgroup passes; the strided k_wave_seg
// (a few passes' worth of resident workgroups) covers the rest, and leaves at
// once when the bound held.  Without a bound the strided kernel does it all.
template <bool THIRD, class P, class PK, int NN, int M, int DFT = GBGPU_TRANSFORM_LDS>
int launch_segment_n(const WaveArgs &a, hipStream_t st) {
    constexpr long long BPW = Geo<NN, M>::BPW;
    // several passes' worth of resident workgroups (initialised once, thread-safe)
    static const int grid = [] {
        int dev = 0, ncu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        return 16 * (ncu > 0 ? ncu : 256);
    }();
    KernelTimer kt(st);
    const long long passes = (a.B + BPW - 1) / BPW;  // every binary of the batch
    long long done = 0;
    if (a.seg_cap > 0) {
        done = (std::min<long long>(a.seg_cap, a.B) + BPW - 1) / BPW;
        hipLaunchKernelGGL((k_wave_seg1<NN, THIRD, P, PK, M, DFT>), dim3((unsigned)done), dim3(Geo<NN, M>::WG), 0,
                           st, a);
        GBGPU_CHECK_LAUNCH("k_wave_seg1");
        if (done >= passes) return GBGPU_OK;
    }
    const long long g = std::min<long long>(passes - done, grid);
    hipLaunchKernelGGL((k_wave_seg<NN, THIRD, P, PK, M, DFT>), dim3((unsigned)g), dim3(Geo<NN, M>::WG), 0, st, a,
                       done);
    GBGPU_CHECK_LAUNCH("k_wave_seg");
    return GBGPU_OK;
}

// variant 0: the waveform kernels' tile (wave_tile); variant 1: half the
// threads per binary of that tile; VARIANT_DFT0.. (mixed N = 256 without a
// third body only): the matrix-core transforms, one wavefront per binary.
template <bool THIRD, class P, class PK, int NN>
int launch_segment_m(int variant, const WaveArgs &a, hipStream_t st) {
    constexpr int M0 = wave_tile<NN>();
    constexpr int M1 = M0 == 0 ? 1 : 2;
    if constexpr (NN == 256 && !THIRD && std::is_same<P, Mixed>::value && std::is_same<PK, Mixed>::value) {
        if (variant >= VARIANT_DFT0 && !a.dftc) {
            set_error("matrix-core transform requested without its constant table");
            return GBGPU_E_ARG;
        }
        switch (variant - VARIANT_DFT0 + 1) {
            case GBGPU_TRANSFORM_MFMA_F32:
                return launch_segment_n<THIRD, P, PK, NN, 1, GBGPU_TRANSFORM_MFMA_F32>(a, st);
            case GBGPU_TRANSFORM_MFMA_BF16X3:
                return launch_segment_n<THIRD, P, PK, NN, 1, GBGPU_TRANSFORM_MFMA_BF16X3>(a, st);
            case GBGPU_TRANSFORM_MFMA_BF16X6:
                return launch_segment_n<THIRD, P, PK, NN, 1, GBGPU_TRANSFORM_MFMA_BF16X6>(a, st);
            default:
                break;
        }
    }
    if constexpr (THIRD) {
        if (variant >= VARIANT_DFT0) {
            set_error("the matrix-core transforms do not take third-body binaries");
            return GBGPU_E_ARG;
        }
    }
    return variant == 1 ? launch_segment_n<THIRD, P, PK, NN, M1>(a, st)
                        : launch_segment_n<THIRD, P, PK, NN, M0>(a, st);
}

// precision: arithmetic of the pass; pack_precision: layout of a.pack;
// variant: tile shape (0: wave_tile, 1: half the threads per binary, or a
// matrix-core transform); N <= 2048.
template <bool THIRD>
int launch_ll_segment_t(int N, int precision, int pack_precision, int variant, const WaveArgs &a,
                        hipStream_t st) {
#define GBGPU_SEG(NN)                                                                                    \
    case NN:                                                                                             \
        if (precision == GBGPU_MIXED) return launch_segment_m<THIRD, Mixed, Mixed, NN>(variant, a, st);  \
        if (pack_precision == GBGPU_MIXED) return launch_segment_m<THIRD, F64, Mixed, NN>(variant, a, st); \
        return launch_segment_m<THIRD, F64, F64, NN>(variant, a, st);
    switch (N) {
        GBGPU_SEG(32)
        GBGPU_SEG(64)
        GBGPU_SEG(128)
        GBGPU_SEG(256)
        GBGPU_SEG(512)
        GBGPU_SEG(1024)
        GBGPU_SEG(2048)
        default:
            set_error("segment kernels take N in [32, 2048]");
            return GBGPU_E_ARG;
    }
#undef GBGPU_SEG
}

}  // namespace gbgpu
