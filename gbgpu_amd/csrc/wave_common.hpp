// Scaffolding shared by the FP64 and mixed-precision waveform kernels:
// launch geometry, kernel arguments and group-level scan/reduction.
#pragma once

#include "common.hpp"

namespace gbgpu {

enum { MODE_WAVE = 0, MODE_LL = 1 };

template <int N>
struct Geo {
    static constexpr int TPB = N / 4;
    static constexpr int WG = TPB >= 256 ? TPB : 256;
    static constexpr int BPW = WG / TPB;
    static constexpr int WPB = (TPB + 63) / 64;  // wavefronts per binary
};

struct WaveArgs {
    const double *params;
    int num_params;
    long long B;
    double T, scale, fstar_inv, armlength;
    int tdi2;
    const double *tab;
    const void *tw;
    // run_wave outputs
    double2 *XYZf, *AETf;
    long long *start_inds;
    // likelihood inputs / outputs
    const double2 *dA, *dE;
    const double *SA, *SE;
    const int *di, *ni;
    long long L, start_freq_ind;
    double df4, d_d;
    int phase_marg;
    double *loglike;
    double2 *d_h, *h_h;
};

// Exclusive prefix sum over the TPB lanes owning one binary.  `tot` is LDS
// scratch of WPB doubles for this binary.  Uniform control flow required.
template <int TPB>
__device__ __forceinline__ double group_excl_scan(double x, int lane, double *tot) {
    constexpr int W = TPB < 64 ? TPB : 64;
    const int wl = lane % W;
    double inc = x;
#pragma unroll
    for (int off = 1; off < W; off <<= 1) {
        const double y = __shfl_up(inc, off, W);
        if (wl >= off) inc += y;
    }
    if constexpr (TPB > 64) {
        const int w = lane / 64;
        if (wl == W - 1) tot[w] = inc;
        __syncthreads();
        double pre = 0.0;
        for (int i = 0; i < w; i++) pre += tot[i];
        __syncthreads();
        inc += pre;
    }
    return inc - x;
}

// Sum over the TPB lanes owning one binary; result valid in every lane.
template <int TPB>
__device__ __forceinline__ double group_total(double x, int lane, double *tot) {
    constexpr int W = TPB < 64 ? TPB : 64;
    x = group_sum<W>(x);
    if constexpr (TPB > 64) {
        if (lane % 64 == 0) tot[lane / 64] = x;
        __syncthreads();
        x = 0.0;
        for (int i = 0; i < TPB / 64; i++) x += tot[i];
        __syncthreads();
    }
    return x;
}

}  // namespace gbgpu
