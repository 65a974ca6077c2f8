// Length-N forward DFT (numpy sign convention, unnormalised) in LDS.
//
// Stockham autosort radix-4 stages followed, when log2 N is odd, by one
// in-place radix-2 stage.  A group of TPB threads (TPB <= N/4) owns one
// binary: each thread performs N/(4 TPB) radix-4 butterflies per channel per
// stage, holding their inputs in registers across the barrier that separates
// the reads from the writes, so a single LDS buffer per channel suffices.
// Output is in natural order.  Every thread of the workgroup must call this
// (it contains __syncthreads); all binaries of a workgroup share N.
#pragma once

#include "common.hpp"

namespace gbgpu {

template <class CT>
__device__ __forceinline__ void bfly4(CT &v0, CT &v1, CT &v2, CT &v3) {
    const CT a0 = cadd(v0, v2), a1 = csub(v0, v2);
    const CT a2 = cadd(v1, v3), a3 = csub(v1, v3);
    v0 = cadd(a0, a2);           // x0 + x1 + x2 + x3
    v2 = csub(a0, a2);           // x0 - x1 + x2 - x3
    v1 = cadd(a1, cmul_mi(a3));  // x0 - i x1 - x2 + i x3
    v3 = csub(a1, cmul_mi(a3));  // x0 + i x1 - x2 - i x3
}

// buf: NCH channels of N contiguous elements for this thread's binary.
// tw:  w[k] = exp(-2 pi i k / N), k < N.
template <int N, int NCH, int TPB, class CT>
__device__ __forceinline__ void fft_lds(CT *__restrict__ buf, int lane, const CT *__restrict__ tw) {
    constexpr int Q = N / 4;
    constexpr int R = Q / TPB;  // butterflies per thread per stage
    static_assert(R >= 1 && R * TPB == Q, "TPB must divide N/4");
#pragma unroll
    for (int Ns = 1; Ns * 4 <= N; Ns *= 4) {
        const int step = N / (4 * Ns);
        CT v[R][NCH][4];
#pragma unroll
        for (int rr = 0; rr < R; rr++)
#pragma unroll
            for (int c = 0; c < NCH; c++)
#pragma unroll
                for (int r = 0; r < 4; r++) v[rr][c][r] = buf[c * N + lane + rr * TPB + r * Q];
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < R; rr++) {
            const int j = lane + rr * TPB;
            const int k = j & (Ns - 1);
            CT w1{}, w2{}, w3{};
            if (Ns > 1) {
                w1 = tw[k * step];
                w2 = tw[2 * k * step];
                w3 = tw[3 * k * step];
            }
            const int base = ((j - k) << 2) + k;
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                CT x0 = v[rr][c][0], x1 = v[rr][c][1], x2 = v[rr][c][2], x3 = v[rr][c][3];
                if (Ns > 1) {  // the first stage's twiddles are all w^0 = 1
                    x1 = cmul(x1, w1);
                    x2 = cmul(x2, w2);
                    x3 = cmul(x3, w3);
                }
                bfly4(x0, x1, x2, x3);
                buf[c * N + base] = x0;
                buf[c * N + base + Ns] = x1;
                buf[c * N + base + 2 * Ns] = x2;
                buf[c * N + base + 3 * Ns] = x3;
            }
        }
        __syncthreads();
    }
    // log2 N odd: one radix-2 stage at Ns = N/2, in place
    constexpr bool odd = (N & 0x55555555) == 0;  // N is 2 * 4^m
    if (odd) {
#pragma unroll
        for (int h = 0; h < 2 * R; h++) {
            const int j = lane + h * TPB;
            const CT w = tw[j];
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const CT a0 = buf[c * N + j];
                const CT a1 = cmul(buf[c * N + j + N / 2], w);
                buf[c * N + j] = cadd(a0, a1);
                buf[c * N + j + N / 2] = csub(a0, a1);
            }
        }
        __syncthreads();
    }
}

}  // namespace gbgpu
