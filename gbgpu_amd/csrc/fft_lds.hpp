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

// The same transform with both ends in registers, for a thread whose
// samples are the inputs of its own first-stage butterflies and whose
// outputs are read by that thread alone.  With SPT = N / TPB:
//   on entry x[c * SPT + q] = sample lane + q TPB of channel c;
//   on return x[c * SPT + p] = bin lane + p TPB of channel c.
// The first radix-4 stage (all twiddles 1) runs on the registers before the
// first LDS write, and the last stage (radix-4 when log2 N is even, radix-2
// when it is odd) reads LDS and leaves its results in registers: two LDS
// round trips and three barriers fewer than fft_lds.  Starts with a barrier,
// so buf may still be in use by earlier readers when it is called.
template <int N, int NCH, int TPB, class CT>
__device__ __forceinline__ void fft_lds_regs(CT *x, CT *__restrict__ buf, int lane, const CT *__restrict__ tw) {
    constexpr int Q = N / 4;
    constexpr int R = Q / TPB;  // butterflies per thread per radix-4 stage
    constexpr int SPT = 4 * R;
    static_assert(R >= 1 && R * TPB == Q, "TPB must divide N/4");
    constexpr bool odd = (N & 0x55555555) == 0;     // N is 2 * 4^m
    constexpr int NS_LAST = odd ? N / 8 : N / 4;    // last radix-4 stage
    static_assert(NS_LAST >= 4, "needs at least two radix-4 stages (N >= 16)");
    __syncthreads();
    // stage Ns = 1: butterfly j = lane + rr TPB takes samples j + r Q, i.e.
    // register q = rr + r R, and writes bins 4 j + r
#pragma unroll
    for (int rr = 0; rr < R; rr++) {
        const int j = lane + rr * TPB;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            CT x0 = x[c * SPT + rr], x1 = x[c * SPT + rr + R], x2 = x[c * SPT + rr + 2 * R],
               x3 = x[c * SPT + rr + 3 * R];
            bfly4(x0, x1, x2, x3);
            buf[c * N + 4 * j] = x0;
            buf[c * N + 4 * j + 1] = x1;
            buf[c * N + 4 * j + 2] = x2;
            buf[c * N + 4 * j + 3] = x3;
        }
    }
    __syncthreads();
    // stages Ns = 4 .. NS_LAST, through LDS; the last one keeps its results
    // in registers when it ends the transform
#pragma unroll
    for (int Ns = 4; Ns <= NS_LAST; Ns *= 4) {
        const int step = N / (4 * Ns);
        const bool to_regs = !odd && Ns == NS_LAST;
        CT v[R][NCH][4];
#pragma unroll
        for (int rr = 0; rr < R; rr++)
#pragma unroll
            for (int c = 0; c < NCH; c++)
#pragma unroll
                for (int r = 0; r < 4; r++) v[rr][c][r] = buf[c * N + lane + rr * TPB + r * Q];
        if (!to_regs) __syncthreads();
#pragma unroll
        for (int rr = 0; rr < R; rr++) {
            const int j = lane + rr * TPB;
            const int k = j & (Ns - 1);
            const CT w1 = tw[k * step], w2 = tw[2 * k * step], w3 = tw[3 * k * step];
            const int base = ((j - k) << 2) + k;
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                CT x0 = v[rr][c][0];
                CT x1 = cmul(v[rr][c][1], w1), x2 = cmul(v[rr][c][2], w2), x3 = cmul(v[rr][c][3], w3);
                bfly4(x0, x1, x2, x3);
                if (to_regs) {
                    // Ns = N/4: j < Ns, so base = j and bin j + r N/4 is register rr + r R
                    x[c * SPT + rr] = x0;
                    x[c * SPT + rr + R] = x1;
                    x[c * SPT + rr + 2 * R] = x2;
                    x[c * SPT + rr + 3 * R] = x3;
                } else {
                    buf[c * N + base] = x0;
                    buf[c * N + base + Ns] = x1;
                    buf[c * N + base + 2 * Ns] = x2;
                    buf[c * N + base + 3 * Ns] = x3;
                }
            }
        }
        if (!to_regs) __syncthreads();
    }
    if constexpr (odd) {
        // radix-2 stage at Ns = N/2: bins j and j + N/2 for j = lane + h TPB,
        // registers h and h + 2R
#pragma unroll
        for (int h = 0; h < 2 * R; h++) {
            const int j = lane + h * TPB;
            const CT w = tw[j];
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const CT a0 = buf[c * N + j];
                const CT a1 = cmul(buf[c * N + j + N / 2], w);
                x[c * SPT + h] = cadd(a0, a1);
                x[c * SPT + h + 2 * R] = csub(a0, a1);
            }
        }
    }
}

}  // namespace gbgpu
