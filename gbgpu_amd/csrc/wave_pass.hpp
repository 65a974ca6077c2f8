// One pass of a workgroup of the uniform waveform kernels (k_wave) and the
// batch segment kernels (k_wave_seg): per-binary constants, the spectrum of
// each binary of the workgroup, and the epilogue of the kernel's mode.
#pragma once

#include <type_traits>

#include "common.hpp"
#include "wave_common.hpp"

namespace gbgpu {

// LDS of one workgroup of k_wave.  PK is the precision policy whose Pack
// layout the per-binary constants were written in (normally P; the FP64
// re-evaluation of guarded mixed binaries reads the mixed packs).
// Arrays a variant does not use have length zero, so that the transform
// buffers alone set the footprint (32 KB for the FP64 likelihood at N = 256:
// five workgroups per CU).
// With the matrix-core transform (DFT != 0) the samples and spectra stay in
// registers and the workgroup needs no LDS.
template <int N, int MODE, bool THIRD, class P, class PK = P, int M = 1, int DFT = GBGPU_TRANSFORM_LDS>
struct WaveLds {
    using G = Geo<N, M>;
    using CT = typename P::CT;
    static constexpr int NCH = (MODE == MODE_WAVE) ? 3 : 2;
    // a wavefront that owns its binary reads the Pack straight into scalar registers
    static constexpr bool PACK_IN_LDS = G::TPB < 64;
    static constexpr int NEDGE = (THIRD && G::WPB > 1) ? 6 * G::WPB : 0;
    CT sbuf[DFT == GBGPU_TRANSFORM_LDS ? G::BPW * NCH * N : 0];
    Pack<PK> spack[PACK_IN_LDS ? G::BPW : 0];
    double stot[G::WPB > 1 ? G::BPW * G::WPB : 0];
    double s_edge[G::BPW * NEDGE];
};

// One pass of a workgroup: items blk*BPW .. blk*BPW + BPW-1 of the
// `count` items (item i is binary order[first + i], or i without an order).
template <class P, class PK>
__device__ __forceinline__ typename P::Aux aux_of(const Pack<PK> &pk) {
    if constexpr (std::is_same<P, PK>::value)
        return pk.aux;
    else
        return P::make_aux(pk.s);
}

template <int N, int MODE, bool THIRD, class P, class PK = P, int M = 1, int DFT = GBGPU_TRANSFORM_LDS,
          bool EXT = false>
__device__ __forceinline__ void wave_pass(const WaveArgs &a, WaveLds<N, MODE, THIRD, P, PK, M, DFT> &L,
                                          long long blk, long long count, long long first, int tid,
                                          const double *tab, const void *twp) {
    using G = Geo<N, M>;
    using W = WaveLds<N, MODE, THIRD, P, PK, M, DFT>;
    using CT = typename P::CT;
    constexpr int TPB = G::TPB, BPW = G::BPW, WPB = G::WPB, SPT = G::SPT;
    constexpr int NCH = W::NCH;
    CT *sbuf = L.sbuf;
    Pack<PK> *spack = L.spack;
    double *stot = L.stot;
    const int slot = per_group<TPB>(tid / TPB);
    const int lane = tid % TPB;
    const long long item = blk * BPW + slot;
    const bool valid = item < count;
    const long long b = (valid && a.order) ? (long long)a.order[first + item] : item;

    const Pack<PK> *pk;
    if constexpr (W::PACK_IN_LDS) {
        if (valid) load_pack<PK, TPB>(a, b, lane, spack + slot);
        __syncthreads();
        pk = spack + slot;
    } else {
        // wave-uniform index; an empty slot reads binary 0 and discards it
        pk = static_cast<const Pack<PK> *>(a.pack) + (valid ? b : 0);
    }
    const Src s = per_group<TPB>(pk->s);
    CT *buf = sbuf + slot * NCH * N;
    double *tot = stot + slot * WPB;
    ThirdScratch ts{L.s_edge + slot * W::NEDGE};

    // spectra in registers: the FP32/bf16 matrix-core transforms' layout
    // (MFMA), or bin lane + q TPB in register q (the LDS transform with
    // register ends for circular binaries, and the FP64 matrix-core transform)
    constexpr bool MFMA = DFT != GBGPU_TRANSFORM_LDS && DFT != GBGPU_TRANSFORM_MFMA_F64;
    constexpr bool REG = DFT != GBGPU_TRANSFORM_LDS || !THIRD;
    CT xr[REG ? NCH * SPT : 1];
    binary_spectrum<N, NCH, THIRD, P, M, DFT, 0, NCH, EXT, DFT == GBGPU_TRANSFORM_LDS && REG>(
        s, per_group<TPB>(aux_of<P, PK>(*pk)), valid, lane, tab, a.fstar_inv, a.scale, buf,
        static_cast<const CT *>(twp), ts, tot, xr, a.dftc, (EXT && valid) ? a.ext + b * 6 * N : nullptr);
    // output q of this thread: fftshifted position pos(q) (natural bin k sits
    // at position (k + N/2) mod N) and its value spec(c, q) in channel c.
    // Positions lane + q TPB keep the data reads at constant offsets; with
    // the register-ended LDS transform and the FP64 matrix-core transform
    // their bins (lane + q TPB + N/2) mod N are registers (q + SPT/2) mod SPT.
    auto pos = [&](int q) {
        if constexpr (MFMA)
            return (dft256_index(lane, q) + N / 2) & (N - 1);
        else
            return lane + q * TPB;
    };
    auto spec = [&](int c, int q) {
        if constexpr (MFMA)
            return xr[c * SPT + q];
        else if constexpr (REG)
            return xr[c * SPT + ((q + SPT / 2) & (SPT - 1))];
        else
            return buf[c * N + ((lane + q * TPB + N / 2) & (N - 1))];  // fftshift
    };

    const double2 fac = P::fac(s);
    if constexpr (MODE == MODE_WAVE) {
        if (valid) {
            double2 *xyz = a.XYZf ? a.XYZf + b * 3 * N : nullptr;
            double2 *aet = a.AETf ? a.AETf + b * 3 * N : nullptr;
#pragma unroll
            for (int q = 0; q < SPT; q++) {
                const int m = pos(q);
                const double2 X = cmul(P::up(spec(0, q)), fac);
                const double2 Y = cmul(P::up(spec(1, q)), fac);
                const double2 Z = cmul(P::up(spec(2, q)), fac);
                if (xyz) {
                    xyz[m] = X;
                    xyz[N + m] = Y;
                    xyz[2 * N + m] = Z;
                }
                if (aet) {
                    aet[m] = cscale(csub(Z, X), INV_SQRT2);
                    aet[N + m] = cscale(cadd(csub(X, cscale(Y, 2.0)), Z), INV_SQRT6);
                    // T = (X + Y + Z)/sqrt3 cancels to O((f/f*)^2) of X: from FP32
                    // spectra it would carry errors far above its own size, so the
                    // mixed policy marks it not-a-number rather than return it
                    if constexpr (std::is_same<P, Mixed>::value)
                        aet[2 * N + m] = make_double2(__builtin_nan(""), __builtin_nan(""));
                    else
                        aet[2 * N + m] = cscale(cadd(cadd(X, Y), Z), INV_SQRT3);
                }
            }
            if (lane == 0) a.start_inds[b] = s.start;
        }
    } else if constexpr (MODE == MODE_AE) {
        if (valid) {
            double2 *ae = a.AETf + b * 2 * N;
#pragma unroll
            for (int q = 0; q < SPT; q++) {
                const int m = pos(q);
                ae[m] = cmul(P::up(spec(0, q)), fac);
                ae[N + m] = cmul(P::up(spec(1, q)), fac);
            }
            if (lane == 0) a.start_inds[b] = s.start;
        }
    } else {
        double dhr = 0.0, dhi = 0.0, hh = 0.0;
        if (valid) {
            const long long doff = (a.di ? (long long)a.di[b] : 0LL) * a.L;
            const long long noff = (a.ni ? (long long)a.ni[b] : 0LL) * a.L;
            if constexpr (REG)
                ll_partial_regs<P, SPT>(spec, pos, fac, s.start - a.start_freq_ind, doff, noff, a, dhr, dhi, hh);
            else
                ll_partial<N, P, M>(buf, fac, lane, s.start - a.start_freq_ind, doff, noff, a, dhr, dhi, hh);
        }
        dhr = group_total<TPB>(dhr, lane, tot);
        dhi = group_total<TPB>(dhi, lane, tot);
        hh = group_total<TPB>(hh, lane, tot);
        if (valid && lane == 0) {
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
    }
}

}  // namespace gbgpu
