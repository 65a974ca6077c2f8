// Scaffolding shared by the waveform kernels: launch geometry, kernel
// arguments, group-level scan/reduction, and the per-binary spectrum routine
// (third-body pre-pass -> slow part -> LDS transform).
#pragma once

#include <type_traits>

#include "common.hpp"
#include "dft_mfma.hpp"
#include "fft_lds.hpp"
#include "prec.hpp"
#include "source.hpp"

namespace gbgpu {

// MODE_WAVE: X, Y, Z (and A, E, T) spectra to HBM; MODE_LL: fused likelihood;
// MODE_AE: the A and E spectra only (two channels through the transform).
enum { MODE_WAVE = 0, MODE_LL = 1, MODE_AE = 2 };

constexpr int WG_SIZE = 256;

// A group of TPB threads owns one binary; each thread holds SPT = N/TPB
// samples.  Tile shapes:
//   M = 1: TPB = min(N/4, 256) in 256-thread workgroups (SPT = 4, 8 at N = 2048);
//   M = 2: half the threads per binary, twice the binaries per workgroup;
//   M = 0 ("wide"): TPB = min(N/4, 1024) in workgroups of max(TPB, 256)
//          threads -- at N = 2048 a 512-thread workgroup per binary, twice
//          the wavefronts per byte of LDS of M = 1; at N = 4096 and 8192
//          one 1024-thread workgroup per binary (SPT = 4 and 8).
// Workgroups hold BPW binaries.
constexpr int WG_MAX = 1024;
template <int N, int M = 1>
struct Geo {
    static constexpr int TPB = M == 0 ? (N / 4 > WG_MAX ? WG_MAX : N / 4) : ((N / 4 < WG_SIZE) ? N / 4 : WG_SIZE) / M;
    static constexpr int WG = TPB > WG_SIZE ? TPB : WG_SIZE;
    static constexpr int BPW = WG / TPB;
    static constexpr int SPT = N / TPB;
    static constexpr int WPB = (TPB + 63) / 64;  // wavefronts per binary
};

// Tile shape of the waveform kernels: wide where N/4 exceeds 256 threads.
template <int N>
constexpr int wave_tile() {
    return N / 4 > WG_SIZE ? 0 : 1;
}

// Per-bucket kernel variants of the batched path: 0 and 1 are tile shapes;
// VARIANT_DFT0 + t - 1 runs the mixed N = 256 bucket with matrix-core
// transform t (GBGPU_TRANSFORM_MFMA_*).
constexpr int VARIANT_DFT0 = 2;

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
    const void *pack;  // per-binary constants written by k_setup (Pack<P>[B])
    const int *order;  // processing order (binaries sorted by start bin) or null
    // Segment of `order` to process, read on the device: item i < *seg_count
    // is binary order[*seg_off + i].  Null: items 0..B-1 (one grid pass).
    const int *seg_count;
    const int *seg_off;
    // Host-side bound on *seg_count, or 0 when none is known.  With a bound
    // the segment runs as one pass of workgroups sized for it (no loop, so
    // the kernel keeps the uniform kernel's registers) followed by the
    // strided kernel, which takes any items beyond the bound.
    long long seg_cap;
    // Accuracy guard of the mixed likelihood: binaries whose error bound
    // GUARD_REL * (<h|h> + |<d|h>|) exceeds redo_tol are appended to
    // redo_list[*seg_off + i] (i from atomicAdd(redo_count)) for an FP64 pass.
    int *redo_list;
    int *redo_count;
    double redo_tol;
    unsigned *keys;    // k_setup output: start-bin sort keys, or null
    int *vals;         // k_setup output: binary indices, or null
    const void *dftc;  // per-lane constants of the matrix-core transform (dft256_table)
    // derived waveforms: (B, 2, 3, N) f_i then argS per binary, spacecraft and
    // sample, as the class's hooks returned them (null otherwise)
    const double *ext;
};

// Bound on |logL_mixed - logL_FP64| / (<h|h> + |<d|h>|) for the mixed
// likelihood: the largest ratio measured over 1e6 binaries at 0.1-20 mHz,
// N = 128-2048, 4 yr (1.9e-6, among templates with logL errors above 1e-9)
// with a 30% margin.
constexpr double GUARD_REL = 2.5e-6;

// Bits of the start-bin sort keys written by k_setup, and the start-bin
// resolution of the order (keys are start >> SORT_KEY_SHIFT).
constexpr int SORT_KEY_BITS = 16;
constexpr int SORT_KEY_SHIFT = 8;

// Per-binary constants for one precision policy, computed once per binary
// by k_setup (64 binaries per wavefront) and copied into LDS by the
// waveform kernels.
template <class P>
struct Pack {
    Src s;
    typename P::Aux aux;
};

// One thread per binary, 256 per workgroup.  The workgroup's packs are
// staged in LDS and written out as consecutive 8-byte words, so that each
// store instruction covers 512 contiguous bytes rather than 64 lines one
// pack (240 or 304 bytes) apart.
template <class P>
__global__ void __launch_bounds__(256) k_setup(WaveArgs a, int N, Pack<P> *__restrict__ out) {
    static_assert(sizeof(Pack<P>) % 8 == 0 && alignof(Pack<P>) <= 8, "packs are whole 8-byte words");
    constexpr int WORDS = (int)(sizeof(Pack<P>) / 8);
    __shared__ unsigned long long stage[256 * WORDS];
    const long long b0 = (long long)blockIdx.x * 256;
    const long long b = b0 + threadIdx.x;
    const bool valid = b < a.B;
    Pack<P> pk;
    if (valid) {
        pk.s = make_src(a.params, a.num_params, a.B, b, N, a.T, a.tdi2, a.armlength);
        pk.aux = P::make_aux(pk.s);
        reinterpret_cast<Pack<P> *>(stage)[threadIdx.x] = pk;
    }
    __syncthreads();
    const int words = (int)(a.B - b0 < 256 ? a.B - b0 : 256) * WORDS;
    unsigned long long *dst = reinterpret_cast<unsigned long long *>(out + b0);
    for (int i = threadIdx.x; i < words; i += 256) dst[i] = stage[i];
    if (valid && a.keys) {
        // order by start bin: neighbouring workgroups then read overlapping
        // data/PSD windows, which stay resident in L2.  Keys are the start
        // bin over 256 in 16 bits (start bins within +-2^23, f0 up to 66 mHz
        // at 4 yr; others clamped, which only affects the order), so the
        // radix sort runs two 8-bit passes; binaries in the same 256-bin cell
        // keep their input order.  Measured at 1e6 x N = 256: the
        // third pass of full start-bin keys cost more (22 us FP64) than the
        // coarser order costs the fused kernel (0-16 us).
        long long k = (pk.s.start >> SORT_KEY_SHIFT) + (1LL << (SORT_KEY_BITS - 1));
        k = k < 0 ? 0 : (k >= (1LL << SORT_KEY_BITS) ? (1LL << SORT_KEY_BITS) - 1 : k);
        a.keys[b] = (unsigned)k;
        a.vals[b] = (int)b;
    }
}

// Values shared by a whole wavefront, moved into scalar registers: once a
// wavefront owns one binary (TPB >= 64) its per-binary constants are
// wave-uniform, and keeping them in SGPRs frees the vector registers they
// would otherwise hold for the whole sample loop.
__device__ __forceinline__ int to_sgpr(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float to_sgpr(float x) { return __int_as_float(to_sgpr(__float_as_int(x))); }
__device__ __forceinline__ long long to_sgpr(long long x) {
    const unsigned lo = (unsigned)to_sgpr((int)(x & 0xffffffffLL));
    const unsigned hi = (unsigned)to_sgpr((int)(x >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double to_sgpr(double x) { return __longlong_as_double(to_sgpr(__double_as_longlong(x))); }

__device__ __forceinline__ Src to_sgpr(const Src &v) {
    Src s;
    s.amp = to_sgpr(v.amp);
    s.f0 = to_sgpr(v.f0);
    s.fdot = to_sgpr(v.fdot);
    s.fddot = to_sgpr(v.fddot);
    s.phi0 = to_sgpr(v.phi0);
    s.om = to_sgpr(v.om);
    s.omq = to_sgpr(v.omq);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        s.k[i] = to_sgpr(v.k[i]);
        s.u[i] = to_sgpr(v.u[i]);
        s.v[i] = to_sgpr(v.v[i]);
    }
    s.DPr = to_sgpr(v.DPr);
    s.DPi = to_sgpr(v.DPi);
    s.DCr = to_sgpr(v.DCr);
    s.DCi = to_sgpr(v.DCi);
    s.A2 = to_sgpr(v.A2);
    s.varpi = to_sgpr(v.varpi);
    s.e2 = to_sgpr(v.e2);
    s.n2 = to_sgpr(v.n2);
    s.T2 = to_sgpr(v.T2);
    s.tdr = to_sgpr(v.tdr);
    s.tdi = to_sgpr(v.tdi);
    s.start = to_sgpr(v.start);
    s.tier = to_sgpr(v.tier);
    return s;
}

__device__ __forceinline__ F64::Aux to_sgpr(const F64::Aux &v) { return v; }
__device__ __forceinline__ SrcF to_sgpr(const SrcF &v) {
    SrcF f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        f.k[i] = to_sgpr(v.k[i]);
        f.u[i] = to_sgpr(v.u[i]);
        f.v[i] = to_sgpr(v.v[i]);
    }
    f.DP = make_float2(to_sgpr(v.DP.x), to_sgpr(v.DP.y));
    f.DC = make_float2(to_sgpr(v.DC.x), to_sgpr(v.DC.y));
    f.f0 = to_sgpr(v.f0);
    f.fdot = to_sgpr(v.fdot);
    f.fddot = to_sgpr(v.fddot);
    return f;
}

// v itself when a group of TPB lanes spans less than a wavefront (several
// binaries per wavefront), else its scalar-register copy.
template <int TPB, class T>
__device__ __forceinline__ T per_group(const T &v) {
    if constexpr (TPB >= 64)
        return to_sgpr(v);
    else
        return v;
}

// Cooperative copy of binary b's Pack into LDS by the TPB lanes owning it.
template <class P, int TPB>
__device__ __forceinline__ void load_pack(const WaveArgs &a, long long b, int lane, Pack<P> *dst) {
    static_assert(sizeof(Pack<P>) % 4 == 0, "Pack must be word-sized");
    constexpr int W = sizeof(Pack<P>) / 4;
    const unsigned *src = reinterpret_cast<const unsigned *>(static_cast<const Pack<P> *>(a.pack) + b);
    unsigned *d = reinterpret_cast<unsigned *>(dst);
    for (int i = lane; i < W; i += TPB) d[i] = src[i];
}

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

// LDS of one binary for the third-body pre-pass: when the binary spans
// several wavefronts, the last sample's xi and g of each wavefront (per
// spacecraft) for the next one (6 WPB doubles); unused otherwise.
struct ThirdScratch {
    double *edge;
};

// Computes the NCH spectra of this group's binary into `buf` (NCH x N,
// natural FFT order, before fftshift and before P::fac()).  NCH == 3 gives
// X, Y, Z; NCH == 2 gives A, E.  Invalid slots (beyond the batch) write
// zeros and still take part in every barrier.
//
// Channel passes (C0, NCP): only channels C0 .. C0 + NCP - 1 are stored and
// transformed, buf then holding NCP x N values.  The large-N kernels call
// this once per group of channels that fits their LDS, recomputing the slow
// part (shared by all channels) in each pass.
//
// Each thread evaluates SPT samples: n = lane + q TPB, or with a third body
// n = SPT lane + q, so that the pre-pass values of a thread's own consecutive
// samples stay in its registers; only the predecessor of its first sample
// comes from the previous lane (a shuffle) or the previous wavefront
// (ThirdScratch).
//
// DFT != GBGPU_TRANSFORM_LDS (N = 256, one wavefront per binary, no third
// body; MFMA_F64 for the FP64 policy, the others for mixed): each thread
// evaluates the samples the transform's register layout gives it, xr[c][i]
// holding sample dft256_index(lane, i) of channel c (lane + 64 i for
// MFMA_F64, as with REGS), and dft256_mfma transforms them in place (same
// index on exit); buf is not touched, and tw only supplies the FP64
// variant's constants.
//
// REGS (LDS transform, circular binaries, all channels in one pass): the
// samples stay in registers, xr[c * SPT + q] = sample lane + q TPB, and
// fft_lds_regs returns the spectrum there, xr[c * SPT + p] = bin lane + p TPB;
// buf is the transform's working storage only.
template <int N, int NCH, bool THIRD, class P, int M = 1, int DFT = GBGPU_TRANSFORM_LDS, int C0 = 0, int NCP = NCH,
          bool EXT = false, bool REGS = false>
__device__ __forceinline__ void binary_spectrum(const Src &s, const typename P::Aux &aux, bool valid, int lane,
                                                const double *__restrict__ tab, double fstar_inv, double scale,
                                                typename P::CT *buf, const typename P::CT *__restrict__ tw,
                                                ThirdScratch ts, double *tot, typename P::CT *xr = nullptr,
                                                const void *dftc = nullptr, const double *ext = nullptr) {
    using G = Geo<N, M>;
    using CT = typename P::CT;
    constexpr int TPB = G::TPB, SPT = G::SPT;
    static_assert(DFT == GBGPU_TRANSFORM_LDS ||
                      (N == 256 && TPB == 64 && !THIRD &&
                       (DFT == GBGPU_TRANSFORM_MFMA_F64) == std::is_same<CT, double2>::value),
                  "the matrix-core transforms cover N = 256 without a third body: the FP64 variant for FP64 "
                  "spectra, the others for mixed");
    static_assert(C0 >= 0 && NCP >= 1 && C0 + NCP <= NCH && (DFT == GBGPU_TRANSFORM_LDS || NCP == NCH),
                  "channel pass out of range");
    static_assert(!EXT || (!THIRD && DFT == GBGPU_TRANSFORM_LDS && std::is_same<P, F64>::value),
                  "derived-waveform hooks run on the FP64 circular path with the LDS transform");
    static_assert(!REGS || (!THIRD && DFT == GBGPU_TRANSFORM_LDS && C0 == 0 && NCP == NCH),
                  "register-ended transform: circular binaries, LDS transform, one pass");
    constexpr int NT = THIRD ? SPT : 1;
    float vf[3][NT];   // v_LOS at xi_sc(t_n)
    double gc[3][NT];  // the cumulative phase term at t_n
    if constexpr (THIRD) {
        // Third-body pre-pass (thirdbody.py:155-226, :382-428): v_LOS and
        // g = v_LOS f_GW at xi_sc(t_n), then the cumulative trapezoid
        // term_n = sum_{m=1..n} (xi_m - xi_{m-1}) (g_{m-1} + g_m)/2 * 2 pi / c.
        constexpr int W = TPB < 64 ? TPB : 64;
        double xs[3][SPT];
#pragma unroll
        for (int i = 0; i < SPT; i++) {
            const int n = SPT * lane + i;
            double kP[3] = {0.0, 0.0, 0.0};
            if (valid) load_kP(s, tab, N, n, kP);
#pragma unroll
            for (int sc = 0; sc < 3; sc++) {
                double xi = 0.0, gg = 0.0;
                float v = 0.0f;
                if (valid) {
                    xi = tab[ROW_T * N + n] - kP[sc];
                    const double vd = third_vlos(s, xi);
                    v = (float)vd;
                    gg = vd * (s.f0 + s.fdot * xi + 0.5 * s.fddot * xi * xi);
                }
                xs[sc][i] = xi;
                vf[sc][i] = v;
                gc[sc][i] = gg;
            }
        }
        // sample SPT lane - 1 is the previous lane's last
        double xprev[3], gprev[3];
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            xprev[sc] = __shfl_up(xs[sc][SPT - 1], 1, W);
            gprev[sc] = __shfl_up(gc[sc][SPT - 1], 1, W);
        }
        if constexpr (TPB > 64) {
            const int w = lane / 64, wl = lane % 64;
            if (wl == 63) {
#pragma unroll
                for (int sc = 0; sc < 3; sc++) {
                    ts.edge[6 * w + 2 * sc] = xs[sc][SPT - 1];
                    ts.edge[6 * w + 2 * sc + 1] = gc[sc][SPT - 1];
                }
            }
            __syncthreads();
            if (wl == 0 && w > 0) {
#pragma unroll
                for (int sc = 0; sc < 3; sc++) {
                    xprev[sc] = ts.edge[6 * (w - 1) + 2 * sc];
                    gprev[sc] = ts.edge[6 * (w - 1) + 2 * sc + 1];
                }
            }
        }
        double loc[3][SPT];
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < SPT; i++) {
                const int n = SPT * lane + i;
                if (n > 0) {
                    const double xp = i > 0 ? xs[sc][i - 1] : xprev[sc];
                    const double gp = i > 0 ? gc[sc][i - 1] : gprev[sc];
                    const double dtt = xs[sc][i] - xp;
                    acc += dtt * (gp + gc[sc][i]) / 2.0 * PI * 2 / C_SI;
                }
                loc[sc][i] = acc;
            }
        }
#pragma unroll
        for (int sc = 0; sc < 3; sc++) {
            const double pre = group_excl_scan<TPB>(loc[sc][SPT - 1], lane, tot);
#pragma unroll
            for (int i = 0; i < SPT; i++) gc[sc][i] = pre + loc[sc][i];
        }
    }
    using R = decltype(CT{}.x);
    // The series tier is fixed per binary: one branch per tier around the
    // whole sample loop gives each tier straight-line code with its own
    // register allocation (a branch per sample keeps every tier's values
    // live at once).
    auto samples = [&](auto tier_c) {
        constexpr int TIER = decltype(tier_c)::value;
#pragma unroll
        for (int q = 0; q < SPT; q++) {
            // third body: the thread's own consecutive samples (pre-pass values in
            // registers); otherwise strided, so orbit-table loads coalesce
            const int n = THIRD ? SPT * lane + q
                                : (DFT != GBGPU_TRANSFORM_LDS && DFT != GBGPU_TRANSFORM_MFMA_F64 ? dft256_index(lane, q)
                                                                                                 : lane + q * TPB);
            CT X{}, Y{}, Z{}, F{};
            if (valid) {
                float v3[3] = {0.f, 0.f, 0.f};
                double t3[3] = {0.0, 0.0, 0.0};
                if constexpr (THIRD) {
#pragma unroll
                    for (int sc = 0; sc < 3; sc++) {
                        v3[sc] = vf[sc][q];
                        t3[sc] = gc[sc][q];
                    }
                }
                if constexpr (EXT) {
                    // this binary's hook outputs: ext[(0, sc, n)] = f_i, ext[(1, sc, n)] = argS
                    double fi3[3], ar3[3];
#pragma unroll
                    for (int sc = 0; sc < 3; sc++) {
                        fi3[sc] = ext[sc * N + n];
                        ar3[sc] = ext[(3 + sc) * N + n];
                    }
                    F64::template sample_ext<TIER>(s, tab, N, n, fstar_inv, scale, fi3, ar3, X, Y, Z, F);
                } else {
                    P::template sample<THIRD, TIER>(s, aux, tab, N, n, fstar_inv, scale, v3, t3, X, Y, Z, F);
                }
            }
            CT out[NCH];
            if constexpr (NCH == 3) {
                out[0] = cmul(F, X);
                out[1] = cmul(F, Y);
                out[2] = cmul(F, Z);
            } else {
                // A = (Z - X)/sqrt2, E = (X - 2Y + Z)/sqrt6 (utility.py:34-38); linear,
                // so it commutes with the transform, and X, Y, Z share the factor F.
                out[0] = cmul(cscale(F, (R)INV_SQRT2), csub(Z, X));
                out[1] = cmul(cscale(F, (R)INV_SQRT6), cadd(csub(X, cscale(Y, (R)2)), Z));
            }
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                if constexpr (DFT != GBGPU_TRANSFORM_LDS || REGS)
                    xr[c * SPT + q] = out[c];
                else if (c >= C0 && c < C0 + NCP)
                    buf[(c - C0) * N + n] = out[c];
            }
        }
    };
    // hooked frequencies may leave the binary's tier: reduced sine/cosine throughout
    if constexpr (EXT)
        samples(std::integral_constant<int, NTIER>{});
    else
        with_tier(s.tier, samples);
    if constexpr (DFT == GBGPU_TRANSFORM_MFMA_F64) {
        const Dft256Lane<DFT> k = load_dft256_f64(tw, lane);
#pragma unroll
        for (int c = 0; c < NCH; c++) dft256_mfma<DFT>(xr + c * SPT, k);
    } else if constexpr (DFT != GBGPU_TRANSFORM_LDS) {
        const Dft256Lane<DFT> k = load_dft256<DFT>(dftc, lane);
#pragma unroll
        for (int c = 0; c < NCH; c++) dft256_mfma<DFT>(xr + c * SPT, k);
    } else if constexpr (REGS) {
        fft_lds_regs<N, NCH, TPB>(xr, buf, lane, tw);
    } else {
        __syncthreads();
        fft_lds<N, NCP, TPB>(buf, lane, tw);
    }
}

// One bin of <d|x> and <x|x> for the A and E channels: xr + i xi +=
// conj(d) x / S, xx += |x|^2 / S, with explicit fused multiply-adds so that
// every caller rounds alike.
__device__ __forceinline__ void ll_bin(double2 hA, double2 hE, double2 dA, double2 dE, double iA, double iE,
                                       double &xr, double &xi, double &xx) {
    const double rA = fma(dA.x, hA.x, dA.y * hA.y), rE = fma(dE.x, hE.x, dE.y * hE.y);
    const double jA = fma(dA.x, hA.y, -(dA.y * hA.x)), jE = fma(dE.x, hE.y, -(dE.y * hE.x));
    const double nA = fma(hA.x, hA.x, hA.y * hA.y), nE = fma(hE.x, hE.x, hE.y * hE.y);
    xr = fma(rE, iE, fma(rA, iA, xr));
    xi = fma(jE, iE, fma(jA, iA, xi));
    xx = fma(nE, iE, fma(nA, iA, xx));
}

// Accumulates this thread's share of <d|h> (complex) and <h|h> for a binary
// whose A/E spectra h = fac x have x in buf (natural order), over output bins
// m = lane + q TPB (fftshifted), data bin j = j0 + m.  The sums are formed
// for x and scaled once: <d|h> = fac <d|x>, <h|h> = |fac|^2 <x|x>.
template <int N, class P, int M = 1>
__device__ __forceinline__ void ll_partial(const typename P::CT *buf, double2 fac, int lane, long long j0,
                                           long long doff, long long noff, const WaveArgs &a, double &dhr,
                                           double &dhi, double &hh) {
    using G = Geo<N, M>;
    double xr = 0.0, xi = 0.0, xx = 0.0;
#pragma unroll
    for (int q = 0; q < G::SPT; q++) {
        const int m = lane + q * G::TPB;
        const int idx = (m + N / 2) & (N - 1);
        const long long j = j0 + m;
        if (j < 0 || j >= a.L) continue;
        ll_bin(P::up(buf[idx]), P::up(buf[N + idx]), a.dA[doff + j], a.dE[doff + j], rcp_f64(a.SA[noff + j]),
               rcp_f64(a.SE[noff + j]), xr, xi, xx);
    }
    dhr += fac.x * xr - fac.y * xi;
    dhi += fac.x * xi + fac.y * xr;
    hh += (fac.x * fac.x + fac.y * fac.y) * xx;
}

// One channel's share of ll_partial: channel CH (0 = A, 1 = E) of x in
// `buf` (natural order), adding conj(d) x / S and |x|^2 / S of this thread's
// bins to xr, xi, xx (before the factor fac, applied once per binary).
template <int N, class P, int CH, int M = 1>
__device__ __forceinline__ void ll_channel(const typename P::CT *buf, int lane, long long j0, long long doff,
                                           long long noff, const WaveArgs &a, double &xr, double &xi, double &xx) {
    using G = Geo<N, M>;
    const double2 *d = CH == 0 ? a.dA : a.dE;
    const double *S = CH == 0 ? a.SA : a.SE;
#pragma unroll
    for (int q = 0; q < G::SPT; q++) {
        const int m = lane + q * G::TPB;
        const long long j = j0 + m;
        if (j < 0 || j >= a.L) continue;
        const double2 h = P::up(buf[(m + N / 2) & (N - 1)]);
        const double2 dv = d[doff + j];
        const double iS = rcp_f64(S[noff + j]);
        xr = fma(fma(dv.x, h.x, dv.y * h.y), iS, xr);
        xi = fma(fma(dv.x, h.y, -(dv.y * h.x)), iS, xi);
        xx = fma(fma(h.x, h.x, h.y * h.y), iS, xx);
    }
}

// ll_partial for spectra held in registers: spec(c, i) is channel c's (A,
// E) value for fftshifted output position pos(i), i < SPT.
template <class P, int SPT, class Pos, class Spec>
__device__ __forceinline__ void ll_partial_regs(Spec spec, Pos pos, double2 fac, long long j0, long long doff,
                                                long long noff, const WaveArgs &a, double &dhr, double &dhi,
                                                double &hh) {
    double xr = 0.0, xi = 0.0, xx = 0.0;
#pragma unroll
    for (int i = 0; i < SPT; i++) {
        const long long j = j0 + pos(i);
        if (j < 0 || j >= a.L) continue;
        ll_bin(P::up(spec(0, i)), P::up(spec(1, i)), a.dA[doff + j], a.dE[doff + j], rcp_f64(a.SA[noff + j]),
               rcp_f64(a.SE[noff + j]), xr, xi, xx);
    }
    dhr += fac.x * xr - fac.y * xi;
    dhi += fac.x * xi + fac.y * xr;
    hh += (fac.x * fac.x + fac.y * fac.y) * xx;
}

}  // namespace gbgpu
