// swap_likelihood_difference: change in log-likelihood when binary h_r is
// replaced by h_a against a residual r (data minus the current global
// template, which contains h_r):
//   dlogL = logL(r + h_r - h_a) - logL(r)
//         = Re<r|h_a> - Re<r|h_r> - 1/2 (<h_a|h_a> + <h_r|h_r> - 2 Re<h_r|h_a>)
// with <x|y> = 4 df sum_{A,E} conj(x) y / S, the reference's inner product
// (src/gbgpu/cutils/gbgpu_utils.cu:156-183).  The reference has no such
// entry point; this is its likelihood algebra for a remove/add proposal.
//
// A workgroup holds proposal pairs: slot 2q computes h_r and slot 2q+1 h_a
// (same N), both spectra stay in LDS, the remove group sums over the remove
// window (including the overlap term, reading h_a from its neighbour's
// buffer) and the add group over the add window.

#include "common.hpp"
#include "wave_common.hpp"

namespace gbgpu {

int check_cfg(const double *params, int num_params, int num_bin, const gbgpu_wave_cfg *cfg);
WaveArgs base_args(const double *params, int num_params, long long num_bin, const gbgpu_wave_cfg *cfg);
void set_data(WaveArgs &a, const gbgpu_data *data);
int check_data(const gbgpu_data *data);
const void *twiddles_for(int N, int precision);

// Slots per workgroup for swap: an even number of binary groups.
template <int N>
struct SwapGeo {
    static constexpr int TPB = Geo<N>::TPB;
    static constexpr int WG = (2 * TPB > WG_SIZE) ? 2 * TPB : WG_SIZE;
    static constexpr int BPW = WG / TPB;  // even
    static constexpr int WPB = Geo<N>::WPB;
};

// LDS of one k_swap workgroup; arrays a variant does not use have length zero.
template <int N, bool THIRD, class P>
struct SwapLds {
    using G = SwapGeo<N>;
    using CT = typename P::CT;
    static constexpr bool PACK_IN_LDS = G::TPB < 64;
    static constexpr int NEDGE = (THIRD && G::WPB > 1) ? 6 * G::WPB : 0;
    CT sbuf[G::BPW * 2 * N];
    Pack<P> spack[PACK_IN_LDS ? G::BPW : 0];
    double stot[G::WPB > 1 ? G::BPW * G::WPB : 0];
    double s_edge[G::BPW * NEDGE];
};

template <int N, bool THIRD, class P>
__global__ void __launch_bounds__(SwapGeo<N>::WG) k_swap(WaveArgs a, double *dll, double *terms) {
    using G = SwapGeo<N>;
    using CT = typename P::CT;
    constexpr int TPB = G::TPB, BPW = G::BPW, WPB = G::WPB;
    constexpr int SPT = Geo<N>::SPT;
    __shared__ SwapLds<N, THIRD, P> L;
    constexpr bool PACK_IN_LDS = SwapLds<N, THIRD, P>::PACK_IN_LDS;
    CT *sbuf = L.sbuf;

    const int tid = threadIdx.x;
    const int slot = per_group<TPB>(tid / TPB);
    const int lane = tid % TPB;
    const long long b = (long long)blockIdx.x * BPW + slot;  // 2p = remove, 2p+1 = add
    const long long prop = b >> 1;
    const bool valid = b < a.B;
    const bool is_add = slot & 1;

    // per-binary constants: scalar loads when a wavefront owns the binary,
    // else shared through LDS; the partner's (slot | 1) likewise
    const Pack<P> *packs = static_cast<const Pack<P> *>(a.pack);
    const Pack<P> *pk, *ppk;
    if constexpr (PACK_IN_LDS) {
        if (valid) load_pack<P, TPB>(a, b, lane, L.spack + slot);
        __syncthreads();
        pk = L.spack + slot;
        ppk = L.spack + (slot | 1);
    } else {
        pk = packs + (valid ? b : 0);
        ppk = packs + ((b | 1) < a.B ? (b | 1) : 0);
    }
    const Src s = per_group<TPB>(pk->s);
    CT *buf = sbuf + slot * 2 * N;
    double *tot = L.stot + slot * WPB;
    ThirdScratch ts{L.s_edge + slot * SwapLds<N, THIRD, P>::NEDGE};
    binary_spectrum<N, 2, THIRD, P>(s, per_group<TPB>(pk->aux), valid, lane, a.tab, a.fstar_inv, a.scale, buf,
                                    static_cast<const CT *>(a.tw), ts, tot);

    // partial sums: 0 Re<r|h_r>, 1 Re<r|h_a>, 2 <h_r|h_r>, 3 <h_a|h_a>, 4 Re<h_r|h_a>
    double rd = 0.0, hh = 0.0, cross = 0.0;
    if (valid) {
        const double2 fac = P::fac(s);
        const long long doff = (a.di ? (long long)a.di[prop] : 0LL) * a.L;
        const long long noff = (a.ni ? (long long)a.ni[prop] : 0LL) * a.L;
        const long long j0 = s.start - a.start_freq_ind;
        // the partner (add binary) of a remove slot
        const Src sp = per_group<TPB>(ppk->s);
        const CT *pbuf = sbuf + (slot | 1) * 2 * N;
        const double2 pfac = P::fac(sp);
        const long long pj0 = sp.start - a.start_freq_ind;
#pragma unroll
        for (int q = 0; q < SPT; q++) {
            const int m = lane + q * TPB;
            const int idx = (m + N / 2) & (N - 1);
            const long long j = j0 + m;
            if (j < 0 || j >= a.L) continue;
            const double2 hA = cmul(P::up(buf[idx]), fac), hE = cmul(P::up(buf[N + idx]), fac);
            const double2 dA = a.dA[doff + j], dE = a.dE[doff + j];
            const double iA = rcp_f64(a.SA[noff + j]), iE = rcp_f64(a.SE[noff + j]);
            rd += (dA.x * hA.x + dA.y * hA.y) * iA + (dE.x * hE.x + dE.y * hE.y) * iE;
            hh += (hA.x * hA.x + hA.y * hA.y) * iA + (hE.x * hE.x + hE.y * hE.y) * iE;
            if (!is_add) {
                const long long mp = j - pj0;  // same bin in the add template
                if (mp >= 0 && mp < N) {
                    const int pidx = ((int)mp + N / 2) & (N - 1);
                    const double2 aA = cmul(P::up(pbuf[pidx]), pfac), aE = cmul(P::up(pbuf[N + pidx]), pfac);
                    cross += (hA.x * aA.x + hA.y * aA.y) * iA + (hE.x * aE.x + hE.y * aE.y) * iE;
                }
            }
        }
    }
    rd = group_total<TPB>(rd, lane, tot);
    hh = group_total<TPB>(hh, lane, tot);
    cross = group_total<TPB>(cross, lane, tot);
    // the pair's sums meet in the first words of each slot's transform
    // buffer, once every group has finished reading the spectra
    __syncthreads();
    double *sred = reinterpret_cast<double *>(buf);
    if (lane == 0) {
        sred[0] = rd;
        sred[1] = hh;
        sred[2] = cross;
    }
    __syncthreads();
    if (valid && !is_add && lane == 0 && b + 1 < a.B) {
        const double *pred = reinterpret_cast<const double *>(sbuf + (slot + 1) * 2 * N);
        const double r_hr = a.df4 * sred[0];
        const double hr_hr = a.df4 * sred[1];
        const double hr_ha = a.df4 * sred[2];
        const double r_ha = a.df4 * pred[0];
        const double ha_ha = a.df4 * pred[1];
        dll[prop] = r_ha - r_hr - 0.5 * (ha_ha + hr_hr - 2.0 * hr_ha);
        if (terms) {
            terms[5 * prop + 0] = r_hr;
            terms[5 * prop + 1] = r_ha;
            terms[5 * prop + 2] = hr_hr;
            terms[5 * prop + 3] = ha_ha;
            terms[5 * prop + 4] = hr_ha;
        }
    }
}

template <bool THIRD, class P>
static int launch_swap(int N, WaveArgs a, double *dll, double *terms, hipStream_t st) {
    Pack<P> *pk = nullptr;
    hipError_t e = hipMallocAsync((void **)&pk, sizeof(Pack<P>) * (size_t)a.B, st);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(swap constants)");
    hipLaunchKernelGGL((k_setup<P>), dim3((unsigned)((a.B + 255) / 256)), dim3(256), 0, st, a, N, pk);
    a.pack = pk;
    int rc = GBGPU_OK;
#define GBGPU_CASE(NN)                                                                                 \
    case NN: {                                                                                         \
        const long long grid = (a.B + SwapGeo<NN>::BPW - 1) / SwapGeo<NN>::BPW;                        \
        KernelTimer kt(st);                                                                            \
        hipLaunchKernelGGL((k_swap<NN, THIRD, P>), dim3((unsigned)grid), dim3(SwapGeo<NN>::WG), 0, st, \
                           a, dll, terms);                                                             \
        break;                                                                                         \
    }
    switch (N) {
        GBGPU_CASE(32)
        GBGPU_CASE(64)
        GBGPU_CASE(128)
        GBGPU_CASE(256)
        GBGPU_CASE(512)
        GBGPU_CASE(1024)
        GBGPU_CASE(2048)
        default:
            set_error("N must be a power of two in [32, 2048]");
            rc = GBGPU_E_ARG;
    }
#undef GBGPU_CASE
    if (!rc) {
        e = hipGetLastError();
        if (e != hipSuccess) rc = hip_fail(e, "k_swap");
    }
    e = hipFreeAsync(pk, st);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipFreeAsync(swap constants)");
    return GBGPU_OK;
}

}  // namespace gbgpu

using namespace gbgpu;

extern "C" int gbgpu_swap_ll_diff(double *dll, double *terms, const double *params_pairs, int num_params,
                                  int num_prop, const gbgpu_wave_cfg *cfg, const gbgpu_data *data, void *stream) {
    int rc = check_cfg(params_pairs, num_params, 2 * num_prop, cfg);
    if (rc) return rc;
    if (cfg->transform != GBGPU_TRANSFORM_LDS) {
        set_error("gbgpu_swap_ll_diff uses the LDS transform only");
        return GBGPU_E_ARG;
    }
    if (cfg->N > 2048) {
        set_error("gbgpu_swap_ll_diff supports N <= 2048");
        return GBGPU_E_ARG;
    }
    if ((rc = check_data(data))) return rc;
    if (!dll) {
        set_error("dll must not be null");
        return GBGPU_E_ARG;
    }
    if (num_prop == 0) return GBGPU_OK;
    WaveArgs a = base_args(params_pairs, num_params, 2LL * num_prop, cfg);
    set_data(a, data);
    a.tw = twiddles_for(cfg->N, cfg->precision);
    if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool third = num_params == GBGPU_NPARAMS_THIRD;
    if (cfg->precision == GBGPU_MIXED)
        return third ? launch_swap<true, Mixed>(cfg->N, a, dll, terms, st)
                     : launch_swap<false, Mixed>(cfg->N, a, dll, terms, st);
    return third ? launch_swap<true, F64>(cfg->N, a, dll, terms, st)
                 : launch_swap<false, F64>(cfg->N, a, dll, terms, st);
}
