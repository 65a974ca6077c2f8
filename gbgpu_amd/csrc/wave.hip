// FastGB waveform kernels for gfx950: run_wave (XYZ/AET spectra to HBM) and
// the fused waveform + likelihood, for both precision policies (prec.hpp).
//
// Work decomposition: a group of TPB = min(N/4, 256) threads owns one binary
// and a 256-thread workgroup holds BPW = 256/TPB binaries (N = 2048: one
// 512-thread workgroup per binary).  Each thread evaluates the slow part at
// SPT = N/TPB samples, the group runs the length-N transform in LDS (for
// circular binaries with its first and last stages in registers), and the
// epilogue either streams the fftshifted, scaled spectra to HBM (run_wave)
// or reads the binary's data/PSD window with coalesced loads and reduces
// <d|h>, <h|h> over the group (wave_ll); one workgroup pass is wave_pass
// (wave_pass.hpp).  N = 4096 and 8192 run the
// large-N kernels of wave_big.hip; the batch segment kernels are in
// wave_seg.hip.
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "dispatch.hpp"
#include "wave_common.hpp"
#include "wave_pass.hpp"

namespace gbgpu {

int launch_wave_big(int N, int mode, bool third, int precision, const WaveArgs &a, hipStream_t st, char *ws,
                    size_t ws_bytes);

template <int N, int MODE, bool THIRD, class P, int DFT = GBGPU_TRANSFORM_LDS>
__global__ void __launch_bounds__((Geo<N, wave_tile<N>()>::WG)) k_wave(WaveArgs a) {
    constexpr int M = wave_tile<N>();
    __shared__ WaveLds<N, MODE, THIRD, P, P, M, DFT> lds;
    wave_pass<N, MODE, THIRD, P, P, M, DFT>(a, lds, blockIdx.x, a.B, 0, threadIdx.x, a.tab, a.tw);
}

// k_wave with the requested transform; the matrix-core variants exist at
// N = 256 without a third body, FP64 for the FP64 policy and the others for
// mixed (check_cfg has rejected every other combination).
template <int NN, int MODE, bool THIRD, class P>
static void launch_k_wave(int transform, long long grid, hipStream_t st, const WaveArgs &aa) {
    constexpr int WG = Geo<NN, wave_tile<NN>()>::WG;
    if constexpr (NN == 256 && !THIRD && std::is_same<P, F64>::value) {
        if (transform == GBGPU_TRANSFORM_MFMA_F64) {
            hipLaunchKernelGGL((k_wave<NN, MODE, THIRD, P, GBGPU_TRANSFORM_MFMA_F64>), dim3((unsigned)grid), dim3(WG),
                               0, st, aa);
            return;
        }
    }
    if constexpr (NN == 256 && !THIRD && std::is_same<P, Mixed>::value) {
        switch (transform) {
            case GBGPU_TRANSFORM_MFMA_F32:
                hipLaunchKernelGGL((k_wave<NN, MODE, THIRD, P, GBGPU_TRANSFORM_MFMA_F32>), dim3((unsigned)grid),
                                   dim3(WG), 0, st, aa);
                return;
            case GBGPU_TRANSFORM_MFMA_BF16X3:
                hipLaunchKernelGGL((k_wave<NN, MODE, THIRD, P, GBGPU_TRANSFORM_MFMA_BF16X3>), dim3((unsigned)grid),
                                   dim3(WG), 0, st, aa);
                return;
            case GBGPU_TRANSFORM_MFMA_BF16X6:
                hipLaunchKernelGGL((k_wave<NN, MODE, THIRD, P, GBGPU_TRANSFORM_MFMA_BF16X6>), dim3((unsigned)grid),
                                   dim3(WG), 0, st, aa);
                return;
            default:
                break;
        }
    }
    hipLaunchKernelGGL((k_wave<NN, MODE, THIRD, P>), dim3((unsigned)grid), dim3(WG), 0, st, aa);
}

// ws / ws_bytes: caller-provided per-binary workspace (nullptr: stream-ordered
// scratch allocated by with_pack).
template <int MODE, bool THIRD, class P>
static int launch_wave(int N, const WaveArgs &a, hipStream_t st, char *ws = nullptr, size_t ws_bytes = 0,
                       int transform = GBGPU_TRANSFORM_LDS) {
#define GBGPU_CASE(NN)                                                                              \
    case NN: {                                                                                      \
        using GG = Geo<NN, wave_tile<NN>()>;                                                        \
        const long long grid = (a.B + GG::BPW - 1) / GG::BPW;                                       \
        return with_pack<P>(                                                                        \
            a, NN, st,                                                                              \
            [&](const WaveArgs &aa) -> int {                                                        \
                KernelTimer kt(st);                                                                 \
                launch_k_wave<NN, MODE, THIRD, P>(transform, grid, st, aa);                         \
                GBGPU_CHECK_LAUNCH("k_wave");                                                       \
                return GBGPU_OK;                                                                    \
            },                                                                                      \
            ws, ws_bytes);                                                                          \
    }
    switch (N) {
        GBGPU_CASE(32)
        GBGPU_CASE(64)
        GBGPU_CASE(128)
        GBGPU_CASE(256)
        GBGPU_CASE(512)
        GBGPU_CASE(1024)
        GBGPU_CASE(2048)
        case 4096:
        case 8192:
            return launch_wave_big(N, MODE, THIRD, std::is_same<P, Mixed>::value ? GBGPU_MIXED : GBGPU_FP64, a, st,
                                   ws, ws_bytes);
        default:
            set_error("N must be a power of two in [32, 8192]");
            return GBGPU_E_ARG;
    }
#undef GBGPU_CASE
}

template <int MODE>
static int launch_any(int N, bool third, int precision, const WaveArgs &a, hipStream_t st, char *ws = nullptr,
                      size_t ws_bytes = 0, int transform = GBGPU_TRANSFORM_LDS) {
    if (precision == GBGPU_MIXED)
        return third ? launch_wave<MODE, true, Mixed>(N, a, st, ws, ws_bytes)
                     : launch_wave<MODE, false, Mixed>(N, a, st, ws, ws_bytes, transform);
    return third ? launch_wave<MODE, true, F64>(N, a, st, ws, ws_bytes)
                 : launch_wave<MODE, false, F64>(N, a, st, ws, ws_bytes, transform);
}

// a.dftc from the per-device constant cache when a matrix-core transform is
// requested (check_cfg has validated the combination).
static int set_transform(const gbgpu_wave_cfg *cfg, WaveArgs &a) {
    // the FP64 variant takes its constants from the FP64 twiddles (a.tw)
    if (cfg->transform == GBGPU_TRANSFORM_LDS || cfg->transform == GBGPU_TRANSFORM_MFMA_F64) return GBGPU_OK;
    a.dftc = dft256_table();
    if (!a.dftc) return hip_fail(hipErrorOutOfMemory, "transform constants");
    return GBGPU_OK;
}

int check_cfg(const double *params, int num_params, int num_bin, const gbgpu_wave_cfg *cfg) {
    if (!cfg || !params || !cfg->orbit_table) {
        set_error("null params, cfg or orbit_table");
        return GBGPU_E_ARG;
    }
    if (num_params != GBGPU_NPARAMS && num_params != GBGPU_NPARAMS_THIRD) {
        set_error("num_params must be 9 or 14");
        return GBGPU_E_ARG;
    }
    if (num_bin < 0 || !is_pow2(cfg->N) || cfg->N < 32 || cfg->N > GBGPU_N_MAX) {
        set_error("num_bin < 0 or N not a power of two in [32, 8192]");
        return GBGPU_E_ARG;
    }
    if (!(cfg->T > 0) || !(cfg->armlength > 0)) {
        set_error("T and armlength must be positive");
        return GBGPU_E_ARG;
    }
    if (cfg->precision != GBGPU_FP64 && cfg->precision != GBGPU_MIXED) {
        set_error("unknown precision");
        return GBGPU_E_ARG;
    }
    if (cfg->transform < GBGPU_TRANSFORM_LDS || cfg->transform > GBGPU_TRANSFORM_MFMA_F64) {
        set_error("unknown transform");
        return GBGPU_E_ARG;
    }
    if (cfg->transform != GBGPU_TRANSFORM_LDS && (cfg->N != 256 || num_params != GBGPU_NPARAMS)) {
        set_error("the matrix-core transforms need N = 256 and no third body");
        return GBGPU_E_ARG;
    }
    if (cfg->transform != GBGPU_TRANSFORM_LDS &&
        (cfg->transform == GBGPU_TRANSFORM_MFMA_F64) != (cfg->precision == GBGPU_FP64)) {
        set_error("GBGPU_TRANSFORM_MFMA_F64 goes with precision GBGPU_FP64, the other matrix-core transforms "
                  "with GBGPU_MIXED");
        return GBGPU_E_ARG;
    }
    return GBGPU_OK;
}

WaveArgs base_args(const double *params, int num_params, long long num_bin, const gbgpu_wave_cfg *cfg) {
    WaveArgs a = {};
    a.params = params;
    a.num_params = num_params;
    a.B = num_bin;
    a.T = cfg->T;
    a.scale = 0.5 * cfg->T / cfg->N;
    a.fstar_inv = 1.0 / (C_SI / (cfg->armlength * 2 * PI));
    a.armlength = cfg->armlength;
    a.tdi2 = cfg->tdi2;
    a.tab = cfg->orbit_table;
    return a;
}

void set_data(WaveArgs &a, const gbgpu_data *data) {
    a.dA = reinterpret_cast<const double2 *>(data->A);
    a.dE = reinterpret_cast<const double2 *>(data->E);
    a.SA = data->psdA;
    a.SE = data->psdE;
    a.di = data->data_index;
    a.ni = data->noise_index;
    a.L = data->data_length;
    a.start_freq_ind = data->start_freq_ind;
    a.df4 = 4.0 * data->df;
}

int check_data(const gbgpu_data *data) {
    if (!data || !data->A || !data->E || !data->psdA || !data->psdE || data->data_length <= 0) {
        set_error("null data pointers or data_length <= 0");
        return GBGPU_E_ARG;
    }
    return GBGPU_OK;
}

const void *twiddles_for(int N, int precision) {
    return precision == GBGPU_MIXED ? (const void *)twiddles_f32(N) : (const void *)twiddles_f64(N);
}

// ---- allocation-free fused likelihood (gbgpu_wave_ll_noalloc) ---------------
//
// `scratch` holds, at fixed offsets: the FP64 and FP32 twiddle tables of every
// supported N (N = 32 << k, k < 9, back to back), the per-lane constants of
// the matrix-core transforms (dft256_table's layout), then the per-binary
// workspace of with_pack (Pack array, sort keys/values, radix-sort storage)
// for up to GBGPU_NOALLOC_MAX_BIN binaries.  Each device holds its own copy
// of the array for the life of the process.  In host code `scratch` names a
// host-side stand-in: every host use goes through the device address that
// hipGetSymbolAddress returns for the current device (noalloc_base).
constexpr int NOALLOC_NK = 9;
constexpr size_t NOALLOC_TW_ENTRIES = 32 * ((1 << NOALLOC_NK) - 1);  // sum of N over 32..8192
constexpr size_t NOALLOC_TW64_BYTES = (16 * NOALLOC_TW_ENTRIES + 255) & ~(size_t)255;
constexpr size_t NOALLOC_TW32_BYTES = (8 * NOALLOC_TW_ENTRIES + 255) & ~(size_t)255;
constexpr size_t NOALLOC_DFT_BYTES = (size_t)DFT_TABLE_ROWS * 64 * 16;
constexpr size_t NOALLOC_TW_BYTES = NOALLOC_TW64_BYTES + NOALLOC_TW32_BYTES + NOALLOC_DFT_BYTES;
constexpr size_t NOALLOC_BYTES = (size_t)24 << 20;
constexpr size_t NOALLOC_WS_BYTES = NOALLOC_BYTES - NOALLOC_TW_BYTES;
static_assert(sizeof(Pack<F64>) * (size_t)GBGPU_NOALLOC_MAX_BIN + ((size_t)2 << 20) < NOALLOC_WS_BYTES,
              "scratch too small for GBGPU_NOALLOC_MAX_BIN FP64 packs");
static_assert(sizeof(Pack<Mixed>) * (size_t)GBGPU_NOALLOC_MAX_BIN + ((size_t)2 << 20) < NOALLOC_WS_BYTES,
              "scratch too small for GBGPU_NOALLOC_MAX_BIN mixed packs");

static __device__ __align__(256) char scratch[NOALLOC_BYTES];

namespace {
std::mutex g_noalloc_mutex;
std::map<int, char *> g_noalloc_base;  // device -> address of its scratch

inline size_t tw_offset(int N) { return (size_t)N - 32; }  // entries before N's table (32 + 64 + ... = N - 32)

// Device address of the current device's scratch, set up on first use: the
// symbol address is looked up once per device, with that device current, and
// the twiddle tables are copied in (blocking; nothing is allocated).
int noalloc_base(char **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    auto it = g_noalloc_base.find(dev);
    if (it != g_noalloc_base.end()) {
        *out = it->second;
        return GBGPU_OK;
    }
    char *ptr = nullptr;
    e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) {
        set_error("hipGetSymbolAddress(scratch) returned a null address");
        return GBGPU_E_HIP;
    }
    // the radix sort's storage for the largest batch must fit beside the packs
    for (int prec = 0; prec < 2; prec++) {
        size_t total = 0, sort_bytes = 0;
        int rc = prec == GBGPU_MIXED
                     ? pack_workspace_bytes<Mixed>(GBGPU_NOALLOC_MAX_BIN, nullptr, &total, &sort_bytes, false)
                     : pack_workspace_bytes<F64>(GBGPU_NOALLOC_MAX_BIN, nullptr, &total, &sort_bytes, false);
        if (rc) return rc;
        if (total > NOALLOC_WS_BYTES) {
            set_error("allocation-free scratch is smaller than the workspace of GBGPU_NOALLOC_MAX_BIN binaries");
            return GBGPU_E_ARG;
        }
    }
    std::vector<double2> t64(NOALLOC_TW_ENTRIES);
    std::vector<float2> t32(NOALLOC_TW_ENTRIES);
    for (int k = 0; k < NOALLOC_NK; k++) {
        const int N = 32 << k;
        host_twiddles_f64(N, t64.data() + tw_offset(N));
        host_twiddles_f32(N, t32.data() + tw_offset(N));
    }
    std::vector<uint32_t> dft(NOALLOC_DFT_BYTES / 4, 0u);
    host_dft256_table(dft.data());
    e = hipMemcpy(ptr, t64.data(), 16 * NOALLOC_TW_ENTRIES, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ptr + NOALLOC_TW64_BYTES, t32.data(), 8 * NOALLOC_TW_ENTRIES, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(ptr + NOALLOC_TW64_BYTES + NOALLOC_TW32_BYTES, dft.data(), NOALLOC_DFT_BYTES,
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(twiddles to scratch)");
    g_noalloc_base[dev] = ptr;
    *out = ptr;
    return GBGPU_OK;
}
}  // namespace

}  // namespace gbgpu

using namespace gbgpu;

extern "C" int gbgpu_run_wave(gbgpu_cmplx *XYZf, gbgpu_cmplx *AETf, long long *start_inds, const double *params,
                              int num_params, int num_bin, const gbgpu_wave_cfg *cfg, void *stream) {
    int rc = check_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if (!start_inds) {
        set_error("start_inds must not be null");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    WaveArgs a = base_args(params, num_params, num_bin, cfg);
    a.XYZf = reinterpret_cast<double2 *>(XYZf);
    a.AETf = reinterpret_cast<double2 *>(AETf);
    a.start_inds = start_inds;
    a.tw = twiddles_for(cfg->N, cfg->precision);
    if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
    if ((rc = set_transform(cfg, a))) return rc;
    return launch_any<MODE_WAVE>(cfg->N, num_params == GBGPU_NPARAMS_THIRD, cfg->precision, a,
                              static_cast<hipStream_t>(stream), nullptr, 0, cfg->transform);
}

extern "C" int gbgpu_run_wave_ae(gbgpu_cmplx *AEf, long long *start_inds, const double *params, int num_params,
                                 int num_bin, const gbgpu_wave_cfg *cfg, void *stream) {
    int rc = check_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if (!AEf || !start_inds) {
        set_error("AEf and start_inds must not be null");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    WaveArgs a = base_args(params, num_params, num_bin, cfg);
    a.AETf = reinterpret_cast<double2 *>(AEf);
    a.start_inds = start_inds;
    a.tw = twiddles_for(cfg->N, cfg->precision);
    if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
    if ((rc = set_transform(cfg, a))) return rc;
    return launch_any<MODE_AE>(cfg->N, num_params == GBGPU_NPARAMS_THIRD, cfg->precision, a,
                              static_cast<hipStream_t>(stream), nullptr, 0, cfg->transform);
}

extern "C" int gbgpu_wave_ll(double *loglike, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const double *params,
                             int num_params, int num_bin, const gbgpu_wave_cfg *cfg, const gbgpu_data *data,
                             double d_d, int phase_marginalize, void *stream) {
    int rc = check_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if ((rc = check_data(data))) return rc;
    if (!loglike) {
        set_error("loglike must not be null");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    WaveArgs a = base_args(params, num_params, num_bin, cfg);
    set_data(a, data);
    a.d_d = d_d;
    a.phase_marg = phase_marginalize;
    a.loglike = loglike;
    a.d_h = reinterpret_cast<double2 *>(d_h);
    a.h_h = reinterpret_cast<double2 *>(h_h);
    a.tw = twiddles_for(cfg->N, cfg->precision);
    if (!a.tw) return hip_fail(hipErrorOutOfMemory, "twiddles");
    if ((rc = set_transform(cfg, a))) return rc;
    return launch_any<MODE_LL>(cfg->N, num_params == GBGPU_NPARAMS_THIRD, cfg->precision, a,
                              static_cast<hipStream_t>(stream), nullptr, 0, cfg->transform);
}

extern "C" int gbgpu_wave_ll_noalloc(double *loglike, gbgpu_cmplx *d_h, gbgpu_cmplx *h_h, const double *params,
                                    int num_params, int num_bin, const gbgpu_wave_cfg *cfg, const gbgpu_data *data,
                                    double d_d, int phase_marginalize, void *stream) {
    int rc = check_cfg(params, num_params, num_bin, cfg);
    if (rc) return rc;
    if ((rc = check_data(data))) return rc;
    if (!loglike) {
        set_error("loglike must not be null");
        return GBGPU_E_ARG;
    }
    if (num_bin > GBGPU_NOALLOC_MAX_BIN) {
        set_error("gbgpu_wave_ll_noalloc: num_bin exceeds GBGPU_NOALLOC_MAX_BIN (65536); split the batch or use "
                  "gbgpu_wave_ll");
        return GBGPU_E_ARG;
    }
    if (num_bin == 0) return GBGPU_OK;
    char *base = nullptr;
    if ((rc = noalloc_base(&base))) return rc;
    WaveArgs a = base_args(params, num_params, num_bin, cfg);
    set_data(a, data);
    a.d_d = d_d;
    a.phase_marg = phase_marginalize;
    a.loglike = loglike;
    a.d_h = reinterpret_cast<double2 *>(d_h);
    a.h_h = reinterpret_cast<double2 *>(h_h);
    a.tw = cfg->precision == GBGPU_MIXED ? (const void *)(reinterpret_cast<const float2 *>(base + NOALLOC_TW64_BYTES) +
                                                         tw_offset(cfg->N))
                                         : (const void *)(reinterpret_cast<const double2 *>(base) + tw_offset(cfg->N));
    a.dftc = base + NOALLOC_TW64_BYTES + NOALLOC_TW32_BYTES;
    return launch_any<MODE_LL>(cfg->N, num_params == GBGPU_NPARAMS_THIRD, cfg->precision, a,
                               static_cast<hipStream_t>(stream), base + NOALLOC_TW_BYTES, NOALLOC_WS_BYTES,
                               cfg->transform);
}
