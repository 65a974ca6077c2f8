// Host-side runtime of the library: per-thread error messages and the
// per-(device, N) twiddle caches.  No device synchronisation happens here
// except the one-time blocking upload of a new twiddle table.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "dft_mfma.hpp"

namespace gbgpu {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char *where) {
    set_error(std::string(where) + ": " + hipGetErrorString(e));
    return GBGPU_E_HIP;
}

namespace {
template <class C>
void host_twiddles(int N, C *out) {
    for (int k = 0; k < N; k++) {
        // long double keeps the table correctly rounded in float64
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)N;
        out[k].x = (decltype(out[k].x))std::cos(ang);
        out[k].y = (decltype(out[k].y))std::sin(ang);
    }
}

std::mutex g_tw_mutex;
std::map<std::pair<int, int>, void *> g_tw64, g_tw32;

template <class C>
const C *get_twiddles(std::map<std::pair<int, int>, void *> &cache, int N) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    auto key = std::make_pair(dev, N);
    auto it = cache.find(key);
    if (it != cache.end()) return static_cast<const C *>(it->second);
    std::vector<C> host(N);
    host_twiddles(N, host.data());
    void *dptr = nullptr;
    if (hipMalloc(&dptr, sizeof(C) * N) != hipSuccess) return nullptr;
    if (hipMemcpy(dptr, host.data(), sizeof(C) * N, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(dptr);
        return nullptr;
    }
    cache[key] = dptr;
    return static_cast<const C *>(dptr);
}
}  // namespace

namespace {
uint32_t f32_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
// bf16 nearest to x (ties to even), as the upper half of a float
uint16_t bf16_rne(double x) {
    const uint32_t u = f32_bits((float)x);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
double bf16_val(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

std::mutex g_dft_mutex;
std::map<int, void *> g_dft;
}  // namespace

void host_dft256_table(uint32_t *out) {
    constexpr long double TWO_PI_L = 6.283185307179586476925286766559005768L;
    auto row = [&](int r, int lane) { return out + (r * 64 + lane) * 4; };
    for (int lane = 0; lane < 64; lane++) {
        const int g = lane >> 4, c = lane & 15;
        double cre[8], cim[8];
        for (int j = 0; j < 8; j++) {
            const int e = ((4 * g + (j & 3)) * c) & 15;
            const double fr = (double)std::cos(-TWO_PI_L * e / 16), fi = (double)std::sin(-TWO_PI_L * e / 16);
            cre[j] = j < 4 ? fr : -fi;
            cim[j] = j < 4 ? fi : fr;
        }
        float tw[8];
        for (int i = 0; i < 4; i++) {
            const int e = (c * (4 * g + i)) & 255;
            tw[2 * i] = (float)std::cos(-TWO_PI_L * e / 256);
            tw[2 * i + 1] = (float)std::sin(-TWO_PI_L * e / 256);
        }
        // FP32 variant: Cre, Cim, twiddles
        for (int j = 0; j < 8; j++) {
            row(DFT_QOFF[1] + j / 4, lane)[j % 4] = f32_bits((float)cre[j]);
            row(DFT_QOFF[1] + 2 + j / 4, lane)[j % 4] = f32_bits((float)cim[j]);
        }
        // bf16 variants: pieces of each constant, 8 bf16 per row (element j in
        // the low half of dword j/2 for even j, the high half for odd j)
        for (int v = 2; v <= 3; v++) {
            const int P = v == 2 ? 2 : 3;
            for (int part = 0; part < 2; part++) {
                const double *src = part == 0 ? cre : cim;
                for (int j = 0; j < 8; j++) {
                    double r = src[j];
                    for (int q = 0; q < P; q++) {
                        const uint16_t h = bf16_rne(r);
                        r -= bf16_val(h);
                        uint32_t &w = row(DFT_QOFF[v] + part * P + q, lane)[j / 2];
                        if (j % 2 == 0)
                            w = (w & 0xffff0000u) | h;
                        else
                            w = (w & 0x0000ffffu) | ((uint32_t)h << 16);
                    }
                }
            }
        }
        for (int v = 1; v <= 3; v++)
            for (int i = 0; i < 8; i++) row(DFT_QOFF[v] + DFT_QN[v] - 2 + i / 4, lane)[i % 4] = f32_bits(tw[i]);
    }
}

const void *dft256_table() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_dft_mutex);
    auto it = g_dft.find(dev);
    if (it != g_dft.end()) return it->second;
    std::vector<uint32_t> host(DFT_TABLE_ROWS * 64 * 4, 0u);
    host_dft256_table(host.data());
    void *dptr = nullptr;
    const size_t bytes = host.size() * 4;
    if (hipMalloc(&dptr, bytes) != hipSuccess) return nullptr;
    if (hipMemcpy(dptr, host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(dptr);
        return nullptr;
    }
    g_dft[dev] = dptr;
    return dptr;
}

void host_twiddles_f64(int N, double2 *out) { host_twiddles(N, out); }
void host_twiddles_f32(int N, float2 *out) { host_twiddles(N, out); }

const double2 *twiddles_f64(int N) { return get_twiddles<double2>(g_tw64, N); }
const float2 *twiddles_f32(int N) { return get_twiddles<float2>(g_tw32, N); }

namespace {
std::mutex g_timing_mutex;
bool g_timing = false;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_events;
}  // namespace

bool timing_enabled() {
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    return g_timing;
}

void timing_push(hipEvent_t e0, hipEvent_t e1) {
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    g_events.emplace_back(e0, e1);
}

}  // namespace gbgpu

using namespace gbgpu;

extern "C" int gbgpu_kernel_timing(int enable) {
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    for (auto &ev : g_events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    g_events.clear();
    g_timing = enable != 0;
    return GBGPU_OK;
}

extern "C" int gbgpu_kernel_timing_read(double *total_ms, int *count) {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
    {
        std::lock_guard<std::mutex> lock(g_timing_mutex);
        evs.swap(g_events);
    }
    double tot = 0.0;
    int rc = GBGPU_OK;
    for (auto &ev : evs) {
        float ms = 0.0f;
        hipError_t e = hipEventSynchronize(ev.second);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev.first, ev.second);
        if (e != hipSuccess && rc == GBGPU_OK) rc = hip_fail(e, "kernel timing");
        tot += ms;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    if (total_ms) *total_ms = tot;
    if (count) *count = (int)evs.size();
    return rc;
}

extern "C" const char *gbgpu_last_error(void) { return g_last_error.c_str(); }

extern "C" int gbgpu_abi_version(void) { return 9; }
