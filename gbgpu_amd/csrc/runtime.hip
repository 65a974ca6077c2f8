// Host-side runtime of the library: per-thread error messages and the
// per-(device, N) twiddle caches.  No device synchronisation happens here
// except the one-time blocking upload of a new twiddle table.
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

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

extern "C" int gbgpu_abi_version(void) { return 4; }
