// Shared device helpers and host plumbing for the gfx950 FastGB library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "gbgpu_hip.h"

namespace gbgpu {

constexpr double C_SI = 299792458.0;
constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double TWO_PI = 6.283185307179586476925286766559005768;
constexpr double INV_TWO_PI = 0.159154943091895335768883763372514362;
constexpr double INV_SQRT2 = 0.707106781186547524400844362104849039;
constexpr double INV_SQRT3 = 0.577350269189625764509148780501957456;
constexpr double INV_SQRT6 = 0.408248290463863016366214012450981899;

// ---------------------------------------------------------------- complex ---
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i and +i
__device__ __forceinline__ double2 cmul_mi(double2 a) { return make_double2(a.y, -a.x); }
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }
__device__ __forceinline__ double2 cmul_pi(double2 a) { return make_double2(-a.y, a.x); }
__device__ __forceinline__ float2 cmul_pi(float2 a) { return make_float2(-a.y, a.x); }

template <class C>
struct real_of;
template <>
struct real_of<double2> {
    using type = double;
};
template <>
struct real_of<float2> {
    using type = float;
};

// ------------------------------------------------------ wavefront reduction ---
// Sum over aligned groups of G lanes (G a power of two <= 64) with xor
// shuffles; every lane of the group ends with the group sum.
template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------ host plumbing ---
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *where);

// Device twiddle table w[k] = exp(-2 pi i k / N), k < N, cached per (device, N).
const double2 *twiddles_f64(int N);
const float2 *twiddles_f32(int N);
// The same tables written to host memory (N entries).
void host_twiddles_f64(int N, double2 *out);
void host_twiddles_f32(int N, float2 *out);

inline bool is_pow2(long long n) { return n > 0 && (n & (n - 1)) == 0; }

// Optional timing of the dominant kernel of each call (gbgpu_kernel_timing):
// HIP events recorded on the launch stream immediately around the launch.
bool timing_enabled();
void timing_push(hipEvent_t e0, hipEvent_t e1);

struct KernelTimer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t st;
    explicit KernelTimer(hipStream_t s) : st(s) {
        if (!timing_enabled()) return;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
            hipEventRecord(e0, st) != hipSuccess)
            e0 = nullptr;
    }
    ~KernelTimer() {
        if (e0 && hipEventRecord(e1, st) == hipSuccess) timing_push(e0, e1);
    }
};

}  // namespace gbgpu

#define GBGPU_CHECK_LAUNCH(where)                                   \
    do {                                                            \
        hipError_t _e = hipGetLastError();                          \
        if (_e != hipSuccess) return ::gbgpu::hip_fail(_e, where);  \
    } while (0)
