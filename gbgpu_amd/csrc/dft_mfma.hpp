// Length-256 forward DFT of one binary by one wavefront on the matrix cores
// (the alternative to fft_lds for the mixed policy at N = 256).
//
// Four-step factorisation N = 16 x 16 with n = 16 n1 + n2, k = k1 + 16 k2:
//   Y[k1, n2] = W256^(k1 n2) sum_n1 F[k1, n1] x[16 n1 + n2]
//   X[k1 + 16 k2] = sum_n2 F[k2, n2] Y[k1, n2],      F[a, b] = W16^(a b),
// W_M = exp(-2 pi i / M) (numpy's forward sign, unnormalised).  Both steps
// are complex 16 x 16 x 16 products, written as real products with K = 32
// by stacking real and imaginary parts along K, and both are evaluated in
// transposed form (Y^T = X^T F, X^T = F Y^T): with the lane layouts of the
// 16 x 16 MFMA tiles, lane l = c + 16 g (c, g < 16, 4) then holds
//   before step 1: x[c + 64 g + 16 i],          i = 0..3  (A operand row c)
//   after  step 1: Y[c, 4 g + i]                          (D rows 4g+i, col c)
//   before step 2: the same values                        (B operand col c)
//   after  step 2: X[c + 64 g + 16 i]                     (D rows 4g+i, col c)
// so no value crosses lanes and nothing goes through LDS.  The K slots a
// lane feeds are s = 8 g + j: j < 4 is Re of element 4 g + j, j >= 4 is Im
// of element 4 g + j - 4.  F is symmetric, so the constant operand of step 1
// (B) and of step 2 (A) hold the same per-lane values:
//   Cre[j] = j < 4 ? Re F[4g+j, c] : -Im F[4g+j-4, c]   (gives Re of the product)
//   Cim[j] = j < 4 ? Im F[4g+j, c] :  Re F[4g+j-4, c]   (gives Im of the product)
//
// Variants (GBGPU_TRANSFORM_*):
//   MFMA_F64     v_mfma_f64_16x16x4_f64, 8 per real output tile: FP64
//                products and accumulation (the FP64 policy).
//   MFMA_F32     v_mfma_f32_16x16x4_f32, 8 per real output tile: FP32 products.
//   MFMA_BF16X3  v_mfma_f32_16x16x32_bf16 on a two-piece bf16 split of data
//                and constants (v = hi + lo), three products hi.hi + hi.lo +
//                lo.hi: about 16 significant bits.
//   MFMA_BF16X6  three-piece split (hi + mid + lo), six products down to the
//                2^-16-relative terms: FP32-level accuracy.
// FP32 accumulation in the FP32 and bf16 variants, smallest products first.
#pragma once

#include "common.hpp"

namespace gbgpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Per-lane constant table: for variant v, QN[v] rows of 64 lanes x 16 bytes
// starting at row QOFF[v]; lane l reads rows q = 0..QN-1 at [q * 64 + l]
// (coalesced).  Rows: the Cre pieces, the Cim pieces (F32: two rows of four
// floats each), then two rows of step twiddles W256^(c (4 g + i)), i = 0..3,
// as (re, im) pairs.
constexpr int DFT_QN[4] = {0, 6, 6, 8};
constexpr int DFT_QOFF[4] = {0, 0, 6, 12};
constexpr int DFT_TABLE_ROWS = 20;  // x 64 lanes x 16 bytes = 20 KiB

// Host: writes the whole table (DFT_TABLE_ROWS * 64 * 4 dwords).
void host_dft256_table(uint32_t *out);
// Device copy of the table, cached per device (nullptr on failure).
const void *dft256_table();

template <int V>
struct Dft256Lane;

template <>
struct Dft256Lane<GBGPU_TRANSFORM_MFMA_F32> {
    float cre[8], cim[8];
    float2 tw[4];
};
template <>
struct Dft256Lane<GBGPU_TRANSFORM_MFMA_BF16X3> {
    bf16x8 cre[2], cim[2];  // hi, lo
    float2 tw[4];
};
template <>
struct Dft256Lane<GBGPU_TRANSFORM_MFMA_BF16X6> {
    bf16x8 cre[3], cim[3];  // hi, mid, lo
    float2 tw[4];
};

// FP64 (v_mfma_f64_16x16x4f64).  Its A and B operand layouts are those of
// the FP32 16 x 16 x 4 instruction, but lane c + 16 g holds rows g + 4 r of
// its 16 x 16 result (r = 0..3; rows 4 g + r for FP32; measured).  With the
// element index e = g + 4 t for K slot t (t < 4: Re, else Im of element
// e(g, t - 4)) in both steps, lane l then holds sample l + 64 i in register
// i before the transform and bin l + 64 i after it: the natural strided
// order of the LDS transform's register ends.  The constants come straight
// from the FP64 twiddle table of N = 256, w[k] = W256^k, since
// F[a, b] = W16^(a b) = w[16 a b mod 256], and the step twiddles are
// W256^(c (g + 4 i)).
template <>
struct Dft256Lane<GBGPU_TRANSFORM_MFMA_F64> {
    double cre[8], cim[8];
    double2 tw[4];
};

__device__ __forceinline__ Dft256Lane<GBGPU_TRANSFORM_MFMA_F64> load_dft256_f64(const double2 *__restrict__ w,
                                                                              int lane) {
    Dft256Lane<GBGPU_TRANSFORM_MFMA_F64> k;
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const double2 f = w[(16 * (g + 4 * j) * c) & 255];  // F[g+4j, c]
        k.cre[j] = f.x;
        k.cim[j] = f.y;
        k.cre[4 + j] = -f.y;
        k.cim[4 + j] = f.x;
        k.tw[j] = w[(c * (g + 4 * j)) & 255];  // W256^(c (g+4j))
    }
    return k;
}

template <int V>
__device__ __forceinline__ Dft256Lane<V> load_dft256(const void *table, int lane) {
    const u32x4 *t = static_cast<const u32x4 *>(table) + DFT_QOFF[V] * 64 + lane;
    Dft256Lane<V> k;
    if constexpr (V == GBGPU_TRANSFORM_MFMA_F32) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            // whole-vector casts: a bit cast of one element of a vector reads
            // element 0 with this compiler
            const f32x4 fr = __builtin_bit_cast(f32x4, t[q * 64]);
            const f32x4 fi = __builtin_bit_cast(f32x4, t[(2 + q) * 64]);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                k.cre[4 * q + e] = fr[e];
                k.cim[4 * q + e] = fi[e];
            }
        }
    } else {
        constexpr int P = V == GBGPU_TRANSFORM_MFMA_BF16X3 ? 2 : 3;
#pragma unroll
        for (int q = 0; q < P; q++) {
            k.cre[q] = __builtin_bit_cast(bf16x8, t[q * 64]);
            k.cim[q] = __builtin_bit_cast(bf16x8, t[(P + q) * 64]);
        }
    }
    constexpr int QT = DFT_QN[V] - 2;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const f32x4 u = __builtin_bit_cast(f32x4, t[(QT + q) * 64]);
        k.tw[2 * q] = make_float2(u[0], u[1]);
        k.tw[2 * q + 1] = make_float2(u[2], u[3]);
    }
    return k;
}

// v = p[0] + p[1] + ... (each piece the bf16 rounding of what the earlier
// pieces leave).
template <int P>
__device__ __forceinline__ void split_bf16(const float (&v)[8], bf16x8 (&p)[P]) {
    float r[8];
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = v[j];
#pragma unroll
    for (int k = 0; k < P; k++) {
#pragma unroll
        for (int j = 0; j < 8; j++) p[k][j] = (__bf16)r[j];
        if (k + 1 < P) {
#pragma unroll
            for (int j = 0; j < 8; j++) r[j] -= (float)p[k][j];
        }
    }
}

template <bool DATA_IS_A>
__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 d, bf16x8 c, f32x4 acc) {
    return DATA_IS_A ? __builtin_amdgcn_mfma_f32_16x16x32_bf16(d, c, acc, 0, 0, 0)
                     : __builtin_amdgcn_mfma_f32_16x16x32_bf16(c, d, acc, 0, 0, 0);
}

// re/im tiles of the complex product of the lane's K-slot values v[8] with
// the constant operand (step 1: data is A; step 2: data is B).
template <int V, bool DATA_IS_A>
__device__ __forceinline__ void cgemm16(const float (&v)[8], const Dft256Lane<V> &k, f32x4 &re, f32x4 &im) {
    re = f32x4{0.f, 0.f, 0.f, 0.f};
    im = f32x4{0.f, 0.f, 0.f, 0.f};
    if constexpr (V == GBGPU_TRANSFORM_MFMA_F32) {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if constexpr (DATA_IS_A) {
                re = __builtin_amdgcn_mfma_f32_16x16x4f32(v[t], k.cre[t], re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_16x16x4f32(v[t], k.cim[t], im, 0, 0, 0);
            } else {
                re = __builtin_amdgcn_mfma_f32_16x16x4f32(k.cre[t], v[t], re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_16x16x4f32(k.cim[t], v[t], im, 0, 0, 0);
            }
        }
    } else if constexpr (V == GBGPU_TRANSFORM_MFMA_BF16X3) {
        bf16x8 d[2];
        split_bf16<2>(v, d);
        re = mfma_bf16<DATA_IS_A>(d[1], k.cre[0], re);  // lo.hi
        im = mfma_bf16<DATA_IS_A>(d[1], k.cim[0], im);
        re = mfma_bf16<DATA_IS_A>(d[0], k.cre[1], re);  // hi.lo
        im = mfma_bf16<DATA_IS_A>(d[0], k.cim[1], im);
        re = mfma_bf16<DATA_IS_A>(d[0], k.cre[0], re);  // hi.hi
        im = mfma_bf16<DATA_IS_A>(d[0], k.cim[0], im);
    } else {
        bf16x8 d[3];
        split_bf16<3>(v, d);
        re = mfma_bf16<DATA_IS_A>(d[2], k.cre[0], re);  // lo.hi
        im = mfma_bf16<DATA_IS_A>(d[2], k.cim[0], im);
        re = mfma_bf16<DATA_IS_A>(d[1], k.cre[1], re);  // mid.mid
        im = mfma_bf16<DATA_IS_A>(d[1], k.cim[1], im);
        re = mfma_bf16<DATA_IS_A>(d[0], k.cre[2], re);  // hi.lo
        im = mfma_bf16<DATA_IS_A>(d[0], k.cim[2], im);
        re = mfma_bf16<DATA_IS_A>(d[1], k.cre[0], re);  // mid.hi
        im = mfma_bf16<DATA_IS_A>(d[1], k.cim[0], im);
        re = mfma_bf16<DATA_IS_A>(d[0], k.cre[1], re);  // hi.mid
        im = mfma_bf16<DATA_IS_A>(d[0], k.cim[1], im);
        re = mfma_bf16<DATA_IS_A>(d[0], k.cre[0], re);  // hi.hi
        im = mfma_bf16<DATA_IS_A>(d[0], k.cim[0], im);
    }
}

// FP64 products, FP64 accumulation.
template <bool DATA_IS_A>
__device__ __forceinline__ void cgemm16_f64(const double (&v)[8], const Dft256Lane<GBGPU_TRANSFORM_MFMA_F64> &k,
                                            f64x4 &re, f64x4 &im) {
    re = f64x4{0.0, 0.0, 0.0, 0.0};
    im = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 8; t++) {
        if constexpr (DATA_IS_A) {
            re = __builtin_amdgcn_mfma_f64_16x16x4f64(v[t], k.cre[t], re, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f64_16x16x4f64(v[t], k.cim[t], im, 0, 0, 0);
        } else {
            re = __builtin_amdgcn_mfma_f64_16x16x4f64(k.cre[t], v[t], re, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f64_16x16x4f64(k.cim[t], v[t], im, 0, 0, 0);
        }
    }
}

// In place: x[i] = element lane + 64 i of the binary's 256 samples on entry,
// of its DFT on exit.  Every lane of the wavefront must call this together.
template <int V>
__device__ __forceinline__ void dft256_mfma(double2 *x, const Dft256Lane<V> &k) {
    static_assert(V == GBGPU_TRANSFORM_MFMA_F64, "FP64 spectra take the FP64 matrix transform");
    double v[8];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i] = x[i].x;
        v[4 + i] = x[i].y;
    }
    f64x4 re, im;
    cgemm16_f64<true>(v, k, re, im);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double2 y = cmul(make_double2(re[i], im[i]), k.tw[i]);
        v[i] = y.x;
        v[4 + i] = y.y;
    }
    cgemm16_f64<false>(v, k, re, im);
#pragma unroll
    for (int i = 0; i < 4; i++) x[i] = make_double2(re[i], im[i]);
}

// In place: x[i] = element c + 64 g + 16 i of the binary's 256 samples on
// entry, of its DFT on exit (lane = c + 16 g).  Every lane of the wavefront
// must call this together.
template <int V>
__device__ __forceinline__ void dft256_mfma(float2 *x, const Dft256Lane<V> &k) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i] = x[i].x;
        v[4 + i] = x[i].y;
    }
    f32x4 re, im;
    cgemm16<V, true>(v, k, re, im);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float2 y = cmul(make_float2(re[i], im[i]), k.tw[i]);
        v[i] = y.x;
        v[4 + i] = y.y;
    }
    cgemm16<V, false>(v, k, re, im);
#pragma unroll
    for (int i = 0; i < 4; i++) x[i] = make_float2(re[i], im[i]);
}

// Sample / output index held in register i of lane `lane` (FP32 and bf16
// variants; the FP64 variant holds lane + 64 i).
__device__ __forceinline__ int dft256_index(int lane, int i) { return (lane & 15) + 64 * (lane >> 4) + 16 * i; }

}  // namespace gbgpu
