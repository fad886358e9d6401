// Device helpers shared by the kernel translation units of liborth_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "orth_kernels.h"

namespace orth {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Group fields by a runtime index without indexing the kernel-argument
// array (which would force a private copy of the arguments).
#define ORTH_G(args_, idx_, fld_) ((idx_) == 0 ? (args_).g[0].fld_ : (args_).g[1].fld_)

// Is Z tile column block nt (16 columns) consumed by some group?
__device__ __forceinline__ bool ztile_used(const PassArgs &a, int nt)
{
    bool u = false;
    for (int g = 0; g < a.ngrp; g++) {
        const int z0 = ORTH_G(a, g, zoff), z1 = z0 + ORTH_G(a, g, w);
        u = u || (nt * 16 < z1 && nt * 16 + 16 > z0);
    }
    return u;
}

// zsteps with the B operand (the coefficients) read from global memory.
__device__ __forceinline__ dbl4 zsteps_g(const double *sm, const double *__restrict__ C, int ao,
                                         int bo, int LD, int N1, int k0, int k1e)
{
    dbl4 z = dbl4{0.0, 0.0, 0.0, 0.0};
    int ks = k0;
    if (k1e - k0 >= 4) {
        double b0 = C[bo + 4 * ks * N1], b1 = C[bo + 4 * (ks + 1) * N1],
               b2 = C[bo + 4 * (ks + 2) * N1], b3 = C[bo + 4 * (ks + 3) * N1];
        double a0 = sm[ao + 4 * ks * LD], a1 = sm[ao + 4 * (ks + 1) * LD],
               a2 = sm[ao + 4 * (ks + 2) * LD], a3 = sm[ao + 4 * (ks + 3) * LD];
        for (; ks + 4 <= k1e; ks += 4) {
            const double c0 = a0, c1 = a1, c2 = a2, c3 = a3;
            const double d0 = b0, d1 = b1, d2 = b2, d3 = b3;
            const int nx = min(ks + 4, k1e - 4);
            b0 = C[bo + 4 * nx * N1];
            b1 = C[bo + 4 * (nx + 1) * N1];
            b2 = C[bo + 4 * (nx + 2) * N1];
            b3 = C[bo + 4 * (nx + 3) * N1];
            a0 = sm[ao + 4 * nx * LD];
            a1 = sm[ao + 4 * (nx + 1) * LD];
            a2 = sm[ao + 4 * (nx + 2) * LD];
            a3 = sm[ao + 4 * (nx + 3) * LD];
            z = __builtin_amdgcn_mfma_f64_16x16x4f64(c0, d0, z, 0, 0, 0);
            z = __builtin_amdgcn_mfma_f64_16x16x4f64(c1, d1, z, 0, 0, 0);
            z = __builtin_amdgcn_mfma_f64_16x16x4f64(c2, d2, z, 0, 0, 0);
            z = __builtin_amdgcn_mfma_f64_16x16x4f64(c3, d3, z, 0, 0, 0);
        }
    }
    for (; ks < k1e; ks++)
        z = __builtin_amdgcn_mfma_f64_16x16x4f64(sm[ao + 4 * ks * LD], C[bo + 4 * ks * N1], z, 0,
                                                 0, 0);
    return z;
}

// ---- reduced-precision MFMA operands (mixed form)

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

// Reduced-precision MFMA operand: 8 values of one lane (k = 8 (l / 16) + e),
// as a high part and, for the split form, the low part FP32(x) - FP32(hi).
template <int MX>
struct LpFrag {
    typedef typename std::conditional<MX == kMxFP16, f16x8, bf16x8>::type V;
    V hi, lo;
};

// Load 8 FP64 values p[0], p[st], ..., p[7 st] from LDS and convert them,
// four at a time (bounded registers).  The values go through FP32 (one
// v_cvt_f32_f64 each, then a packed FP32 -> bf16 / fp16 conversion): the
// FP32 value is kept opaque, since a fused FP64 -> bf16 conversion becomes
// an eight-instruction sequence of FP64 compares and conversions per value,
// which share the FP64 pipeline with the MFMA.  Rounding twice differs from
// rounding once only on ties of the reduced format, and not at all for the
// FP32 input columns.
template <int MX>
__device__ __forceinline__ LpFrag<MX> lp_frag(const double *p, int st)
{
    LpFrag<MX> f;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        float x[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            x[e] = (float)p[(4 * h + e) * st];
            asm volatile("" : "+v"(x[e]));
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            if constexpr (MX == kMxFP16) {
                f.hi[4 * h + e] = (_Float16)x[e];
            } else {
                f.hi[4 * h + e] = (__bf16)x[e];
                if constexpr (MX == kMxBF16X3) f.lo[4 * h + e] = (__bf16)(x[e] - (float)f.hi[4 * h + e]);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    return f;
}

// The same from global memory with a validity mask (coefficients past k1).
template <int MX>
__device__ __forceinline__ LpFrag<MX> lp_frag_g(const double *p, int st, int nvalid)
{
    LpFrag<MX> f;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        float x[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            x[e] = 4 * h + e < nvalid ? (float)p[(4 * h + e) * st] : 0.0f;
            asm volatile("" : "+v"(x[e]));
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            if constexpr (MX == kMxFP16) {
                f.hi[4 * h + e] = (_Float16)x[e];
            } else {
                f.hi[4 * h + e] = (__bf16)x[e];
                if constexpr (MX == kMxBF16X3) f.lo[4 * h + e] = (__bf16)(x[e] - (float)f.hi[4 * h + e]);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    return f;
}

// One 16 x 16 x 32 product in the reduced precision, accumulated into t
// (FP32): lane l holds rows 4 (l / 16) + v of column l % 16 of the result.
template <int MX>
__device__ __forceinline__ f32x4 lp_mma(const LpFrag<MX> &a, const LpFrag<MX> &b, f32x4 t)
{
    if constexpr (MX == kMxFP16) {
        t = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.hi, t, 0, 0, 0);
    } else {
        if constexpr (MX == kMxBF16X3) {
            t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, t, 0, 0, 0);
            t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, t, 0, 0, 0);
        }
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, t, 0, 0, 0);
    }
    return t;
}

// Is Z tile nt consumed by a group whose product is marked reduced-precision?
__device__ __forceinline__ bool ztile_lp(const PassArgs &a, int nt)
{
    bool u = false;
    for (int g = 0; g < a.ngrp; g++) {
        const int z0 = ORTH_G(a, g, zoff), z1 = z0 + ORTH_G(a, g, w);
        u = u || (ORTH_G(a, g, lp) && nt * 16 < z1 && nt * 16 + 16 > z0);
    }
    return u;
}

}  // namespace orth
