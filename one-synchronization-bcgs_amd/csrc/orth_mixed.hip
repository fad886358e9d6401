// Mixed-precision fused pass for block width 16 (gfx950).
//
// Same schedule as the FP64 s = 16 pass (stage a 32-row tile of the slab in
// LDS, apply the pending block updates, write the updated blocks, accumulate
// the next projection), with two differences:
//   * columns that no earlier pass has written are read from the caller's
//     FP32 input X32 (converted exactly to FP64 in the tile); every written
//     column goes to the FP64 output Q;
//   * the products marked by the host -- the first-pass update of a block
//     (X - Q S) and the first-pass projection Q^T X of the next block with
//     its Gram X^T X -- run on bf16 / fp16 MFMA (v_mfma_f32_16x16x32_*) with
//     FP32 accumulation per 32-row tile, summed in FP64 across tiles.
//     The split bf16x3 form represents each FP32 value as hi + lo (two bf16)
//     and adds hi*lo' + lo*hi' + hi*hi'.
// Everything that the reorthogonalization depends on stays FP64: the
// second-pass update and projection, the Pythagorean subtraction, the
// Cholesky factors and the triangular solves.
#include "orth_kernels.h"

#include <type_traits>

#include "orth_device.h"
#include "orth_pass16b.h"
#include "orth_pass16d.h"

namespace orth {

constexpr int kM = 512;    // threads per workgroup
constexpr int kMPF = 16;   // prefetch registers per thread (32 rows x 256 columns)

template <int NB, int MX>
__global__ __launch_bounds__(kM) __attribute__((amdgpu_waves_per_eu(MX == kMxFP64 ? 4 : 2, MX == kMxFP64 ? 4 : 2))) void k_pass16m(PassArgs a)
{
    constexpr int N1 = 32, TS = 32, LD = TS + 1, PSTEP = kM / TS;
    extern __shared__ double sm[];
    const uint64_t kst = kernarg_stash();
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int k1p = (a.k1 + 3) & ~3;
    const int R0 = a.wl * LD;  // step-1 half-sum scratch (4 x 256 doubles)
    const double *__restrict__ Cg = a.C;
    for (int e = tid; e < (a.wl - a.ncols) * TS; e += kM)
        sm[(a.ncols + (e >> 5)) * LD + (e & (TS - 1))] = 0.0;

    // Step 2: wave w owns projection M-tiles w and w + 8 for every B tile.
    const int Ma = (a.na + 15) >> 4;
    const bool proj = a.na > 0;
    bool lpt[NB];
#pragma unroll
    for (int tb = 0; tb < NB; tb++) lpt[tb] = MX != kMxFP64 && a.b0 + 16 * tb >= a.lpb;
    int acol[2];
#pragma unroll
    for (int j = 0; j < 2; j++) acol[j] = a.a0 + min(wave + 8 * j, max(Ma - 1, 0)) * 16 + li;
    dbl4 acc[2][NB];
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int tb = 0; tb < NB; tb++) acc[j][tb] = dbl4{0.0, 0.0, 0.0, 0.0};

    // Triangular-solve operands of this wave's group (as in the FP64 pass).
    const int gq = wave & 1;
    const double *Tq = gq < a.ngrp ? ORTH_G(a, gq, T) : nullptr;
    double bw[4];
#pragma unroll
    for (int b = 0; b < 4; b++) bw[b] = Tq ? Tq[kSlotW + 128 * b + lk * kLdT + li] : 0.0;

    const int64_t rb = (int64_t)blockIdx.x * a.rows_per;
    const int64_t re = min(a.m, rb + a.rows_per);
    const int prow = tid & (TS - 1), pcol = tid >> 5;
    const int nld = (a.ncols - pcol + PSTEP - 1) / PSTEP;
    const int pl = pcol * LD + prow;
    const double *px = a.X + prow + (int64_t)pcol * a.ldX;
    const float *px32 = a.X32 + prow + (int64_t)pcol * a.ldX32;
    const int64_t pxs = (int64_t)PSTEP * a.ldX, pxs32 = (int64_t)PSTEP * a.ldX32;
    // raw0 is a multiple of 16 (a block boundary) or past the tile, so column
    // pcol + 16 i is raw exactly when i >= raw0 / 16: a uniform choice.
    const int iraw = a.raw0 >> 4;
    double pf[kMPF];
    auto load = [&](int64_t r0) {
        const bool ok = r0 + prow < re;
        const double *q = px + r0;
        const float *q32 = px32 + r0;
#pragma unroll
        for (int i = 0; i < kMPF; i++)
            if (i < nld) {
                if (i >= iraw)
                    pf[i] = ok ? (double)__builtin_nontemporal_load(q32 + i * pxs32) : 0.0;
                else
                    pf[i] = ok ? __builtin_nontemporal_load(q + i * pxs) : 0.0;
            }
    };
    if (rb < re) load(rb);

    for (int64_t r0 = rb; r0 < re; r0 += TS) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kMPF; i++)
            if (i < nld) sm[pl + i * PSTEP * LD] = pf[i];
        __syncthreads();

        if (a.ngrp > 0) {
            const int nphase = a.seq ? a.ngrp : 1;
            for (int p = 0; p < nphase; p++) {
                const int glo = a.seq ? p : 0, ghi = a.seq ? p + 1 : a.ngrp;
                const int kp = a.seq ? ((ORTH_G(a, p, k) + 3) & ~3) : k1p;
                const int nks = a.k1 > 0 ? kp >> 2 : 0;
                // Step 1: Z tile (wave & 3) over half (wave >> 2) of the inner
                // dimension, halves added in a fixed order through LDS.
                const int item = wave & 3, mt = item & 1, nt = item >> 1, kh = wave >> 2;
                const bool mine = nks > 0 && (a.seq ? nt == (ORTH_G(a, p, zoff) >> 4)
                                                    : ztile_used(a, nt));
                const bool zlp = MX != kMxFP64 && !a.seq && ztile_lp(a, nt);
                dbl4 z = dbl4{0.0, 0.0, 0.0, 0.0};
                if (mine) {
                    if (zlp) {
                        // K in chunks of 32 columns; coefficient rows past k1
                        // are zero, tile columns past the tile are clamped.
                        const int nch = (a.k1 + 31) >> 5, half = nch >> 1;
                        f32x4 t = f32x4{0.f, 0.f, 0.f, 0.f};
                        for (int q = kh ? half : 0; q < (kh ? nch : half); q++) {
                            const int k0 = 32 * q + 8 * lk;
                            const LpFrag<MX> fa =
                                lp_frag<MX>(sm + min(k0, a.wl - 8) * LD + mt * 16 + li, LD);
                            const LpFrag<MX> fb =
                                lp_frag_g<MX>(Cg + k0 * N1 + nt * 16 + li, N1, a.k1 - k0);
                            t = lp_mma<MX>(fa, fb, t);
                        }
                        z = dbl4{t[0], t[1], t[2], t[3]};
                    } else {
                        const int half = nks >> 1;
                        z = zsteps_g(sm, Cg, lk * LD + mt * 16 + li, lk * N1 + nt * 16 + li, LD,
                                     N1, kh ? half : 0, kh ? nks : half);
                    }
                    if (kh) {
#pragma unroll
                        for (int v = 0; v < 4; v++) sm[R0 + item * 256 + v * 64 + lane] = z[v];
                    }
                }
                __syncthreads();
                if (mine && !kh) {
#pragma unroll
                    for (int v = 0; v < 4; v++) z[v] += sm[R0 + item * 256 + v * 64 + lane];
                    const int zc = nt * 16 + li;
                    for (int g = glo; g < ghi; g++) {
                        const int off = zc - ORTH_G(a, g, zoff);
                        if (off >= 0 && off < ORTH_G(a, g, w)) {
                            const int o = (ORTH_G(a, g, t0) + off) * LD + mt * 16;
#pragma unroll
                            for (int v = 0; v < 4; v++)
                                sm[o + (zlp ? 4 * lk + v : lk + 4 * v)] -= z[v];
                        }
                    }
                }
                __syncthreads();
                // Triangular solve X <- X T^{-1}, blocked by four columns on
                // FP64 MFMA with the W_b row blocks (see the FP64 pass); wave
                // w takes group (w & 1) and 16-row chunk w >> 1.
                if (Tq && gq >= glo && gq < ghi && (wave >> 1) < TS / 16) {
                    const int t0 = ORTH_G(a, gq, t0), w = ORTH_G(a, gq, w);
                    const int ch = wave >> 1;
                    const int base = (t0 + li) * LD + ch * 16 + lk;
                    const int abase = t0 * LD + ch * 16 + li + lk * LD;
                    dbl4 x;
#pragma unroll
                    for (int v = 0; v < 4; v++) x[v] = sm[base + 4 * v];
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const bool mineb = (li >> 2) == b && li < w;
                        if (mineb && b > 0) {
#pragma unroll
                            for (int v = 0; v < 4; v++) sm[base + 4 * v] = x[v];
                        }
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                        const double av = sm[abase + 4 * b * LD];
                        if ((li >> 2) == b) x = dbl4{0.0, 0.0, 0.0, 0.0};
                        x = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bw[b], x, 0, 0, 0);
                    }
                    if (li < w) {
#pragma unroll
                        for (int v = 0; v < 4; v++) sm[base + 4 * v] = x[v];
                    }
                }
                __syncthreads();
            }
            if (r0 + prow < re) {
                for (int g = 0; g < a.ngrp; g++) {
                    const int t0 = ORTH_G(a, g, t0);
                    for (int cc = pcol; cc < ORTH_G(a, g, w); cc += PSTEP)
                        __builtin_nontemporal_store(sm[(t0 + cc) * LD + prow],
                                                    a.X + r0 + prow + (int64_t)(t0 + cc) * a.ldX);
                }
            }
        }
        if (r0 + TS < re) load(r0 + TS);

        if (proj) {
#pragma unroll
            for (int tb = 0; tb < NB; tb++) {
                if (16 * tb >= a.nb) continue;
                const int bcol = a.b0 + 16 * tb + li;
                if (lpt[tb]) {
                    const LpFrag<MX> fb = lp_frag<MX>(sm + bcol * LD + 8 * lk, 1);
#pragma unroll
                    for (int j = 0; j < 2; j++) {
                        if (wave + 8 * j >= Ma) continue;
                        const LpFrag<MX> fa = lp_frag<MX>(sm + acol[j] * LD + 8 * lk, 1);
                        const f32x4 t = lp_mma<MX>(fa, fb, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                        for (int v = 0; v < 4; v++) acc[j][tb][v] += (double)t[v];
                    }
                    continue;
                }
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    if (wave + 8 * j >= Ma) continue;
                    {
#pragma unroll
                        for (int ks = 0; ks < TS / 4; ks++) {
                            acc[j][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                sm[acol[j] * LD + lk + 4 * ks], sm[bcol * LD + lk + 4 * ks],
                                acc[j][tb], 0, 0, 0);
                            if (ks & 1) __builtin_amdgcn_sched_barrier(0);
                        }
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
    }

    if (proj) {
        double *out = a.part + (int64_t)blockIdx.x * a.na * a.nb;
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int ta = wave + 8 * j;
            if (ta >= Ma) continue;
#pragma unroll
            for (int tb = 0; tb < NB; tb++) {
                const int col = 16 * tb + li;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = ta * 16 + (lpt[tb] ? 4 * lk + v : lk + 4 * v);
                    if (row < a.na && col < a.nb) out[row + (int64_t)col * a.na] = acc[j][tb][v];
                }
            }
        }
    }
    pass_tail(a, sm, kst, kM);
}

size_t pass16m_lds_bytes(const PassArgs &a)
{
    if (a.n1 != 32 || a.ncols * 32 > kMPF * kM || a.nb > 32 || a.na > 256) return 0;
    if (a.na > 0 && (a.a0 + ((a.na + 15) / 16) * 16 > a.wl || a.b0 + ((a.nb + 15) / 16) * 16 > a.wl))
        return 0;
    for (int g = 0; g < a.ngrp; g++)
        if (a.g[g].w > 16 || a.g[g].t0 + a.g[g].w > a.ncols) return 0;
    const size_t b = ((size_t)a.wl * 33 + 1024) * sizeof(double);
    return b <= 160 * 1024 ? b : 0;
}

template <int NB, int MX>
hipError_t launch_pass16m_t(const PassArgs &a, size_t lds, int nwg, hipStream_t st)
{
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)k_pass16m<NB, MX>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(160 * 1024));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    k_pass16m<NB, MX><<<dim3(nwg), dim3(kM), pass_tail_lds(a, lds), st>>>(a);
    return hipGetLastError();
}

template <int MX>
hipError_t launch_pass16m_mx(const PassArgs &a, size_t lds, int nwg, hipStream_t st)
{
    return a.nb > 16 ? launch_pass16m_t<2, MX>(a, lds, nwg, st)
                     : launch_pass16m_t<1, MX>(a, lds, nwg, st);
}

hipError_t launch_pass16m(const PassArgs &a, int mx, int nwg, hipStream_t st)
{
    const size_t lds = pass16m_lds_bytes(a);
    if (lds == 0 || !a.X32) return hipErrorInvalidValue;
    switch (mx) {
    case kMxBF16: return launch_pass16m_mx<kMxBF16>(a, lds, nwg, st);
    case kMxFP16: return launch_pass16m_mx<kMxFP16>(a, lds, nwg, st);
    case kMxBF16X3: return launch_pass16m_mx<kMxBF16X3>(a, lds, nwg, st);
    case kMxFP64: return launch_pass16m_mx<kMxFP64>(a, lds, nwg, st);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// The mixed form on the 512-thread pass (k_pass16b with MX = the precision).
// ---------------------------------------------------------------------------

int pass16bm_plan(const PassArgs &a)
{
    int mt = 0, nb = 0;
    const int ts = pass16b_plan(a, &mt, &nb);
    if (ts == 0 || !a.X32) return 0;
    // 16-byte loads of the FP64 columns (the instantiated form).
    if (a.ldX % 2 != 0 || (uintptr_t)a.X % 16 != 0) return 0;
    // Reduced-precision coefficient chunks held per wave.
    const int nch = (a.k1 + 31) >> 5;
    if (ts == 32 ? nch - (nch >> 1) > kLpC : nch > kLpC) return 0;
    if (a.raw0 < a.ncols && a.raw0 % 16 != 0) return 0;
    // The FP32 loads move 16-byte groups of four rows (leading dimension a
    // multiple of 4, 16-byte aligned slab); FP32 and FP64 loads use 32-bit
    // per-lane byte offsets of up to 4 columns.
    if (a.ldX32 % 4 != 0 || (uintptr_t)a.X32 % 16 != 0 || a.ldX32 > (int64_t)1 << 27 ||
        a.ldX > (int64_t)1 << 27)
        return 0;
    return ts;
}

template <int MT, int NB, int F, bool EARLY, int MX>
hipError_t launch_bm_t(const PassArgs &a, size_t lds, int nwg, hipStream_t st)
{
    return launch_pass16b_v<MT, NB, F, EARLY, true, MX>(a, lds, nwg, st);
}

template <int MX>
hipError_t launch_bm_mx(const PassArgs &a, int ts, int mt, int nb, size_t lds, int nwg,
                        hipStream_t st)
{
    if (a.na == 0) {
        switch (ts) {
        case 32: return launch_bm_t<1, 1, 1, true, MX>(a, lds, nwg, st);
        case 64: return launch_bm_t<1, 1, 2, true, MX>(a, lds, nwg, st);
        case 128: return launch_bm_t<1, 1, 4, true, MX>(a, lds, nwg, st);
        case 256: return launch_bm_t<1, 1, 8, true, MX>(a, lds, nwg, st);
        default: return hipErrorInvalidValue;
        }
    }
#define ORTH_BM(M, F)                                                                  \
    if (mt == M)                                                                       \
        return nb == 2 ? launch_bm_t<M, 2, F, false, MX>(a, lds, nwg, st)              \
                       : launch_bm_t<M, 1, F, false, MX>(a, lds, nwg, st);
    switch (ts) {
    case 32:
        ORTH_BM(1, 1) ORTH_BM(2, 1) ORTH_BM(3, 1) ORTH_BM(4, 1)
        break;
    case 64:
        ORTH_BM(1, 2) ORTH_BM(2, 2)
        break;
    case 128:
        ORTH_BM(1, 4)
        break;
    case 256:
        ORTH_BM(1, 8)
        break;
    default: break;
    }
#undef ORTH_BM
    return hipErrorInvalidValue;
}

hipError_t launch_pass16bm(const PassArgs &a, int mx, int nwg, hipStream_t st)
{
    int mt = 0, nb = 0;
    const int ts = pass16bm_plan(a) ? pass16b_plan(a, &mt, &nb) : 0;
    if (ts == 0 || a.rows_per % ts != 0) return hipErrorInvalidValue;
    const size_t lds = pass16b_lds_bytes(a, ts);
    switch (mx) {
    case kMxBF16: return launch_bm_mx<kMxBF16>(a, ts, mt, nb, lds, nwg, st);
    case kMxFP16: return launch_bm_mx<kMxFP16>(a, ts, mt, nb, lds, nwg, st);
    case kMxBF16X3: return launch_bm_mx<kMxBF16X3>(a, ts, mt, nb, lds, nwg, st);
    case kMxFP64: return launch_bm_mx<kMxFP64>(a, ts, mt, nb, lds, nwg, st);
    default: return hipErrorInvalidValue;
    }
}

template <int NW, int NB, int MX>
hipError_t launch_pass16d_mixed_t(const PassArgs &a, int nwg, hipStream_t st)
{
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)k_pass16d<NW, NB, MX>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(160 * 1024));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const size_t lds = (size_t)pass16d_lds(a.wl, NW) * sizeof(double);
    k_pass16d<NW, NB, MX><<<dim3(nwg), dim3(64 * NW), pass_tail_lds(a, lds), st>>>(a);
    return hipGetLastError();
}

template <int MX>
hipError_t launch_pass16d_mx(const PassArgs &a, int nw, int nb, int nwg, hipStream_t st)
{
    if (nw == 16)
        return nb == 2 ? launch_pass16d_mixed_t<16, 2, MX>(a, nwg, st)
                       : launch_pass16d_mixed_t<16, 1, MX>(a, nwg, st);
    return nb == 2 ? launch_pass16d_mixed_t<8, 2, MX>(a, nwg, st)
                   : launch_pass16d_mixed_t<8, 1, MX>(a, nwg, st);
}

hipError_t launch_pass16d_mixed(const PassArgs &a, int mx, int nw, int nb, int nwg, hipStream_t st)
{
    if ((nw != 8 && nw != 16) || a.rows_per % 32 != 0 || !a.X32) return hipErrorInvalidValue;
    switch (mx) {
    case kMxFP64: return launch_pass16d_mx<kMxFP64>(a, nw, nb, nwg, st);
    case kMxBF16: return launch_pass16d_mx<kMxBF16>(a, nw, nb, nwg, st);
    case kMxFP16: return launch_pass16d_mx<kMxFP16>(a, nw, nb, nwg, st);
    case kMxBF16X3: return launch_pass16d_mx<kMxBF16X3>(a, nw, nb, nwg, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace orth
