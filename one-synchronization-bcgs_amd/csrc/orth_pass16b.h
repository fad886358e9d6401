// The 512-thread fused pass for block width 16 (FP64 form and, through the
// MX parameter, the mixed form), shared by the kernel translation units.
#pragma once

#include <type_traits>

#include "orth_device.h"
#include "orth_tail.h"

namespace orth {

// ---------------------------------------------------------------------------
// Fused pass, block width 16, 512 threads (two waves per SIMD).
//
// Rows are processed in super-tiles of TS = 32 f rows (f = 1, 2, 4, 8) so
// that a super-tile holds ~10K doubles whatever the column count: the
// prefetch of the next super-tile keeps ~80 KB of loads in flight per CU.
// Step 1 (f == 1): wave w computes Z tile (w & 3) over half of the inner
// dimension (w >> 2); the two halves are added in a fixed order through LDS.
// Step 1 (f >= 2): wave w computes Z tiles (rows 16 ((w >> 1) + 4 j), cols
// 16 (w & 1)).  Step 2: wave w owns projection N-tile (w & 1) and M-tiles
// (w >> 1) + 4 j when nb > 16, else M-tiles w + 8 j.
// ---------------------------------------------------------------------------

constexpr int kB = 512;
constexpr int kBTile = 10240;  // doubles of the LDS tile
constexpr int kBPF = 16;       // prefetch registers per thread (ncols * rows <= 8192)

// Reduced-precision operand chunks of 32 coefficient rows one wave keeps in
// registers (mixed form): a step-1 product over at most 10 chunks (k1 <= 320)
// split in halves, or 5 chunks in one piece.
constexpr int kLpC = 5;

// EARLY: passes without a projection issue the next tile's loads as soon as
// the current tile is in LDS (nothing else would hide their latency).
// V2: every thread loads two consecutive rows with one 16-byte load (needs
// an even leading dimension and a 16-byte aligned slab).
// MX (mixed form, kMx*; 0 = FP64 form): columns at or past raw0 come from
// the FP32 input X32; when MX is a reduced precision, the products the host
// marks (Z tiles consumed by lp groups, projection tiles whose B columns
// start at or past lpb) run on bf16 / fp16 MFMA with FP32 accumulation per
// 32-row chunk, summed in FP64.
template <int MT, int NB, int F, bool EARLY, bool V2, int MX = 0>
// Reduced-precision forms run one workgroup per CU (their operand
// conversions need more than the 128 registers of two).
__global__ __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(
    MX != 0 && MX != kMxFP64 ? 2 : 4, MX != 0 && MX != kMxFP64 ? 2 : 4))) void k_pass16b(PassArgs a)
{
    constexpr int N1 = 32, TS = 32 * F, LD = TS + 1;
    constexpr int LTS = F == 1 ? 5 : F == 2 ? 6 : F == 4 ? 7 : 8;
    extern __shared__ double sm[];
    const uint64_t kst = kernarg_stash();
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int k1p = (a.k1 + 3) & ~3;
    // LDS: the row tile, then the scratch of the step-1 half-sum.  The
    // coefficients C stay in global memory (one small block shared by every
    // workgroup, so it is served from L2).
    const int R0 = a.wl * LD;
    const double *__restrict__ Cg = a.C;
    for (int e = tid; e < (a.wl - a.ncols) * TS; e += kB)
        sm[(a.ncols + (e >> LTS)) * LD + (e & (TS - 1))] = 0.0;

    const int Ma = (a.na + 15) >> 4;
    const bool proj = a.na > 0;
    const int tb = NB == 2 ? (wave & 1) : 0;
    const int mb = NB == 2 ? (wave >> 1) : wave;
    constexpr int MS = NB == 2 ? 4 : 8;
    // Tiles beyond Ma read zero columns or other valid tile columns; their
    // products are never stored.
    int aoff2[MT];
#pragma unroll
    for (int j = 0; j < MT; j++) {
        const int ta = min(mb + MS * j, Ma - 1);
        aoff2[j] = (a.a0 + ta * 16 + li) * LD + lk;
    }
    const int boff2 = (a.b0 + tb * 16 + li) * LD + lk;
    const bool doproj = proj && tb * 16 < a.nb && mb < Ma;
    // Triangular solve operands of this wave's group (wave & 1), constant
    // during the pass: B fragments (lane: row lk, column li) of the inverted
    // diagonal blocks placed at columns 4b..4b+3, and of -T(4b+lk, li) for
    // the columns right of block b.
    const int gq = wave & 1;
    const double *Tq = gq < a.ngrp ? ORTH_G(a, gq, T) : nullptr;
    double bw[4];
#pragma unroll
    for (int b = 0; b < 4; b++) bw[b] = Tq ? Tq[kSlotW + 128 * b + lk * kLdT + li] : 0.0;
    dbl4 acc[MT];
#pragma unroll
    for (int j = 0; j < MT; j++) acc[j] = dbl4{0.0, 0.0, 0.0, 0.0};

    // Mixed form: is this wave's Z tile (N-tile znt) a reduced-precision
    // product, and its coefficient chunks [zq0, zq0 + zqn) as MFMA B
    // fragments (lane: k = 8 lk + e of column li), converted once.
    constexpr bool LP = MX != 0 && MX != kMxFP64;
    constexpr int MXL = LP ? MX : kMxBF16;  // operand format of the reduced-precision paths
    const int znt = F == 1 ? (wave >> 1) & 1 : wave & 1;
    const bool zlp = LP && !a.seq && ztile_lp(a, znt);
    int zq0 = 0, zqn = 0;
    LpFrag<MXL> fbz[LP ? kLpC : 1];
    if constexpr (LP) {
        const int nch = (a.k1 + 31) >> 5;
        zq0 = F == 1 && (wave >> 2) ? nch >> 1 : 0;
        zqn = F == 1 ? ((wave >> 2) ? nch - (nch >> 1) : nch >> 1) : nch;
#pragma unroll
        for (int c = 0; c < kLpC; c++) {
            const int k0 = 32 * (zq0 + c) + 8 * lk;
            fbz[c] = lp_frag_g<MXL>(Cg + (int64_t)k0 * N1 + znt * 16 + li, N1,
                                   zlp && c < zqn ? a.k1 - k0 : 0);
        }
    }
    const bool plp = LP && a.b0 + 16 * tb >= a.lpb;

    const int64_t rb = (int64_t)blockIdx.x * a.rows_per;
    const int64_t re = min(a.m, rb + a.rows_per);
    // Thread tid moves row (tid % TS) of columns tid / TS + i (512 / TS).
    constexpr int PSTEP = kB / TS;
    const int prow = tid & (TS - 1), pcol = tid >> LTS;
    // Loads: thread tid moves rows (2 (tid % (TS/2)), +1) of columns
    // tid / (TS/2) + i (1024 / TS) when V2, else row tid % TS of columns
    // tid / TS + i (512 / TS).
    typedef double dbl2v __attribute__((ext_vector_type(2)));
    constexpr int VR = V2 ? 2 : 1;               // rows per thread
    constexpr int LSTEP = VR * kB / TS;          // columns per load step
    constexpr int NPF = kBPF / VR;               // load steps
    const int lrow = VR * (tid & (TS / VR - 1)), lcol = tid / (TS / VR);
    const int lnld = (a.ncols - lcol + LSTEP - 1) / LSTEP;
    const double *lx = a.X + lrow + (int64_t)lcol * a.ldX;
    const int64_t lxs = (int64_t)LSTEP * a.ldX;
    const int lpl = lcol * LD + lrow;
    typename std::conditional<V2, dbl2v, double>::type pf[NPF];
    // Raw (FP32) columns: a.raw0 is a multiple of 16 and a wave's lanes move
    // at most four consecutive columns of a 4-aligned group, so whether load
    // step i is raw is uniform per wave.  A raw step moves RF = 2 VR FP32
    // rows per lane (the prefetch register's width) with the first half of
    // each column's lanes; the second half repeats the first half's rows and
    // is not stored.  Its bits stay in the prefetch register and are
    // converted (exactly) when the tile is written to LDS, so the load is
    // one instruction of the FP64 step's width at a per-wave selected
    // address and nothing waits for its data at issue.  The FP32 columns are
    // addressed from a wave-uniform base (the wave's first load column) plus
    // one 32-bit byte offset per lane, below 16 ldX32 bytes (the host checks
    // the size and 16-byte alignment).
    typedef float flt2v __attribute__((ext_vector_type(2)));
    typedef float flt4v __attribute__((ext_vector_type(4)));
    constexpr int RF = 2 * VR;
    const int lq = lrow / VR, rq = lq & (TS / RF - 1);
    const bool rstore = lq < TS / RF;
    auto raw_step = [&](int i) {
        return MX != 0 && __builtin_amdgcn_readfirstlane((int)(lcol + i * LSTEP >= a.raw0)) != 0;
    };
    // The mixed form addresses its FP64 columns the same way, so a load step
    // keeps no 64-bit lane address; row checks compare against the
    // wave-uniform count of rows left in the super-tile.
    const char *wx32 = nullptr, *wx64 = nullptr;
    int64_t lxs32 = 0, lxs64 = 0;
    uint32_t lb32 = 0, lb64 = 0;
    if constexpr (MX != 0) {
        const int wcol = __builtin_amdgcn_readfirstlane(lcol);
        wx32 = (const char *)(a.X32 + (int64_t)wcol * a.ldX32);
        wx64 = (const char *)(a.X + (int64_t)wcol * a.ldX);
        lxs32 = (int64_t)LSTEP * a.ldX32 * 4;
        lxs64 = (int64_t)LSTEP * a.ldX * 8;
        lb32 = 4u * (uint32_t)(RF * rq + (lcol - wcol) * a.ldX32);
        lb64 = 8u * (uint32_t)(lrow + (lcol - wcol) * a.ldX);
    }
    auto load = [&](int64_t r0) {
        if constexpr (MX != 0) {
            const int rem = (int)min(re - r0, (int64_t)TS);
            // The lane offsets are opaque to the optimizer here, so each
            // address stays (uniform base) + (32-bit lane offset) instead of
            // eight hoisted 64-bit lane addresses per source.
            uint32_t o32 = lb32, o64 = lb64;
            asm volatile("" : "+v"(o32), "+v"(o64));
            const char *b32 = wx32 + 4 * r0, *b64 = wx64 + 8 * r0;
            const bool ok1 = lrow < rem, ok2 = lrow + VR - 1 < rem, ok4 = RF * rq + RF - 1 < rem;
#pragma unroll
            for (int i = 0; i < NPF; i++)
                if (i < lnld) {
                    using PV = typename std::remove_reference<decltype(pf[0])>::type;
                    if (raw_step(i)) {
                        const float *q32 = (const float *)(b32 + i * lxs32 + o32);
                        if (ok4) {
                            pf[i] = __builtin_nontemporal_load((const PV *)q32);
                        } else {
                            float v[RF];
#pragma unroll
                            for (int e = 0; e < RF; e++) v[e] = RF * rq + e < rem ? q32[e] : 0.f;
                            if constexpr (V2)
                                pf[i] = __builtin_bit_cast(dbl2v, flt4v{v[0], v[1], v[2], v[3]});
                            else
                                pf[i] = __builtin_bit_cast(double, flt2v{v[0], v[1]});
                        }
                    } else {
                        const double *q64 = (const double *)(b64 + i * lxs64 + o64);
                        if (ok2)
                            pf[i] = __builtin_nontemporal_load((const PV *)q64);
                        else if constexpr (V2)
                            pf[i] = dbl2v{ok1 ? q64[0] : 0.0, 0.0};
                        else
                            pf[i] = 0.0;
                    }
                }
        } else if constexpr (V2) {
            const double *q = lx + r0;
            const bool ok2 = r0 + lrow + 1 < re, ok1 = r0 + lrow < re;
#pragma unroll
            for (int i = 0; i < NPF; i++)
                if (i < lnld) {
                    if (ok2)
                        pf[i] = __builtin_nontemporal_load((const dbl2v *)(q + i * lxs));
                    else
                        pf[i] = dbl2v{ok1 ? q[i * lxs] : 0.0, 0.0};
                }
        } else {
            const double *q = lx + r0;
            const bool ok = r0 + lrow < re;
#pragma unroll
            for (int i = 0; i < NPF; i++)
                if (i < lnld) pf[i] = ok ? __builtin_nontemporal_load(q + i * lxs) : 0.0;
        }
    };
    if (rb < re) load(rb);

    for (int64_t r0 = rb; r0 < re; r0 += TS) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NPF; i++)
            if (i < lnld) {
                if (raw_step(i)) {
                    if (rstore) {
                        double *d = sm + lcol * LD + RF * rq + i * LSTEP * LD;
                        if constexpr (V2) {
                            const flt4v v = __builtin_bit_cast(flt4v, pf[i]);
#pragma unroll
                            for (int e = 0; e < 4; e++) d[e] = (double)v[e];
                        } else {
                            const flt2v v = __builtin_bit_cast(flt2v, pf[i]);
                            d[0] = (double)v.x;
                            d[1] = (double)v.y;
                        }
                    }
                } else if constexpr (V2) {
                    sm[lpl + i * LSTEP * LD] = pf[i].x;
                    sm[lpl + 1 + i * LSTEP * LD] = pf[i].y;
                } else {
                    sm[lpl + i * LSTEP * LD] = pf[i];
                }
            }
        __syncthreads();
        if constexpr (EARLY) {
            if (r0 + TS < re) load(r0 + TS);
        }

        if (a.ngrp > 0) {
            const int nphase = a.seq ? a.ngrp : 1;
            for (int p = 0; p < nphase; p++) {
                const int glo = a.seq ? p : 0, ghi = a.seq ? p + 1 : a.ngrp;
                const int kp = a.seq ? ((ORTH_G(a, p, k) + 3) & ~3) : k1p;
                const int nks = a.k1 > 0 ? kp >> 2 : 0;
                if constexpr (F == 1) {
                    // Z tile (wave & 3) over half (wave >> 2) of the k-steps;
                    // the halves are added in a fixed order through LDS.
                    const int item = wave & 3, mt = item & 1, nt = item >> 1, kh = wave >> 2;
                    const bool mine = nks > 0 && (a.seq ? nt == (ORTH_G(a, p, zoff) >> 4)
                                                        : ztile_used(a, nt));
                    dbl4 z = dbl4{0.0, 0.0, 0.0, 0.0};
                    if (mine) {
                        if (zlp) {
                            // 32 coefficient rows per reduced-precision MFMA;
                            // rows past k1 have zero coefficients and read
                            // clamped tile columns.
                            f32x4 t = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                            for (int c = 0; c < kLpC; c++)
                                if (c < zqn) {
                                    const int k0 = 32 * (zq0 + c) + 8 * lk;
                                    const LpFrag<MXL> fa = lp_frag<MXL>(
                                        sm + min(k0, a.wl - 8) * LD + mt * 16 + li, LD);
                                    t = lp_mma<MXL>(fa, fbz[c], t);
                                    __builtin_amdgcn_sched_barrier(0);
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
                                // Result rows: lk + 4 v (FP64 MFMA) or 4 lk + v
                                // (reduced-precision MFMA).
                                const int o = (ORTH_G(a, g, t0) + off) * LD + mt * 16;
#pragma unroll
                                for (int v = 0; v < 4; v++)
                                    sm[o + (zlp ? 4 * lk + v : lk + 4 * v)] -= z[v];
                            }
                        }
                    }
                } else {
                    // Z tiles (rows 16 ((wave >> 1) + 4 j), cols 16 (wave & 1)).
                    constexpr int NM = F > 1 ? F / 2 : 1;
                    const int nt = wave & 1, mt0 = wave >> 1;
                    const bool mine = nks > 0 && (a.seq ? nt == (ORTH_G(a, p, zoff) >> 4)
                                                        : ztile_used(a, nt));
                    dbl4 z[NM];
#pragma unroll
                    for (int j = 0; j < NM; j++) z[j] = dbl4{0.0, 0.0, 0.0, 0.0};
                    if (mine && zlp) {
#pragma unroll
                        for (int j = 0; j < NM; j++) {
                            f32x4 t = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                            for (int c = 0; c < kLpC; c++)
                                if (c < zqn) {
                                    const int k0 = 32 * c + 8 * lk;
                                    const LpFrag<MXL> fa = lp_frag<MXL>(
                                        sm + min(k0, a.wl - 8) * LD + (mt0 + 4 * j) * 16 + li, LD);
                                    t = lp_mma<MXL>(fa, fbz[c], t);
                                    __builtin_amdgcn_sched_barrier(0);
                                }
                            z[j] = dbl4{t[0], t[1], t[2], t[3]};
                        }
                    } else if (mine) {
                        const int bo = lk * N1 + nt * 16 + li;
                        const int ao = lk * LD + mt0 * 16 + li;
                        double bn = Cg[bo], an[NM];
#pragma unroll
                        for (int j = 0; j < NM; j++) an[j] = sm[ao + 64 * j];
                        for (int ks = 0; ks < nks; ks++) {
                            const double bc = bn;
                            double ac[NM];
#pragma unroll
                            for (int j = 0; j < NM; j++) ac[j] = an[j];
                            const int nx = min(ks + 1, nks - 1);
                            bn = Cg[bo + 4 * nx * N1];
#pragma unroll
                            for (int j = 0; j < NM; j++) an[j] = sm[ao + 4 * nx * LD + 64 * j];
#pragma unroll
                            for (int j = 0; j < NM; j++)
                                z[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[j], bc, z[j], 0, 0,
                                                                           0);
                        }
                    }
                    __syncthreads();
                    if (mine) {
                        const int zc = nt * 16 + li;
                        for (int g = glo; g < ghi; g++) {
                            const int off = zc - ORTH_G(a, g, zoff);
                            if (off >= 0 && off < ORTH_G(a, g, w)) {
#pragma unroll
                                for (int j = 0; j < NM; j++) {
                                    const int o = (ORTH_G(a, g, t0) + off) * LD + (mt0 + 4 * j) * 16;
#pragma unroll
                                    for (int v = 0; v < 4; v++)
                                        sm[o + (zlp ? 4 * lk + v : lk + 4 * v)] -= z[j][v];
                                }
                            }
                        }
                    }
                }
                __syncthreads();
                // Triangular solve X <- X T^{-1} of each group, blocked by
                // four columns on FP64 MFMA (the inverted-diagonal-block
                // scheme of blocked dtrsm): block b's columns become
                // b_b D_b and the columns right of it lose b_b D_b T(b, >),
                // one product with W_b per block.  Wave w takes group (w & 1)
                // and 16-row chunks (w >> 1) + 4 i; the chunk lives in MFMA
                // accumulator layout (lane: column li, rows lk + 4 v) and each
                // block goes through LDS once to become an A operand.
                if (Tq && gq >= glo && gq < ghi) {
                    const int t0 = ORTH_G(a, gq, t0), w = ORTH_G(a, gq, w);
                    for (int ch = wave >> 1; ch < TS / 16; ch += 4) {
                        const int base = (t0 + li) * LD + ch * 16 + lk;
                        const int abase = t0 * LD + ch * 16 + li + lk * LD;
                        dbl4 x;
#pragma unroll
                        for (int v = 0; v < 4; v++) x[v] = sm[base + 4 * v];
#pragma unroll
                        for (int b = 0; b < 4; b++) {
                            // Columns at or past the group width keep their
                            // tile contents (they only feed each other).
                            const bool mineb = (li >> 2) == b && li < w;
                            if (mineb && b > 0) {
#pragma unroll
                                for (int v = 0; v < 4; v++) sm[base + 4 * v] = x[v];
                            }
                            // Other lanes' LDS writes become visible to this
                            // lane (the wave's LDS accesses execute in order);
                            // the fence stops the compiler from forwarding the
                            // lane's own earlier value instead of reloading.
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
        // Next tile's loads: issued after the triangular solve, so the
        // prefetch registers are live only across the projection.
        if constexpr (!EARLY) {
            if (r0 + TS < re) load(r0 + TS);
        }

        if (doproj && plp) {
            // One reduced-precision MFMA per 32-row chunk and tile, FP32
            // accumulation within the chunk, FP64 across chunks.
#pragma unroll 1
            for (int c = 0; c < F; c++) {
                const LpFrag<MXL> fb =
                    lp_frag<MXL>(sm + boff2 - lk + 32 * c + 8 * lk, 1);
#pragma unroll
                for (int j = 0; j < MT; j++) {
                    const LpFrag<MXL> fa =
                        lp_frag<MXL>(sm + aoff2[j] - lk + 32 * c + 8 * lk, 1);
                    const f32x4 t = lp_mma<MXL>(fa, fb, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[j][v] += (double)t[v];
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        } else if (doproj) {
#pragma unroll
            for (int ks = 0; ks < TS / 4; ks++) {
                const double bc = sm[boff2 + 4 * ks];
                double ac[MT];
#pragma unroll
                for (int j = 0; j < MT; j++) ac[j] = sm[aoff2[j] + 4 * ks];
#pragma unroll
                for (int j = 0; j < MT; j++)
                    acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[j], bc, acc[j], 0, 0, 0);
                // Bound the operand loads in flight (register pressure): two
                // k-steps, one in the mixed form's four-tile instances (their
                // raw-column loader holds a few more registers).
                if ((MX != 0 && MT >= 4) || (ks & 1) == 1) __builtin_amdgcn_sched_barrier(0);
            }
        }
    }

    if (doproj) {
        double *out = a.part + (int64_t)blockIdx.x * a.na * a.nb;
#pragma unroll
        for (int j = 0; j < MT; j++) {
            const int ta = mb + MS * j;
            if (ta < Ma) {
                const int col = tb * 16 + li;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = ta * 16 + (plp ? 4 * lk + v : lk + 4 * v);
                    if (row < a.na && col < a.nb) out[row + (int64_t)col * a.na] = acc[j][v];
                }
            }
        }
    }
    pass_tail(a, sm, kst, kB);
}

template <int MT, int NB, int F, bool EARLY, bool V2, int MX = 0>
hipError_t launch_pass16b_v(const PassArgs &a, size_t lds, int nwg, hipStream_t st)
{
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)k_pass16b<MT, NB, F, EARLY, V2, MX>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(160 * 1024));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    k_pass16b<MT, NB, F, EARLY, V2, MX><<<dim3(nwg), dim3(kB), pass_tail_lds(a, lds), st>>>(a);
    return hipGetLastError();
}


}  // namespace orth
