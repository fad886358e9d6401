// Device kernels of liborth_hip.so (gfx950).
//
//  k_update  row-local block updates, one row per lane, coefficients read as
//            wave-uniform (scalar) loads
//  k_upd     the schedules' paired block updates on v_mfma_f64_16x16x4f64
//            (shared prefix reads, blocked MFMA triangular solve)
//  k_proj    A^T B over a row slab on v_mfma_f64_16x16x4f64, 16-row tiles
//            staged through LDS with the next tile prefetched into
//            registers; one partial per workgroup
//  k_pass    the fused pass for any block width (256 threads; s = 16 mostly
//            runs the 512-thread k_pass16b of orth_pass16b.h)
//  k_reduce  fixed-order sum of the partials (the in-kernel replacement of
//            the reference's MPI_Allreduce)
//  k_small   the s x s / c x s steps between passes, one workgroup, FP64
//  k_resid   ||Q R - X||_F and ||X||_F partial sums (accuracy harness)
#include "orth_kernels.h"

#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <type_traits>

#include "orth_device.h"
#include "orth_pass16b.h"
#include "orth_pass16d.h"
#include "orth_small.h"

namespace orth {

// ---------------------------------------------------------------------------
// Row-local updates
// ---------------------------------------------------------------------------

template <int S>
__device__ __forceinline__ void apply_upd(double *__restrict__ xr, int64_t ldX,
                                          const Upd &u)
{
    double acc[S];
#pragma unroll
    for (int i = 0; i < S; i++) acc[i] = (i < u.w) ? xr[(u.col0 + i) * ldX] : 0.0;
    const int c = u.c;
    int j = 0;
    for (; j + 4 <= c; j += 4) {
        const double q0 = xr[(int64_t)j * ldX];
        const double q1 = xr[(int64_t)(j + 1) * ldX];
        const double q2 = xr[(int64_t)(j + 2) * ldX];
        const double q3 = xr[(int64_t)(j + 3) * ldX];
        const double *w = u.W + j;
#pragma unroll
        for (int i = 0; i < S; i++) {
            const double *wi = w + i * u.ldW;
            double a = acc[i];
            a = fma(-q0, wi[0], a);
            a = fma(-q1, wi[1], a);
            a = fma(-q2, wi[2], a);
            a = fma(-q3, wi[3], a);
            acc[i] = a;
        }
    }
    for (; j < c; j++) {
        const double q = xr[(int64_t)j * ldX];
#pragma unroll
        for (int i = 0; i < S; i++) acc[i] = fma(-q, u.W[j + i * u.ldW], acc[i]);
    }
    if (u.T) {
        // Row-wise forward substitution with the upper factor (dtrsm R,U,N,N).
#pragma unroll
        for (int i = 0; i < S; i++) {
            if (i < u.w) {
                double v = acc[i];
#pragma unroll
                for (int l = 0; l < i; l++) v = fma(-acc[l], u.T[l + i * kLdT], v);
                acc[i] = v * (1.0 / u.T[i + i * kLdT]);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < S; i++)
        if (i < u.w) xr[(u.col0 + i) * ldX] = acc[i];
}

template <int S>
__global__ __launch_bounds__(256) void k_update(double *X, int64_t ldX, int64_t m,
                                                Upd u0, Upd u1, int nupd)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    double *xr = X + r;
    apply_upd<S>(xr, ldX, u0);
    if (nupd > 1) apply_upd<S>(xr, ldX, u1);
}

// Block updates on FP64 MFMA (s = 16, 32: NT = s / 16 column tiles).  Each
// wave owns 16 MT rows.  The products over the shared prefix [0, u0.c) of
// both updates run in one K-loop (one read of the prefix); the triangular
// solve X <- X T^{-1} is the blocked scheme of the s = 16 pass (one MFMA per
// 4-column block with the row blocks [D_b | -D_b T(b, >)] the Cholesky step
// stores beside T).  u1's prefix is u0's prefix followed by u0's block
// (checked by the host): that block's final values go from registers
// through LDS into u1's product.  No workgroup barrier: waves past the last
// row return.
// TWO: u1 follows u0 (the pairing); else u0 alone.
template <int NT, int MT, bool TWO>
__global__ __launch_bounds__(256) void k_upd(double *X, int64_t ldX, int64_t m, Upd u0, Upd u1)
{
    constexpr int S = 16 * NT, LDY = S + 1, RW = 16 * MT;
    __shared__ double ys_all[4][RW * LDY];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * RW;
    if (row0 >= m) return;
    double *ys = ys_all[wave];
    int64_t ra[MT];  // A-operand row of this lane per M-tile (clamped to the slab)
#pragma unroll
    for (int t = 0; t < MT; t++) ra[t] = min(row0 + 16 * t + li, m - 1);
    dbl4 z0[MT][NT], z1[TWO ? MT : 1][NT];
#pragma unroll
    for (int t = 0; t < MT; t++)
#pragma unroll
        for (int n = 0; n < NT; n++) {
            z0[t][n] = dbl4{0.0, 0.0, 0.0, 0.0};
            if constexpr (TWO) z1[t][n] = dbl4{0.0, 0.0, 0.0, 0.0};
        }
    // Z += X(:, j0:j0+4) W(j0:j0+4, :) for j0 in [lo, hi), for u0 and (when
    // paired) u1: A lane (li, lk) holds X(row li, j0 + lk), B lane holds
    // W(j0 + lk, 16 n + li); columns past hi read column hi - 1 with zero
    // coefficients.
    auto kloop = [&](int lo, int hi) {
#pragma unroll 4
        for (int j0 = lo; j0 < hi; j0 += 4) {
            const int j = j0 + lk;
            const bool jok = j < hi;
            const int64_t jc = jok ? j : hi - 1;
            double a[MT];
#pragma unroll
            for (int t = 0; t < MT; t++) a[t] = X[ra[t] + jc * ldX];
#pragma unroll
            for (int n = 0; n < NT; n++) {
                const int col = 16 * n + li;
                {
                    const double b = jok && col < u0.w ? u0.W[j + (int64_t)col * u0.ldW] : 0.0;
#pragma unroll
                    for (int t = 0; t < MT; t++)
                        z0[t][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b, z0[t][n], 0, 0, 0);
                }
                if constexpr (TWO) {
                    const double b = jok && col < u1.w ? u1.W[j + (int64_t)col * u1.ldW] : 0.0;
#pragma unroll
                    for (int t = 0; t < MT; t++)
                        z1[t][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b, z1[t][n], 0, 0, 0);
                }
            }
        }
    };
    // Y = X(:, col0:col0+w) - Z, then Y <- Y T^{-1}, then store (D layout:
    // lane (li, lk) value v is row lk + 4 v, column 16 n + li).
    auto finish = [&](const Upd &u, auto &z) {
#pragma unroll
        for (int t = 0; t < MT; t++)
#pragma unroll
            for (int n = 0; n < NT; n++) {
                const int col = 16 * n + li;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int64_t row = row0 + 16 * t + lk + 4 * v;
                    const double x = row < m && col < u.w ? X[row + (u.col0 + col) * ldX] : 0.0;
                    z[t][n][v] = x - z[t][n][v];
                }
            }
        if (u.T) {
            const double *T = u.T;
            for (int b = 0; b < (u.w + 3) >> 2; b++) {
                const int nt = b >> 2, cb = b & 3;
                // Block b's current values go through LDS to become A operands.
                if ((li >> 2) == cb) {
#pragma unroll
                    for (int t = 0; t < MT; t++)
#pragma unroll
                        for (int n = 0; n < NT; n++)
                            if (n == nt) {
#pragma unroll
                                for (int v = 0; v < 4; v++)
                                    ys[(16 * t + lk + 4 * v) * LDY + 4 * b + (li & 3)] = z[t][n][v];
                            }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                double av[MT];
#pragma unroll
                for (int t = 0; t < MT; t++) av[t] = ys[(16 * t + li) * LDY + 4 * b + lk];
#pragma unroll
                for (int n = 0; n < NT; n++) {
                    if (n < nt) continue;
                    const double bv = T[kSlotW + 128 * b + lk * kLdT + 16 * n + li];
#pragma unroll
                    for (int t = 0; t < MT; t++) {
                        if (n == nt && (li >> 2) == cb) z[t][n] = dbl4{0.0, 0.0, 0.0, 0.0};
                        z[t][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t], bv, z[t][n], 0, 0, 0);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            }
        }
#pragma unroll
        for (int t = 0; t < MT; t++)
#pragma unroll
            for (int n = 0; n < NT; n++) {
                const int col = 16 * n + li;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int64_t row = row0 + 16 * t + lk + 4 * v;
                    if (row < m && col < u.w) X[row + (u.col0 + col) * ldX] = z[t][n][v];
                }
            }
    };
    if (u0.c > 0) kloop(0, u0.c);
    finish(u0, z0);
    if constexpr (TWO) {
        // u1 += (u0's finished block) W1(u0.c : u0.c + u0.w, :).
#pragma unroll
        for (int t = 0; t < MT; t++)
#pragma unroll
            for (int n = 0; n < NT; n++)
                if (16 * n + li < u0.w) {
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        ys[(16 * t + lk + 4 * v) * LDY + 16 * n + li] = z0[t][n][v];
                }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        for (int j0 = 0; j0 < u0.w; j0 += 4) {
            const int j = j0 + lk;
            const bool jok = j < u0.w;
            double av[MT];
#pragma unroll
            for (int t = 0; t < MT; t++) av[t] = jok ? ys[(16 * t + li) * LDY + j] : 0.0;
#pragma unroll
            for (int n = 0; n < NT; n++) {
                const int col = 16 * n + li;
                const double b = jok && col < u1.w ? u1.W[u0.c + j + (int64_t)col * u1.ldW] : 0.0;
#pragma unroll
                for (int t = 0; t < MT; t++)
                    z1[t][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t], b, z1[t][n], 0, 0, 0);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        finish(u1, z1);
    }
}

// Paired updates, 32 rows per wave.
template <int NT>
hipError_t launch_upd_pair(double *X, int64_t ldX, int64_t m, const Upd *u, hipStream_t st)
{
    const int64_t rows = 4 * 16 * 2;
    k_upd<NT, 2, true><<<dim3((unsigned)((m + rows - 1) / rows)), dim3(256), 0, st>>>(X, ldX, m,
                                                                                    u[0], u[1]);
    return hipGetLastError();
}

hipError_t launch_update(int S, double *X, int64_t ldX, int64_t m,
                         const Upd *u, int nupd, hipStream_t st)
{
    if (m <= 0 || nupd <= 0) return hipSuccess;
    // The MFMA form takes the schedules' pairing (u1's prefix is u0's prefix
    // followed by u0's block), where it shares the prefix and coefficient
    // loads between the two products; single updates stay row-wise (there,
    // even at 64 rows per wave, the MFMA form's coefficient loads cost more
    // than it saves).
    const bool paired = nupd == 2 && u[0].col0 == u[0].c && u[1].c == u[0].c + u[0].w;
    if ((S == 16 || S == 32) && paired && !getenv("ORTH_UPDATE_ROWWISE"))
        return S == 16 ? launch_upd_pair<1>(X, ldX, m, u, st) : launch_upd_pair<2>(X, ldX, m, u, st);
    dim3 grid((unsigned)((m + 255) / 256)), block(256);
    Upd u1 = nupd > 1 ? u[1] : u[0];
    switch (S) {
    case 4: k_update<4><<<grid, block, 0, st>>>(X, ldX, m, u[0], u1, nupd); break;
    case 8: k_update<8><<<grid, block, 0, st>>>(X, ldX, m, u[0], u1, nupd); break;
    case 16: k_update<16><<<grid, block, 0, st>>>(X, ldX, m, u[0], u1, nupd); break;
    case 32: k_update<32><<<grid, block, 0, st>>>(X, ldX, m, u[0], u1, nupd); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Projection A^T B on FP64 MFMA
// ---------------------------------------------------------------------------

constexpr int kRT = 16;       // rows per LDS tile
constexpr int kLdsLd = kRT + 1;  // padded column stride in LDS (doubles)
constexpr int kMaxTilesPerWave = 12;
constexpr int kMaxLdsCols = 20 * 16;

ProjPlan plan_proj(int64_t m, int64_t na, int64_t nb, int target_wg)
{
    ProjPlan p;
    p.Ma = (int)((na + 15) / 16);
    p.Nb = (int)((nb + 15) / 16);
    p.gb = p.Nb <= 4 ? p.Nb : (p.Nb < 8 ? p.Nb : 8);
    int ga = 4 * kMaxTilesPerWave / p.gb;
    if (ga > kMaxLdsCols / 16 - p.gb) ga = kMaxLdsCols / 16 - p.gb;
    if (ga > p.Ma) ga = p.Ma;
    if (ga < 1) ga = 1;
    p.ga = ga;
    p.ngy = ((p.Ma + p.ga - 1) / p.ga) * ((p.Nb + p.gb - 1) / p.gb);
    int64_t chunks = (m + kRT - 1) / kRT;
    int64_t want = target_wg / p.ngy;
    if (want < 1) want = 1;
    if (want > chunks) want = chunks;
    if (want < 1) want = 1;
    int64_t rows = (m + want - 1) / want;
    rows = ((rows + kRT - 1) / kRT) * kRT;
    if (rows < kRT) rows = kRT;
    p.rows_per = rows;
    p.nwg = (int)((m + rows - 1) / rows);
    if (p.nwg < 1) p.nwg = 1;
    p.lds_bytes = (size_t)(p.ga + p.gb) * 16 * kLdsLd * sizeof(double);
    return p;
}

// PF: load steps per tile (16 columns each), at least the column group's
// width / 16; every step loads (columns past the group read column a0).
template <int PF>
__global__ __launch_bounds__(256) void k_proj(const double *__restrict__ X, int64_t ldX,
                                              int64_t m, int64_t a0, int na, int64_t b0,
                                              int nb, int Ma, int Nb, int ga, int gb,
                                              int64_t rows_per, double *__restrict__ part)
{
    extern __shared__ double lds[];
    const int ngb = (Nb + gb - 1) / gb;
    const int gy = blockIdx.y;
    const int ta0 = (gy / ngb) * ga, tb0 = (gy % ngb) * gb;
    const int nta = min(ga, Ma - ta0), ntb = min(gb, Nb - tb0);
    const int colsA = nta * 16, ncol = colsA + ntb * 16;
    const int ntiles = nta * ntb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;

    dbl4 acc[kMaxTilesPerWave];
#pragma unroll
    for (int t = 0; t < kMaxTilesPerWave; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};

    const int64_t rb = (int64_t)blockIdx.x * rows_per;
    const int64_t re = min(m, rb + rows_per);
    // The next tile is prefetched into registers while the current one is
    // multiplied.  Element tid + 256 i of a tile is column (tid >> 4) + 16 i,
    // row tid & 15, so load step i of wave w moves columns 4 w + 16 i .. + 3,
    // which never straddle the A / B boundary (colsA is a multiple of 16):
    // the step's first column is wave-uniform and each lane adds its column
    // offset (lane >> 4).  Every lane loads unconditionally from a valid
    // address (columns past na / nb read column a0, rows past the end the
    // last row); the LDS store zeroes those elements.
    const int rr = threadIdx.x & (kRT - 1), lc = lane >> 4;
    const int64_t lcoff = (int64_t)lc * ldX;
    double pf[PF];
    // Group column cl (wave-uniform cw = cl - lc) -> matrix column and validity.
    auto colbase = [&](int cw, int64_t &base, bool &lok) {
        if (cw < colsA) {
            const int ac = ta0 * 16 + cw;
            lok = ac + lc < na;
            base = ac < na ? a0 + ac : a0;
        } else {
            const int bc = tb0 * 16 + (cw - colsA);
            lok = bc + lc < nb;
            base = bc < nb ? b0 + bc : a0;
        }
    };
    auto load = [&](int64_t r0) {
        int64_t rowc = min(r0 + rr, re - 1);
        // Opaque to the optimizer, so the addresses are not hoisted.
        asm volatile("" : "+v"(rowc));
#pragma unroll
        for (int i = 0; i < PF; i++) {
            const int cw = 4 * wave + 16 * i;
            int64_t base = a0;
            bool lok = false;
            if (cw < ncol) colbase(cw, base, lok);
            pf[i] = X[base * ldX + (lok ? lcoff : 0) + rowc];
        }
    };
    if (rb < re) load(rb);
    for (int64_t r0 = rb; r0 < re; r0 += kRT) {
        __syncthreads();
        const bool rok = r0 + rr < re;
#pragma unroll
        for (int i = 0; i < PF; i++) {
            const int cw = 4 * wave + 16 * i;
            if (cw < ncol) {
                int64_t base;
                bool lok;
                colbase(cw, base, lok);
                lds[(cw + lc) * kLdsLd + rr] = lok && rok ? pf[i] : 0.0;
            }
        }
        __syncthreads();
        if (r0 + kRT < re) load(r0 + kRT);
#pragma unroll
        for (int ks = 0; ks < kRT / 4; ks++) {
            const int rk = ks * 4 + lk;
#pragma unroll
            for (int t = 0; t < kMaxTilesPerWave; t++) {
                const int tile = wave + 4 * t;
                if (tile < ntiles) {
                    const int ta = tile / ntb, tb = tile - ta * ntb;
                    const double a = lds[(ta * 16 + li) * kLdsLd + rk];
                    const double b = lds[(colsA + tb * 16 + li) * kLdsLd + rk];
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
                }
            }
        }
    }
    // D layout of v_mfma_f64_16x16x4f64: lane (li, lk), value v holds row
    // lk + 4 v, column li of the tile.
    double *out = part + (int64_t)blockIdx.x * na * nb;
#pragma unroll
    for (int t = 0; t < kMaxTilesPerWave; t++) {
        const int tile = wave + 4 * t;
        if (tile < ntiles) {
            const int ta = tile / ntb, tb = tile - ta * ntb;
            const int j = (tb0 + tb) * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = (ta0 + ta) * 16 + lk + 4 * v;
                if (i < na && j < nb) out[i + (int64_t)j * na] = acc[t][v];
            }
        }
    }
}

hipError_t launch_proj(const ProjPlan &pl, const double *X, int64_t ldX, int64_t m,
                       int64_t a0, int64_t na, int64_t b0, int64_t nb, double *part,
                       hipStream_t st)
{
    dim3 grid(pl.nwg, pl.ngy), block(256);
    const int steps = (pl.ga + pl.gb + 3) / 4 * 4;  // 16-column load steps, rounded up to 4
#define ORTH_PROJ(P)                                                                        \
    case P:                                                                                 \
        k_proj<P><<<grid, block, pl.lds_bytes, st>>>(X, ldX, m, a0, (int)na, b0, (int)nb,   \
                                                     pl.Ma, pl.Nb, pl.ga, pl.gb,            \
                                                     pl.rows_per, part);                    \
        break;
    switch (steps) {
        ORTH_PROJ(4) ORTH_PROJ(8) ORTH_PROJ(12) ORTH_PROJ(16) ORTH_PROJ(20)
    default: return hipErrorInvalidValue;
    }
#undef ORTH_PROJ
    return hipGetLastError();
}

// Sums of the partials of 32 elements in the fixed order of the in-pass tail
// (tail_group): thread (el, q) sums groups q, q + 8, ... of element
// blockIdx.x * 32 + el into red, then q == 0 adds the group sums in order.
static __device__ double reduce_fixed_order(const double *__restrict__ part, int nwg, int grp,
                                            int64_t E, double (*red)[32])
{
    const int el = threadIdx.x & 31, q = threadIdx.x >> 5;
    const int64_t e = (int64_t)blockIdx.x * 32 + el;
    const int ngrp = (nwg + grp - 1) / grp;
    if (e < E)
        for (int g = q; g < ngrp; g += 8)
            red[g][el] = part_range_sum(part, E, e, g * grp, min(nwg, g * grp + grp));
    __syncthreads();
    double t = 0.0;
    if (q == 0 && e < E) {
        t = red[0][el];
        for (int g = 1; g < ngrp; g++) t += red[g][el];
    }
    return t;
}

__global__ __launch_bounds__(256) void k_reduce(const double *__restrict__ part, int nwg, int grp,
                                                int64_t na, int64_t E,
                                                double *__restrict__ out, int64_t ldo)
{
    __shared__ double red[kTailMaxGroups][32];
    const double t = reduce_fixed_order(part, nwg, grp, E, red);
    const int64_t e = (int64_t)blockIdx.x * 32 + (threadIdx.x & 31);
    if ((threadIdx.x >> 5) == 0 && e < E) {
        const int64_t i = e % na, j = e / na;
        out[i + j * ldo] = t;
    }
}

hipError_t launch_reduce(const double *part, int nwg, int64_t na, int64_t nb, double *out,
                         int64_t ldo, hipStream_t st)
{
    const int64_t E = na * nb;
    if (E <= 0) return hipSuccess;
    k_reduce<<<dim3((unsigned)((E + 31) / 32)), dim3(256), 0, st>>>(part, nwg, tail_group(nwg), na,
                                                                   E, out, ldo);
    return hipGetLastError();
}


// k_reduce's fixed-order sum, then the small step in the last workgroup to
// finish (ticket counter; release by every workgroup, acquire by the last).
__global__ __launch_bounds__(256) void k_reduce_small(const double *__restrict__ part, int nwg,
                                                      int grp, int64_t E, double *out,
                                                      int *ticket, SmallArgs sm)
{
    __shared__ double red[kTailMaxGroups][32];
    __shared__ double lds[kSmallLds];
    __shared__ int last;
    const double t = reduce_fixed_order(part, nwg, grp, E, red);
    const int64_t e = (int64_t)blockIdx.x * 32 + (threadIdx.x & 31);
    if ((threadIdx.x >> 5) == 0 && e < E) out[e] = t;
    // Release this workgroup's outputs, take a ticket; the last workgroup
    // acquires everyone's and runs the small step.
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(ticket, 1) == (int)gridDim.x - 1 ? 1 : 0;
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (threadIdx.x == 0) *ticket = 0;
    small_step((int)blockDim.x, sm, lds);
}

hipError_t launch_reduce_small(const double *part, int nwg, int64_t na, int64_t nb, double *out,
                               int *ticket, const SmallArgs &sm, hipStream_t st)
{
    const int64_t E = na * nb;
    if (E <= 0) return hipErrorInvalidValue;
    k_reduce_small<<<dim3((unsigned)((E + 31) / 32)), dim3(256), 0, st>>>(
        part, nwg, tail_group(nwg), E, out, ticket, sm);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Small-matrix steps on their own (one workgroup of 256 threads, FP64)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_small(SmallArgs a)
{
    __shared__ double lds[kSmallLds];
    small_step((int)blockDim.x, a, lds);
}

hipError_t launch_small(const SmallArgs &a, hipStream_t st)
{
    k_small<<<dim3(1), dim3(256), 0, st>>>(a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fused pass: updates + projection in one sweep over the row slab
// ---------------------------------------------------------------------------

constexpr int kPT = kPassRows;                   // rows per tile (32)
constexpr int kPLd = kPT + 1;                    // padded column stride (doubles)
constexpr int kPF = kPassMaxCols * kPT / 256;    // prefetch registers per thread
constexpr int kPassMaxTiles = 12;                // projection tiles per wave
constexpr int kPassMaxC = 8192;                  // doubles of C held in LDS

size_t pass_lds_bytes(const PassArgs &a, int F)
{
    if (a.ncols > kPassMaxCols || a.wl < a.ncols || a.wl > kPassMaxCols + 32) return 0;
    const int k1p = (a.k1 + 3) & ~3;
    if ((size_t)k1p * a.n1 > (size_t)kPassMaxC || k1p > a.wl) return 0;
    if (a.k1 > 0 && (a.n1 % 16 != 0 || (F * kPT / 16) * (a.n1 / 16) > 8)) return 0;
    if (a.na > 0) {
        const int Ma = (a.na + 15) / 16, Nb = (a.nb + 15) / 16;
        if (Ma * Nb > 4 * kPassMaxTiles) return 0;
        if (a.a0 + Ma * 16 > a.wl || a.b0 + Nb * 16 > a.wl) return 0;
    }
    for (int g = 0; g < a.ngrp; g++)
        if (a.g[g].w > kMaxS || a.g[g].t0 + a.g[g].w > a.ncols) return 0;
    return ((size_t)a.wl * (F * kPT + 1) + (size_t)k1p * a.n1 + 2 * kMaxS * kMaxS) * sizeof(double);
}

// Waves per SIMD (= 256-thread workgroups per CU) the generic pass is
// compiled for: the wide instance uses up to 512 registers per lane (one
// workgroup per CU); the narrow one fits three (S <= 8) or two (S = 16).
constexpr int pass_waves(int S, int MT) { return MT > 4 || S > 16 ? 1 : S <= 8 ? 3 : 2; }

// MT: projection tiles per wave, PF: prefetch registers per thread.  The
// narrow instance (fewer of both) fits several workgroups per CU.  V2: every
// lane moves two consecutive rows with one 16-byte load (even ldX and m,
// 16-byte aligned slab).  F: rows per tile = 32 F (the narrow instance with
// few columns takes longer runs of each column per tile).
template <int S, int MT = kPassMaxTiles, int PF = kPF, bool V2 = false, int F = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    pass_waves(S, MT), pass_waves(S, MT)))) void k_pass(PassArgs a)
{
    constexpr int TS = kPT * F, PLD = TS + 1;
    extern __shared__ double sm[];
    const uint64_t kst = kernarg_stash();
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int k1p = (a.k1 + 3) & ~3;
    double *tile = sm;
    double *Cs = tile + a.wl * PLD;
    double *Ts = Cs + k1p * a.n1;  // one S x S factor per group, ld S

    for (int e = tid; e < k1p * a.n1; e += 256) Cs[e] = a.C[e];
    for (int g = 0; g < a.ngrp; g++) {
        const double *T = a.g[g].T;
        const int w = a.g[g].w;
        if (T)
            for (int e = tid; e < S * S; e += 256) {
                const int i = e % S, j = e / S;
                Ts[g * S * S + e] = (i < w && j < w) ? T[i + j * kLdT] : 0.0;
            }
    }
    for (int e = tid; e < (a.wl - a.ncols) * TS; e += 256)
        tile[(a.ncols + e / TS) * PLD + (e % TS)] = 0.0;

    const int Ma = (a.na + 15) >> 4, Nb = (a.nb + 15) >> 4;
    const int ntiles = a.na > 0 ? Ma * Nb : 0;

    dbl4 acc[MT];
#pragma unroll
    for (int t = 0; t < MT; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};

    const int64_t rb = (int64_t)blockIdx.x * a.rows_per;
    const int64_t re = min(a.m, rb + a.rows_per);
    // Load step i moves columns lcol + CPS i: LPC = TS / VR lanes per column,
    // VR consecutive rows per lane, so each group of LPC lanes reads TS
    // consecutive rows of one column.  A wave's lanes touch at most four
    // consecutive columns; the loads are addressed from a wave-uniform base
    // (the wave's first column) plus one 32-bit byte offset per lane (below
    // 8 (3 ldX + TS) bytes; the host keeps ldX under 2^27), so no per-step
    // 64-bit lane address is kept.  Every lane loads unconditionally (one
    // code path, so no register copies or zero defaults wait for loads or
    // stores in flight): lanes past the last row re-read the last valid
    // row(s), and the LDS store zeroes them.  With V2 the row count of every
    // tile is even.
    typedef double dbl2v __attribute__((ext_vector_type(2)));
    using PV = typename std::conditional<V2, dbl2v, double>::type;
    constexpr int VR = V2 ? 2 : 1, NPF = PF / VR;
    constexpr int LPC = TS / VR, CPS = 256 / LPC;
    static_assert(LPC >= 16 && LPC <= 256 && 256 % LPC == 0, "lane mapping");
    PV pf[NPF];
    const int lrow = VR * (tid % LPC), lcol = tid / LPC;
    const int wcol = __builtin_amdgcn_readfirstlane(lcol);
    const uint32_t lb = 8u * (uint32_t)((lcol - wcol) * a.ldX + lrow);
    const int64_t wcolb = (int64_t)wcol * a.ldX * 8, cstep = (int64_t)CPS * a.ldX * 8;
    auto load = [&](int64_t r0) {
        const int rem = (int)min(re - r0, (int64_t)TS);
        // Opaque to the optimizer, so the addresses are not hoisted.
        uint32_t o = lb - 8u * (uint32_t)max(0, lrow - (rem - VR));
        asm volatile("" : "+v"(o));
        const char *b = (const char *)(a.X + r0) + wcolb;
#pragma unroll
        for (int i = 0; i < NPF; i++)
            if (lcol + CPS * i < a.ncols) pf[i] = *(const PV *)(b + i * cstep + o);
    };
    if (rb < re) load(rb);
    __syncthreads();

    for (int64_t r0 = rb; r0 < re; r0 += TS) {
        __syncthreads();
        const bool rok = lrow < min(re - r0, (int64_t)TS);
#pragma unroll
        for (int i = 0; i < NPF; i++)
            if (lcol + CPS * i < a.ncols) {
                double *d = tile + (lcol + CPS * i) * PLD + lrow;
                if constexpr (V2) {
                    d[0] = rok ? pf[i].x : 0.0;
                    d[1] = rok ? pf[i].y : 0.0;
                } else {
                    d[0] = rok ? pf[i] : 0.0;
                }
            }
        __syncthreads();
        if (r0 + TS < re) load(r0 + TS);

        if (a.ngrp > 0) {
            // Step 1 in one phase (all groups from the same product) or, in
            // sequential mode, one phase per group.
            const int nphase = a.seq ? a.ngrp : 1;
            for (int p = 0; p < nphase; p++) {
                const int glo = a.seq ? p : 0, ghi = a.seq ? p + 1 : a.ngrp;
                int zlo = 1 << 30, zhi = 0;
                for (int g = glo; g < ghi; g++) {
                    zlo = min(zlo, a.g[g].zoff);
                    zhi = max(zhi, a.g[g].zoff + a.g[g].w);
                }
                const int ntlo = zlo >> 4, nN = ((zhi + 15) >> 4) - ntlo;
                const int kp = a.seq ? ((a.g[p].k + 3) & ~3) : k1p;
                const int ncombo = a.k1 > 0 ? (TS / 16) * nN : 0;
                // Z = tile(:, 0:kp) C(:, N-tiles), one 16 x 16 tile of Z per combo.
                dbl4 z[2];
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    z[q] = dbl4{0.0, 0.0, 0.0, 0.0};
                    const int combo = wave + 4 * q;
                    if (combo < ncombo) {
                        const int mt = combo / nN, nt = ntlo + combo % nN;
                        const double *ap = tile + lk * PLD + mt * 16 + li;
                        const double *bp = Cs + lk * a.n1 + nt * 16 + li;
                        for (int ks = 0; ks < kp / 4; ks++)
                            z[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                ap[4 * ks * PLD], bp[4 * ks * a.n1], z[q], 0, 0, 0);
                    }
                }
                __syncthreads();
                // Subtract Z from the target blocks (D layout: row lk + 4v, col li).
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const int combo = wave + 4 * q;
                    if (combo < ncombo) {
                        const int mt = combo / nN, nt = ntlo + combo % nN;
                        const int zc = nt * 16 + li;
                        for (int g = glo; g < ghi; g++) {
                            const int off = zc - a.g[g].zoff;
                            if (off >= 0 && off < a.g[g].w) {
                                double *col = tile + (a.g[g].t0 + off) * PLD + mt * 16 + lk;
#pragma unroll
                                for (int v = 0; v < 4; v++) col[4 * v] -= z[q][v];
                            }
                        }
                    }
                }
                __syncthreads();
                // Row-wise forward substitution with each group's factor
                // (dtrsm_('R','U','N','N') on the rows of the tile).
                for (int task = tid; task < (ghi - glo) * TS; task += 256) {
                    const int g = glo + task / TS, row = task % TS;
                    if (a.g[g].T) {
                        const int w = a.g[g].w;
                        double *col = tile + a.g[g].t0 * PLD + row;
                        const double *Tg = Ts + g * S * S;
                        double x[S];
#pragma unroll
                        for (int i = 0; i < S; i++) x[i] = i < w ? col[i * PLD] : 0.0;
#pragma unroll
                        for (int i = 0; i < S; i++) {
                            if (i < w) {
                                double v = x[i];
#pragma unroll
                                for (int l = 0; l < i; l++) v = fma(-x[l], Tg[l + i * S], v);
                                x[i] = v * (1.0 / Tg[i + i * S]);
                            }
                        }
#pragma unroll
                        for (int i = 0; i < S; i++)
                            if (i < w) col[i * PLD] = x[i];
                    }
                }
                __syncthreads();
            }
            // Write the finished target blocks back to X.
            for (int g = 0; g < a.ngrp; g++) {
                const int t0 = a.g[g].t0;
                for (int e = tid; e < a.g[g].w * TS; e += 256) {
                    const int cc = e / TS, rr = e - cc * TS;
                    const int64_t row = r0 + rr;
                    if (row < re) a.X[row + (int64_t)(t0 + cc) * a.ldX] = tile[(t0 + cc) * PLD + rr];
                }
            }
        }

        // Step 2: projection partials.
        if (ntiles > 0) {
#pragma unroll
            for (int ks = 0; ks < TS / 4; ks++) {
                const int rr = 4 * ks + lk;
#pragma unroll
                for (int t = 0; t < MT; t++) {
                    const int tl = wave + 4 * t;
                    if (tl < ntiles) {
                        const int ta = tl / Nb, tb = tl - ta * Nb;
                        const double av = tile[(a.a0 + ta * 16 + li) * PLD + rr];
                        const double bv = tile[(a.b0 + tb * 16 + li) * PLD + rr];
                        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
                    }
                }
            }
        }
    }

    if (ntiles > 0) {
        double *out = a.part + (int64_t)blockIdx.x * a.na * a.nb;
#pragma unroll
        for (int t = 0; t < MT; t++) {
            const int tl = wave + 4 * t;
            if (tl < ntiles) {
                const int ta = tl / Nb, tb = tl - ta * Nb;
                const int j = tb * 16 + li;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = ta * 16 + lk + 4 * v;
                    if (i < a.na && j < a.nb) out[i + (int64_t)j * a.na] = acc[t][v];
                }
            }
        }
    }
    pass_tail(a, sm, kst, 256);
}

// Block width 16, the benchmark shape: compile-time tile counts, 32-bit
// LDS offsets, four independent MFMA chains in step 1.  Wave w computes
// Z tile (rows 16 (w & 1), cols 16 (w >> 1)) in step 1 and projection tiles
// [w MT, (w+1) MT) x [0, NB) in step 2.
template <int MT, int NB>
__global__ __launch_bounds__(256) void k_pass16(PassArgs a)
{
    constexpr int S = 16, N1 = 32, LD = kPLd;
    extern __shared__ double sm[];
    const uint64_t kst = kernarg_stash();
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int k1p = (a.k1 + 3) & ~3;
    const int C0 = a.wl * LD, T0 = C0 + k1p * N1;

    for (int e = tid; e < k1p * N1; e += 256) sm[C0 + e] = a.C[e];
    for (int g = 0; g < a.ngrp; g++) {
        const double *T = a.g[g].T;
        const int w = a.g[g].w;
        if (T)
            for (int e = tid; e < S * S; e += 256) {
                const int i = e % S, j = e / S;
                sm[T0 + g * S * S + e] = (i < w && j < w) ? T[i + j * kLdT] : 0.0;
            }
    }
    for (int e = tid; e < (a.wl - a.ncols) * kPT; e += 256)
        sm[(a.ncols + e / kPT) * LD + (e % kPT)] = 0.0;

    const int Ma = (a.na + 15) >> 4;
    const bool proj = a.na > 0;
    const int mt1 = wave & 1, nt1 = wave >> 1;
    const int aoff1 = lk * LD + mt1 * 16 + li;
    const int boff1 = C0 + lk * N1 + nt1 * 16 + li;
    int aoff2[MT], boff2[NB];
#pragma unroll
    for (int i = 0; i < MT; i++) aoff2[i] = (a.a0 + (wave * MT + i) * 16 + li) * LD + lk;
#pragma unroll
    for (int j = 0; j < NB; j++) boff2[j] = (a.b0 + j * 16 + li) * LD + lk;

    dbl4 acc[MT][NB];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    const int64_t rb = (int64_t)blockIdx.x * a.rows_per;
    const int64_t re = min(a.m, rb + a.rows_per);
    const int nload = a.ncols * kPT;
    const double *Xl = a.X + (tid & (kPT - 1));
    double pf[kPF];
    auto load = [&](int64_t r0) {
        const bool full = r0 + kPT <= re;
#pragma unroll
        for (int i = 0; i < kPF; i++) {
            const int e = tid + 256 * i;
            if (e < nload) {
                const int64_t off = r0 + (int64_t)(e >> 5) * a.ldX;
                pf[i] = (full || r0 + (e & 31) < re) ? Xl[off] : 0.0;
            }
        }
    };
    if (rb < re) load(rb);

    for (int64_t r0 = rb; r0 < re; r0 += kPT) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kPF; i++) {
            const int e = tid + 256 * i;
            if (e < nload) sm[(e >> 5) * LD + (e & 31)] = pf[i];
        }
        __syncthreads();
        if (r0 + kPT < re) load(r0 + kPT);

        if (a.ngrp > 0) {
            const int nphase = a.seq ? a.ngrp : 1;
            for (int p = 0; p < nphase; p++) {
                const int glo = a.seq ? p : 0, ghi = a.seq ? p + 1 : a.ngrp;
                const int kp = a.seq ? ((a.g[p].k + 3) & ~3) : k1p;
                const bool mine = a.k1 > 0 && (!a.seq || nt1 == (a.g[p].zoff >> 4));
                dbl4 z = dbl4{0.0, 0.0, 0.0, 0.0};
                if (mine) {
                    dbl4 z1 = z, z2 = z, z3 = z;
                    const int nks = kp >> 2;
                    int ks = 0;
                    for (; ks + 4 <= nks; ks += 4) {
                        const int ao = aoff1 + 4 * ks * LD, bo = boff1 + 4 * ks * N1;
                        const double x0 = sm[ao], x1 = sm[ao + 4 * LD], x2 = sm[ao + 8 * LD],
                                     x3 = sm[ao + 12 * LD];
                        const double c0 = sm[bo], c1 = sm[bo + 4 * N1], c2 = sm[bo + 8 * N1],
                                     c3 = sm[bo + 12 * N1];
                        z = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, c0, z, 0, 0, 0);
                        z1 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, c1, z1, 0, 0, 0);
                        z2 = __builtin_amdgcn_mfma_f64_16x16x4f64(x2, c2, z2, 0, 0, 0);
                        z3 = __builtin_amdgcn_mfma_f64_16x16x4f64(x3, c3, z3, 0, 0, 0);
                    }
                    for (; ks < nks; ks++)
                        z = __builtin_amdgcn_mfma_f64_16x16x4f64(sm[aoff1 + 4 * ks * LD],
                                                                 sm[boff1 + 4 * ks * N1], z, 0,
                                                                 0, 0);
                    z = (z + z1) + (z2 + z3);
                }
                __syncthreads();
                if (mine) {
                    const int zc = nt1 * 16 + li;
                    for (int g = glo; g < ghi; g++) {
                        const int off = zc - a.g[g].zoff;
                        if (off >= 0 && off < a.g[g].w) {
                            const int o = (a.g[g].t0 + off) * LD + mt1 * 16 + lk;
#pragma unroll
                            for (int v = 0; v < 4; v++) sm[o + 4 * v] -= z[v];
                        }
                    }
                }
                __syncthreads();
                // Row-wise forward substitution (dtrsm_('R','U','N','N')).
                if (tid < (ghi - glo) * kPT) {
                    const int g = glo + tid / kPT, row = tid % kPT;
                    if (a.g[g].T) {
                        const int w = a.g[g].w;
                        const int o = a.g[g].t0 * LD + row;
                        const int tg = T0 + g * S * S;
                        double x[S];
#pragma unroll
                        for (int i = 0; i < S; i++) x[i] = i < w ? sm[o + i * LD] : 0.0;
#pragma unroll
                        for (int i = 0; i < S; i++) {
                            if (i < w) {
                                double v = x[i];
#pragma unroll
                                for (int l = 0; l < i; l++) v = fma(-x[l], sm[tg + l + i * S], v);
                                x[i] = v * (1.0 / sm[tg + i + i * S]);
                            }
                        }
#pragma unroll
                        for (int i = 0; i < S; i++)
                            if (i < w) sm[o + i * LD] = x[i];
                    }
                }
                __syncthreads();
            }
            for (int g = 0; g < a.ngrp; g++) {
                const int t0 = a.g[g].t0;
                for (int e = tid; e < a.g[g].w * kPT; e += 256) {
                    const int cc = e >> 5, rr = e & 31;
                    if (r0 + rr < re)
                        a.X[r0 + rr + (int64_t)(t0 + cc) * a.ldX] = sm[(t0 + cc) * LD + rr];
                }
            }
        }

        if (proj) {
#pragma unroll
            for (int ks = 0; ks < kPT / 4; ks++) {
                double bv[NB], av[MT];
#pragma unroll
                for (int j = 0; j < NB; j++) bv[j] = sm[boff2[j] + 4 * ks];
#pragma unroll
                for (int i = 0; i < MT; i++) av[i] = sm[aoff2[i] + 4 * ks];
#pragma unroll
                for (int i = 0; i < MT; i++)
#pragma unroll
                    for (int j = 0; j < NB; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j],
                                                                        0, 0, 0);
            }
        }
    }

    if (proj) {
        double *out = a.part + (int64_t)blockIdx.x * a.na * a.nb;
#pragma unroll
        for (int i = 0; i < MT; i++) {
            const int ta = wave * MT + i;
            if (ta < Ma) {
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const int col = j * 16 + li;
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int row = ta * 16 + lk + 4 * v;
                        if (row < a.na && col < a.nb) out[row + (int64_t)col * a.na] = acc[i][j][v];
                    }
                }
            }
        }
    }
    pass_tail(a, sm, kst, 256);
}

int pass16b_rows(const PassArgs &a)
{
    for (int f = 8; f >= 1; f >>= 1) {
        const int ts = 32 * f;
        if ((int64_t)a.wl * (ts + 1) <= kBTile && a.ncols * ts <= kBPF * kB) return ts;
    }
    return 0;
}

size_t pass16b_lds_bytes(const PassArgs &a, int ts)
{
    return ((size_t)a.wl * (ts + 1) + 1024) * sizeof(double);
}

// 16-byte loads when the slab allows them: even leading dimension and a
// 16-byte aligned base (rows_per and the first row of every chunk are even).
template <int MT, int NB, int F, bool EARLY = false>
hipError_t launch_pass16b_t(const PassArgs &a, size_t lds, int nwg, hipStream_t st)
{
    const bool v2 = (a.ldX % 2 == 0) && ((uintptr_t)a.X % 16 == 0) && (a.rows_per % 2 == 0);
    return v2 ? launch_pass16b_v<MT, NB, F, EARLY, true>(a, lds, nwg, st)
              : launch_pass16b_v<MT, NB, F, EARLY, false>(a, lds, nwg, st);
}

// Passes without a projection.
hipError_t launch_pass16b_noproj(const PassArgs &a, int ts, size_t lds, int nwg, hipStream_t st)
{
    switch (ts) {
    case 32: return launch_pass16b_t<1, 1, 1, true>(a, lds, nwg, st);
    case 64: return launch_pass16b_t<1, 1, 2, true>(a, lds, nwg, st);
    case 128: return launch_pass16b_t<1, 1, 4, true>(a, lds, nwg, st);
    case 256: return launch_pass16b_t<1, 1, 8, true>(a, lds, nwg, st);
    default: return hipErrorInvalidValue;
    }
}

// Only the (MT, F) pairs the plan can produce are instantiated: a tile of
// TS = 32 F rows holds at most 8192 / TS columns, so MT <= 4 / F.
template <int MT, int NB>
hipError_t launch_pass16b_f(const PassArgs &a, int ts, size_t lds, int nwg, hipStream_t st)
{
    switch (ts) {
    case 32: return launch_pass16b_t<MT, NB, 1>(a, lds, nwg, st);
    case 64:
        if constexpr (MT <= 2) return launch_pass16b_t<MT, NB, 2>(a, lds, nwg, st);
        break;
    case 128:
        if constexpr (MT == 1) return launch_pass16b_t<MT, NB, 4>(a, lds, nwg, st);
        break;
    case 256:
        if constexpr (MT == 1) return launch_pass16b_t<MT, NB, 8>(a, lds, nwg, st);
        break;
    default: break;
    }
    return hipErrorInvalidValue;
}

// Super-tile rows of the 512-thread s = 16 pass, or 0 when it does not apply.
int pass16b_plan(const PassArgs &a, int *mt, int *nb)
{
    if (a.n1 != 32 || (a.na > 0 && a.nb > 32)) return 0;
    const int ts = pass16b_rows(a);
    if (ts == 0) return 0;
    if (pass16b_lds_bytes(a, ts) > 160 * 1024) return 0;
    if (a.na > 0) {
        const int Ma = (a.na + 15) / 16;
        if (a.a0 + Ma * 16 > a.wl || a.b0 + ((a.nb + 15) / 16) * 16 > a.wl) return 0;
        *nb = a.nb > 16 ? 2 : 1;
        *mt = (Ma + (*nb == 2 ? 3 : 7)) / (*nb == 2 ? 4 : 8);
        if (*mt > 4) return 0;
    } else {
        *nb = 1;
        *mt = 1;
    }
    for (int g = 0; g < a.ngrp; g++)
        if (a.g[g].w > 16 || a.g[g].t0 + a.g[g].w > a.ncols) return 0;
    return ts;
}

hipError_t launch_pass16b(const PassArgs &a, int nwg, hipStream_t st)
{
    int mt = 0, nb = 0;
    const int ts = pass16b_plan(a, &mt, &nb);
    if (ts == 0) return hipErrorInvalidValue;
    const size_t lds = pass16b_lds_bytes(a, ts);
    if (a.na == 0) return launch_pass16b_noproj(a, ts, lds, nwg, st);
#define ORTH_P16B(M)                                                                 \
    case M:                                                                          \
        return nb == 2 ? launch_pass16b_f<M, 2>(a, ts, lds, nwg, st)                \
                       : launch_pass16b_f<M, 1>(a, ts, lds, nwg, st);
    switch (mt) {
        ORTH_P16B(1)
        ORTH_P16B(2)
        ORTH_P16B(3)
        ORTH_P16B(4)
    default: return hipErrorInvalidValue;
    }
#undef ORTH_P16B
}

// The double-buffered s = 16 pass (orth_pass16d.h) with 32-row tiles: waves
// per workgroup (8: two workgroups per CU, when both buffers of two
// workgroups fit a CU's LDS; 16: one workgroup per CU, up to the whole
// 160 KB), or 0 when it does not apply.  It needs FP64 data with an even
// row count and leading dimension and a 16-byte aligned slab (16-byte loads
// of row pairs), the folded (non-sequential) step 1, and at least
// kPass16dMinCols columns: measured at m = 2^24, n = 256, it beats k_pass16b
// on the passes over 128 columns and more (by up to 0.8 ms each) and loses
// on the narrower ones, where k_pass16b runs two workgroups per CU with 64-
// to 256-row super-tiles (ORTH_PASS16D_MINCOLS overrides the bound).
constexpr int kPass16dMinCols = 128;

int pass16d_plan(const PassArgs &a, int *mt, int *nb, bool any_width, int min_cols)
{
    if (getenv("ORTH_PASS_REGPF")) return 0;
    if (a.n1 != 32 || a.seq || (a.na > 0 && a.nb > 32)) return 0;
    if (a.X32) {
        // Mixed form: at most 32 raw (FP32) columns, none of them a step-1
        // operand, a 16-column aligned frontier, no reduced-precision update.
        const int ndma = std::min(a.ncols, a.raw0);
        if (a.raw0 % 16 != 0 || a.ncols - ndma > 32 || ((a.k1 + 3) & ~3) > ndma) return 0;
        for (int g = 0; g < a.ngrp; g++)
            if (a.g[g].lp) return 0;
    }
    // Row pairs are loaded whole: an odd row count would read past the last
    // row of the last column.
    if (a.ldX % 2 != 0 || a.m % 2 != 0 || (uintptr_t)a.X % 16 != 0) return 0;
    int minc = any_width ? 1 : min_cols > 0 ? min_cols : kPass16dMinCols;
    if (const char *e = getenv("ORTH_PASS16D_MINCOLS")) minc = atoi(e);
    if (a.ncols < minc || a.ncols > a.wl) return 0;
    const bool force16 = getenv("ORTH_PASS16D_ONEWG") != nullptr;
    int nw = 0;
    if (!force16 && 2 * pass16d_lds(a.wl, 8) <= kDLds) nw = 8;
    else if (pass16d_lds(a.wl, 16) <= kDLds) nw = 16;
    if (nw == 0) return 0;
    const int ks = nw / 4;
    const int nks = a.k1 > 0 ? ((a.k1 + 3) & ~3) / 4 : 0;
    if ((nks + ks - 1) / ks > kDKq || (a.k1 > 0 && ((a.k1 + 3) & ~3) > a.wl)) return 0;
    // Each group's Z columns are one whole Z tile (zoff 0 or 16, distinct):
    // the wave that sums that tile solves the group.
    for (int g = 0; g < a.ngrp; g++)
        if (a.g[g].w > 16 || a.g[g].t0 + a.g[g].w > a.ncols || a.g[g].t0 % 16 != 0 ||
            a.g[g].t0 + 16 > a.wl || (a.g[g].zoff != 0 && a.g[g].zoff != 16))
            return 0;
    if (a.ngrp == 2 && a.g[0].zoff == a.g[1].zoff) return 0;
    *mt = 1;
    *nb = 1;
    if (a.na > 0) {
        const int Ma = (a.na + 15) / 16, Nb = (a.nb + 15) / 16;
        if (a.a0 + Ma * 16 > a.wl || a.b0 + Nb * 16 > a.wl) return 0;
        // The LDS bound (wl <= 128 with 8 waves, <= 256 with 16) keeps Ma
        // within one M-tile per wave.
        *mt = (Ma + nw - 1) / nw;
        *nb = Nb;
        if (*mt > 1) return 0;
    }
    return nw;
}

template <int NW, int NB>
hipError_t launch_pass16d_t(const PassArgs &a, int nwg, hipStream_t st)
{
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)k_pass16d<NW, NB, 0>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(160 * 1024));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const size_t lds = (size_t)pass16d_lds(a.wl, NW) * sizeof(double);
    k_pass16d<NW, NB, 0><<<dim3(nwg), dim3(64 * NW), pass_tail_lds(a, lds), st>>>(a);
    return hipGetLastError();
}

template <int NW>
hipError_t launch_pass16d_nw(const PassArgs &a, int mt, int nb, int nwg, hipStream_t st)
{
    if (mt != 1) return hipErrorInvalidValue;
    return nb == 2 ? launch_pass16d_t<NW, 2>(a, nwg, st) : launch_pass16d_t<NW, 1>(a, nwg, st);
}

hipError_t launch_pass16d(const PassArgs &a, int nw, int nb, int nwg, hipStream_t st)
{
    if ((nw != 8 && nw != 16) || a.rows_per % 32 != 0) return hipErrorInvalidValue;
    return nw == 16 ? launch_pass16d_nw<16>(a, 1, nb, nwg, st) : launch_pass16d_nw<8>(a, 1, nb, nwg, st);
}

// Tiles per wave of the s = 16 kernel for a pass, or 0 when it does not apply.
static int pass16_mt(const PassArgs &a)
{
    if (a.n1 != 32) return 0;
    if (a.na == 0) return 1;
    const int Ma = (a.na + 15) / 16, Nb = (a.nb + 15) / 16;
    if (Nb > 2) return 0;
    const int mt = (Ma + 3) / 4;
    if (mt > 5 || a.a0 + 64 * mt > a.wl) return 0;
    return mt;
}

template <int MT, int NB>
hipError_t launch_pass16_t(const PassArgs &a, size_t lds, int nwg, hipStream_t st)
{
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)k_pass16<MT, NB>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(160 * 1024));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    k_pass16<MT, NB><<<dim3(nwg), dim3(256), pass_tail_lds(a, lds), st>>>(a);
    return hipGetLastError();
}

hipError_t launch_pass16(const PassArgs &a, size_t lds, int nwg, hipStream_t st)
{
    const int mt = pass16_mt(a);
    const int nb = a.na > 0 ? (a.nb + 15) / 16 : 1;
#define ORTH_P16(M)                                                                  \
    case M:                                                                          \
        return nb == 2 ? launch_pass16_t<M, 2>(a, lds, nwg, st)                     \
                       : launch_pass16_t<M, 1>(a, lds, nwg, st);
    switch (mt) {
        ORTH_P16(1)
        ORTH_P16(2)
        ORTH_P16(3)
        ORTH_P16(4)
        ORTH_P16(5)
    default: return hipErrorInvalidValue;
    }
#undef ORTH_P16
}

// The narrow generic pass: at most kNarrowCols columns and 4 projection
// tiles per wave.
constexpr int kNarrowCols = 160;
constexpr int kNarrowPF = kNarrowCols * kPT / 256;

bool pass_narrow(const PassArgs &a)
{
    const int Ma = (a.na + 15) / 16, Nb = (a.nb + 15) / 16;
    return a.ncols <= kNarrowCols && (a.na == 0 || Ma * Nb <= 4 * 4);
}

template <int S, int MT, int PF, bool V2 = false, int F = 1>
hipError_t launch_pass_v(const PassArgs &a, size_t lds, int nwg, hipStream_t st)
{
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)k_pass<S, MT, PF, V2, F>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(160 * 1024));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    k_pass<S, MT, PF, V2, F><<<dim3(nwg), dim3(256), pass_tail_lds(a, lds), st>>>(a);
    return hipGetLastError();
}

// Rows per tile of the generic pass, 32 F: the narrow instance takes F = 4
// (S <= 8) or 2 (S <= 16) when the tile's columns fit its prefetch
// registers (ncols * 32 F <= 256 * kNarrowPF) and the step-1 tiles fit.
int pass_f(int S, const PassArgs &a)
{
    if (!pass_narrow(a) || getenv("ORTH_PASS_TILE32")) return 1;
    for (int F = S <= 8 ? 4 : S <= 16 ? 2 : 1; F > 1; F >>= 1)
        if (a.ncols * kPT * F <= 256 * kNarrowPF && pass_lds_bytes(a, F) != 0) return F;
    return 1;
}

int pass_rows(int S, const PassArgs &a)
{
    if (S == 16 && pass16_mt(a) > 0 && !getenv("ORTH_GENERIC_PASS")) return kPT;  // k_pass16
    return kPT * pass_f(S, a);
}

// Two-row (16-byte) loads: even leading dimension and row count, 16-byte
// aligned slab.
bool pass_v2_ok(const PassArgs &a)
{
    return a.ldX % 2 == 0 && a.m % 2 == 0 && (uintptr_t)a.X % 16 == 0 &&
           !getenv("ORTH_PASS_LOAD8");
}

template <int S, int F>
hipError_t launch_pass_n(const PassArgs &a, size_t lds, int nwg, hipStream_t st)
{
    return pass_v2_ok(a) ? launch_pass_v<S, 4, kNarrowPF, true, F>(a, lds, nwg, st)
                         : launch_pass_v<S, 4, kNarrowPF, false, F>(a, lds, nwg, st);
}

template <int S>
hipError_t launch_pass_t(const PassArgs &a, int nwg, hipStream_t st)
{
    const int F = pass_f(S, a);
    const size_t lds = pass_lds_bytes(a, F);
    if (lds == 0 || a.rows_per % (kPT * F) != 0) return hipErrorInvalidValue;
    if (pass_narrow(a)) {
        if constexpr (S <= 8)
            if (F == 4) return launch_pass_n<S, 4>(a, lds, nwg, st);
        if constexpr (S <= 16)
            if (F == 2) return launch_pass_n<S, 2>(a, lds, nwg, st);
        return launch_pass_n<S, 1>(a, lds, nwg, st);
    }
    return launch_pass_v<S, kPassMaxTiles, kPF>(a, lds, nwg, st);
}

int pass_wg_per_cu(int S, const PassArgs &a, bool tail)
{
    const size_t lds = pass_lds_bytes(a, pass_f(S, a));
    if (lds == 0) return 1;
    // Register-limited residency (k_pass16 and the wide generic pass: one
    // workgroup per CU).
    const bool p16 = S == 16 && pass16_mt(a) > 0 && !getenv("ORTH_GENERIC_PASS");
    const int by_regs = p16 ? 1 : pass_narrow(a) ? pass_waves(S, 4) : pass_waves(S, kPassMaxTiles);
    // The in-pass tail raises small LDS footprints to its own.
    const int by_lds =
        (int)std::max<size_t>(1, (160 * 1024) / (tail ? std::max(lds, kTailLdsBytes) : lds));
    return std::max(1, std::min(by_regs, by_lds));
}

size_t pass_tail_lds(const PassArgs &a, size_t lds)
{
    return a.tail_on ? std::max(lds, kTailLdsBytes) : lds;
}

hipError_t launch_pass(int S, const PassArgs &a, int nwg, hipStream_t st)
{
    if (S == 16 && pass16_mt(a) > 0 && !getenv("ORTH_GENERIC_PASS")) {
        const size_t lds = pass_lds_bytes(a);
        if (lds == 0) return hipErrorInvalidValue;
        return launch_pass16(a, lds, nwg, st);
    }
    switch (S) {
    case 4: return launch_pass_t<4>(a, nwg, st);
    case 8: return launch_pass_t<8>(a, nwg, st);
    case 16: return launch_pass_t<16>(a, nwg, st);
    case 32: return launch_pass_t<32>(a, nwg, st);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Streaming probes (measurement of the device's streaming ceiling)
// ---------------------------------------------------------------------------

typedef double dbl2p __attribute__((ext_vector_type(2)));

// dst = src over n2 16-byte elements, 8 non-temporal loads in flight per
// thread before the stores.
__global__ __launch_bounds__(256) void k_probe_copy(const dbl2p *__restrict__ src,
                                                    dbl2p *__restrict__ dst, int64_t n2)
{
    constexpr int U = 8;
    const int64_t stride = (int64_t)gridDim.x * 256 * U;
    for (int64_t base = (int64_t)blockIdx.x * 256 * U + threadIdx.x; base < n2; base += stride) {
        dbl2p v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t i = base + (int64_t)u * 256;
            v[u] = i < n2 ? __builtin_nontemporal_load(src + i) : dbl2p{0.0, 0.0};
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t i = base + (int64_t)u * 256;
            if (i < n2) __builtin_nontemporal_store(v[u], dst + i);
        }
    }
}

// Reads n2 16-byte elements; one per-workgroup sum goes to part[blockIdx.x]
// (so that the loads are kept).
__global__ __launch_bounds__(256) void k_probe_read(const dbl2p *__restrict__ src, double *part,
                                                    int64_t n2)
{
    constexpr int U = 8;
    __shared__ double red[256];
    const int64_t stride = (int64_t)gridDim.x * 256 * U;
    double t = 0.0;
    for (int64_t base = (int64_t)blockIdx.x * 256 * U + threadIdx.x; base < n2; base += stride) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t i = base + (int64_t)u * 256;
            if (i < n2) {
                const dbl2p v = __builtin_nontemporal_load(src + i);
                t += v.x + v.y;
            }
        }
    }
    red[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < 256; k++) s += red[k];
        part[blockIdx.x] = s;
    }
}

hipError_t launch_stream_probe(int mode, double *dst, const double *src, int64_t bytes,
                               double *part, int nwg, hipStream_t st)
{
    const int64_t n2 = bytes / 16;
    if (n2 < 1 || nwg < 1) return hipErrorInvalidValue;
    if (mode == 0)
        k_probe_copy<<<dim3(nwg), dim3(256), 0, st>>>((const dbl2p *)src, (dbl2p *)dst, n2);
    else
        k_probe_read<<<dim3(nwg), dim3(256), 0, st>>>((const dbl2p *)src, part, n2);
    return hipGetLastError();
}

// Q (m x n, ld ldq) = X32 widened exactly to FP64.
__global__ __launch_bounds__(256) void k_widen(const float *__restrict__ x, int64_t ldx,
                                                double *__restrict__ q, int64_t ldq, int64_t m,
                                                int64_t n)
{
    const int64_t total = m * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % m, j = e / m;
        q[i + j * ldq] = (double)x[i + j * ldx];
    }
}

hipError_t launch_widen(const float *x, int64_t ldx, double *q, int64_t ldq, int64_t m, int64_t n,
                        hipStream_t st)
{
    const int64_t total = m * n;
    const int nb = (int)std::min<int64_t>((total + 255) / 256, 8192);
    k_widen<<<dim3(std::max(nb, 1)), dim3(256), 0, st>>>(x, ldx, q, ldq, m, n);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        p[i] = T(0);
}

hipError_t launch_fill(double *p, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    size_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    k_fill<double><<<dim3((unsigned)g), dim3(256), 0, st>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_fill_int(int *p, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    size_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    k_fill<int><<<dim3((unsigned)g), dim3(256), 0, st>>>(p, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Residual of the accuracy harness
// ---------------------------------------------------------------------------

constexpr int kResidWG = 1024;

int resid_wg_count(int64_t m)
{
    int64_t g = (m + 255) / 256;
    return (int)(g < kResidWG ? g : kResidWG);
}

__global__ __launch_bounds__(256) void k_resid(int64_t m, int n, const double *__restrict__ X,
                                               int64_t ldX, const double *__restrict__ Q,
                                               int64_t ldQ, const double *__restrict__ R,
                                               int64_t ldR, double *part2)
{
    __shared__ double red[2][256];
    double sr = 0.0, sx = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * 256) {
        for (int j0 = 0; j0 < n; j0 += 16) {
            double t[16];
#pragma unroll
            for (int jj = 0; jj < 16; jj++) {
                const int j = j0 + jj;
                const double x = j < n ? X[r + j * ldX] : 0.0;
                sx = fma(x, x, sx);
                t[jj] = -x;
            }
            for (int i = 0; i < n; i++) {
                const double q = Q[r + i * ldQ];
#pragma unroll
                for (int jj = 0; jj < 16; jj++) {
                    const int j = j0 + jj;
                    if (j < n) t[jj] = fma(q, R[i + j * ldR], t[jj]);
                }
            }
#pragma unroll
            for (int jj = 0; jj < 16; jj++) sr = fma(t[jj], t[jj], sr);
        }
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = sx;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            red[0][threadIdx.x] += red[0][threadIdx.x + h];
            red[1][threadIdx.x] += red[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part2[2 * blockIdx.x] = red[0][0];
        part2[2 * blockIdx.x + 1] = red[1][0];
    }
}

hipError_t launch_resid(int64_t m, int64_t n, const double *X, int64_t ldX, const double *Q,
                        int64_t ldQ, const double *R, int64_t ldR, double *part2,
                        hipStream_t st)
{
    k_resid<<<dim3(resid_wg_count(m)), dim3(256), 0, st>>>(m, (int)n, X, ldX, Q, ldQ, R, ldR,
                                                           part2);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Accuracy harness: the Gram Q^T Q with compensated accumulation (Dot2 of
// Ogita, Rump and Oishi, "Accurate sum and dot product", SIAM J. Sci. Comput.
// 26 (2005): TwoProduct by FMA, TwoSum, the error terms summed in FP64), so
// that the Gram's own rounding stays far below the loss of orthogonality it
// measures.  Workgroup (x, y): row chunk x, 64 x 64 block y of the upper
// block triangle; thread t owns the 16 pairs (I0 + 4 (t % 16) + di,
// J0 + 4 (t / 16) + dj).  part: (hi, lo) of each pair, per chunk and block.
// ---------------------------------------------------------------------------
constexpr int kGramRows = 32;

__global__ __launch_bounds__(256) void k_gram_dd(const double *__restrict__ Q, int64_t ldQ,
                                                 int64_t m, int n, int64_t rows_per,
                                                 double *__restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double sa[kGramRows][65], sb[kGramRows][65];
    const int t = threadIdx.x;
    const int nb = (n + 63) / 64;
    // Block y -> (I, J), I <= J, row-major over the upper block triangle.
    int y = blockIdx.y, I = 0;
    while (y >= nb - I) {
        y -= nb - I;
        I++;
    }
    const int J = I + y, I0 = 64 * I, J0 = 64 * J;
    const int ti = 4 * (t % 16), tj = 4 * (t / 16);
    double hi[16], lo[16];
#pragma unroll
    for (int k = 0; k < 16; k++) hi[k] = lo[k] = 0.0;
    const int64_t rb = (int64_t)blockIdx.x * rows_per, re = min(m, rb + rows_per);
    for (int64_t r0 = rb; r0 < re; r0 += kGramRows) {
        for (int e = t; e < kGramRows * 64; e += 256) {
            const int c = e / kGramRows, r = e % kGramRows;
            const bool in = r0 + r < re;
            sa[r][c] = in && I0 + c < n ? Q[r0 + r + (int64_t)(I0 + c) * ldQ] : 0.0;
            sb[r][c] = in && J0 + c < n ? Q[r0 + r + (int64_t)(J0 + c) * ldQ] : 0.0;
        }
        __syncthreads();
        for (int r = 0; r < kGramRows; r++) {
            double a[4], b[4];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                a[d] = sa[r][ti + d];
                b[d] = sb[r][tj + d];
            }
#pragma unroll
            for (int di = 0; di < 4; di++)
#pragma unroll
                for (int dj = 0; dj < 4; dj++) {
                    const int k = 4 * di + dj;
                    const double p = a[di] * b[dj];
                    const double pe = __builtin_fma(a[di], b[dj], -p);
                    const double sum = hi[k] + p;
                    const double bp = sum - hi[k];
                    const double se = (hi[k] - (sum - bp)) + (p - bp);
                    hi[k] = sum;
                    lo[k] += se + pe;
                }
        }
        __syncthreads();
    }
    double *o = part + ((int64_t)(blockIdx.x * gridDim.y + blockIdx.y) * 256 + t) * 32;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        o[2 * k] = hi[k];
        o[2 * k + 1] = lo[k];
    }
}

int gram_dd_chunks(int64_t m)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(64, (m + 4095) / 4096));
}

hipError_t launch_gram_dd(const double *Q, int64_t ldQ, int64_t m, int64_t n, double *part,
                          hipStream_t st)
{
    const int nch = gram_dd_chunks(m), nb = (int)((n + 63) / 64);
    int64_t rows = (m + nch - 1) / nch;
    rows = (rows + kGramRows - 1) / kGramRows * kGramRows;
    k_gram_dd<<<dim3(nch, nb * (nb + 1) / 2), dim3(256), 0, st>>>(Q, ldQ, m, (int)n, rows, part);
    return hipGetLastError();
}

}  // namespace orth
