// The FP64 fused pass for block width 16 with double-buffered direct-to-LDS
// loads (32-row tiles; 512 threads and two workgroups per CU for slabs of at
// most 128 columns, 1024 threads and one workgroup per CU up to 256).
//
// Two LDS buffers hold consecutive tiles: right after the barrier that opens
// tile i, every wave issues global_load_lds_dwordx4 for tile i+1 into the
// other buffer, so the next tile streams from HBM during all of tile i's
// arithmetic without holding a register.  Vector-memory loads complete in
// issue order and share one counter, so any other load in the tile body
// (a coefficient read from L2, a spill reload) would wait for the whole
// next tile: the step-1 coefficient fragments and the triangular-solve
// operands are loaded into registers once per pass, and the body's only
// vector-memory wait is the one that opens the next tile.
//
// LDS layout of a tile buffer: column c at c * kDLd doubles (kDLd = 34: 32
// rows and one 16-byte pad slot), its two 16-row halves swapped when c is
// odd (row r at c * kDLd + (r ^ 16 (c & 1)), see dpos).  A load instruction
// fills 64 consecutive 16-byte slots (lane l -> slot l of its span); lanes
// that land on a pad slot load a row pair of their column again.  With 64
// LDS banks serving 32 lanes of a 64-bit read per cycle, the stride of 34
// doubles spreads the 16 columns x 2 rows of a projection operand read over
// all banks, and the half swap does the same for the 16 rows x 2 columns of
// a step-1 operand read.  Every operand address of the pass is a per-lane
// base plus compile-time offsets.
//
// Per tile: step 1 (Z = tile(:, 0:k1) C on v_mfma_f64_16x16x4f64, the inner
// dimension split over the waves, partial products summed in a fixed
// order), then, by the wave that holds each summed 16 x 16 Z tile, the
// subtraction from its group's 16-row chunk and the chunk's blocked
// triangular solve (as in k_pass16b) in registers, the groups written back
// to X, and step 2 (the projection partial tile(:, a0:a0+na)^T
// tile(:, b0:b0+nb)).
#pragma once

#include "orth_device.h"
#include "orth_tail.h"

namespace orth {

constexpr int kDLds = 20480;  // doubles of LDS a CU has (160 KB)
constexpr int kDKq = 17;      // step-1 coefficient fragments per wave (k-steps)
constexpr int kDLd = 34;      // doubles per column of a tile buffer
constexpr int kDSl = kDLd / 2;  // 16-byte slots per column (16 row pairs + pad)

// Doubles of the step-1 partial-product scratch for NW waves (4 Z tiles of
// 16 x 16, the inner dimension split NW / 4 ways).
__host__ __device__ constexpr int dzs(int NW)
{
    return (NW / 4 - 1) * 4 * 256;
}

// Doubles of dynamic LDS for a slab of wl columns.
__host__ __device__ constexpr int pass16d_lds(int wl, int NW)
{
    return 2 * wl * kDLd + dzs(NW);
}

// Index (in doubles) of row r of column c in a tile buffer.
__device__ __forceinline__ int dpos(int c, int r)
{
    return c * kDLd + (r ^ ((c & 1) << 4));
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void glb_void_t;

// A copy of v the optimizer cannot see through: addresses derived from it
// inside the row loop are recomputed there instead of being hoisted out of
// it into registers that would stay live for the whole pass.
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Workgroup barrier over LDS: the wave's LDS accesses complete first, and
// nothing waits for its vector-memory operations, so the next tile's loads
// into the other buffer stay in flight.  (__syncthreads() is a workgroup
// fence as well, which waits for every load, the LDS-bound ones included.)
// The memory clobber keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// MX = 0: the FP64 form.  MX != 0 (mixed form): columns from raw0 on (at
// most 32, never step-1 operands) have not been written yet and come from the
// FP32 input X32; they are loaded into registers together with the next
// tile's DMA and written to LDS, widened, when that tile opens (a barrier
// separates that write from their first reader).  With MX a reduced precision
// (kMxBF16, kMxFP16, kMxBF16X3), the projection products whose B columns
// start at or past lpb run as one 16 x 16 x 32 MFMA per 32-row tile with FP32
// accumulation, summed across tiles in FP64; every other product is FP64.
template <int NW, int NB, int MX>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_pass16d(
    PassArgs a)
{
    constexpr int TS = 32, N1 = 32, kD = 64 * NW, LDc = kDLd;
    constexpr bool MIX = MX != 0, LPP = MX != 0 && MX != kMxFP64;
    constexpr int MXL = LPP ? MX : kMxBF16;  // operand format of the reduced-precision products
    // Step-1 work items: 2 x 2 Z tiles of 16 x 16, the inner dimension
    // split KS ways so that the items fill the NW waves.
    constexpr int ZT = 4, KS = NW / ZT;
    extern __shared__ double sm[];
    const uint64_t kst = kernarg_stash();
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lk = lane >> 4;
    const int tsz = a.wl * LDc;  // doubles per tile buffer
    double *const zs = sm + 2 * tsz;

    // Padding columns [ncols, wl) of both buffers are never loaded: zero.
    for (int e = tid; e < (a.wl - a.ncols) * TS; e += kD) {
        const int c = a.ncols + e / TS, r = e % TS;
        sm[dpos(c, r)] = 0.0;
        sm[tsz + dpos(c, r)] = 0.0;
    }

    // Step-1 item of this wave: Z tile (zm, zn), k-steps [klo, khi).  The A
    // operand of k-step k (row 16 zm + li of column 4 k + lk) sits at
    // zab + 4 LDc (k - klo).
    const int item = wave % ZT, ksp = wave / ZT;
    const int zm = item >> 1, zn = item & 1;
    const int nks = a.k1 > 0 ? ((a.k1 + 3) & ~3) >> 2 : 0;
    const int klo = nks * ksp / KS, khi = nks * (ksp + 1) / KS;
    const bool zuse = nks > 0 && ztile_used(a, zn);
    const int zab = (4 * klo + lk) * LDc + 16 * (zm ^ (lk & 1)) + li;
    double cf[kDKq];
#pragma unroll
    for (int j = 0; j < kDKq; j++)
        cf[j] = zuse && klo + j < khi ? a.C[(4 * (klo + j) + lk) * N1 + zn * 16 + li] : 0.0;

    // The group whose Z columns are this wave's Z tile (zoff = 16 zn); the
    // ksp = 0 wave subtracts the summed tile from the group's chunk zm and
    // solves it.  Its triangular-solve operands are held for the pass: B
    // fragments (lane: row lk, column li) of the inverted diagonal blocks and
    // of -T(4b + lk, li) right of block b.
    int gs = -1;
    for (int g = 0; g < a.ngrp; g++)
        if (ORTH_G(a, g, zoff) == 16 * zn) gs = g;
    const bool solver = ksp == 0 && gs >= 0;
    const double *Tq = gs >= 0 ? ORTH_G(a, gs, T) : nullptr;
    const int gt0 = gs >= 0 ? ORTH_G(a, gs, t0) : 0, gw = gs >= 0 ? ORTH_G(a, gs, w) : 0;
    double bw[4];
#pragma unroll
    for (int b = 0; b < 4; b++) bw[b] = Tq ? Tq[kSlotW + 128 * b + lk * kLdT + li] : 0.0;
    // Make the compiler wait for these loads here, once: a use it could not
    // prove complete inside the row loop would get a wait there, on every
    // tile, which would also wait for the next tile's loads.
#pragma unroll
    for (int j = 0; j < kDKq; j++) asm volatile("" ::"v"(cf[j]));
#pragma unroll
    for (int b = 0; b < 4; b++) asm volatile("" ::"v"(bw[b]));

    // Projection: wave w owns M-tile w (the host keeps Ma <= NW; waves past
    // Ma have none) and all NB N-tiles.  The operands of k-step ks (row
    // 4 ks + lk of columns a0 + 16 w + li and b0 + 16 t + li) sit at per-lane
    // bases plus compile-time offsets; rows 0..15 and 16..31 of an odd column
    // are swapped, so each operand has one base for k-steps 0..3 and one for
    // 4..7.
    const int Ma = (a.na + 15) >> 4;
    const bool proj = a.na > 0 && wave < Ma;
    // Reduced-precision N-tiles; their accumulators hold rows lk 4 + v.
    bool lowt[NB];
#pragma unroll
    for (int t = 0; t < NB; t++) lowt[t] = LPP && a.b0 + 16 * t >= a.lpb;
    const bool anylow = LPP && lowt[NB - 1];
    dbl4 acc[NB];
#pragma unroll
    for (int t = 0; t < NB; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    const int ca = a.a0 + 16 * wave + li, cb = a.b0 + li;
    const int pab = ca * LDc + lk, pbb = cb * LDc + lk;
    const int sa = 16 * (ca & 1), sb = 16 * (cb & 1);

    const int64_t rb = (int64_t)blockIdx.x * a.rows_per;
    const int64_t re = min(a.m, rb + a.rows_per);
    // Loads: slot span i (64 slots) goes to wave i % NW.  Slot q of column c
    // holds row pair q ^ 8 (c & 1) (the half swap); the pad slot repeats
    // slot 15's row pair.  Mixed form: the DMA covers the columns below raw0.
    const int ndma = MIX ? min(a.ncols, a.raw0) : a.ncols;
    const int nspan = (ndma * kDSl + 63) / 64;
    const int64_t rlast = (a.m - 1) & ~(int64_t)1;  // last row-pair start
    // Spans of this wave: i = wave + NW j; those with j in [j0, j1) are
    // issued by one call.
    auto issue = [&](int64_t r0, double *buf, int j0, int j1) {
        const int ln = opaque(lane);
        for (int i = wave + NW * j0; i < nspan && i < wave + NW * j1; i += NW) {
            const int s = i * 64 + ln, c = s / kDSl;
            const int q = min(s - c * kDSl, TS / 2 - 1) ^ ((c & 1) << 3);
            if (c < ndma) {
                const int64_t row = min(r0 + 2 * q, rlast);
                const double *g = a.X + row + (int64_t)c * a.ldX;
                __builtin_amdgcn_global_load_lds((glb_void_t *)g, (lds_void_t *)(buf + i * 128),
                                                 16, 0, 0);
            }
        }
    };
    // Raw FP32 columns [ndma, ncols): thread e < 8 nraw moves rows
    // 4 (e % 8) .. 4 (e % 8) + 3 of column ndma + e / 8, with one 16-byte
    // load when the FP32 slab is 16-byte aligned with a leading dimension
    // that is a multiple of 4 (a quarter of the load instructions of 4-byte
    // loads: the address path is what limits the pass), else with four
    // 4-byte loads; rows past the matrix are loaded from the last row and
    // never used.
    typedef float flt4 __attribute__((ext_vector_type(4)));
    const int nraw = a.ncols - ndma;
    const bool rv4 = MIX && a.ldX32 % 4 == 0 && (uintptr_t)a.X32 % 16 == 0;
    flt4 rv = flt4{0.f, 0.f, 0.f, 0.f};
    auto raw_issue = [&](int64_t r0) {
        if constexpr (MIX) {
            const int e = opaque(tid), c = e >> 3;
            if (c < nraw) {
                const int64_t row = r0 + 4 * (e & 7);
                const float *p = a.X32 + (int64_t)(ndma + c) * a.ldX32;
                if (rv4 && row + 3 < a.m) {
                    rv = __builtin_nontemporal_load((const flt4 *)(p + row));
                } else {
#pragma unroll
                    for (int q = 0; q < 4; q++) rv[q] = __builtin_nontemporal_load(p + min(row + q, a.m - 1));
                }
            }
        }
    };
    if (rb < re) {
        issue(rb, sm, 0, 1 << 20);
        raw_issue(rb);
    }

    int it = 0;
    for (int64_t r0 = rb; r0 < re; r0 += TS, it++) {
        double *cur = sm + (it & 1) * tsz;
        double *nxt = sm + ((it + 1) & 1) * tsz;
        // Tile i has landed (every wave waits for its own loads; the barrier
        // makes all of them visible), and every wave is done with the other
        // buffer: stream tile i+1 into it.
        asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
        if constexpr (MIX) {
            // This tile's raw columns, widened (rows past the slab: zero),
            // before any load of the next tile is issued (a wait for the
            // registers then waits for nothing in flight).
            {
                const int e = opaque(tid), c = e >> 3, r4 = 4 * (e & 7);
                if (c < nraw) {
                    double *d = cur + dpos(ndma + c, r4);
#pragma unroll
                    for (int q = 0; q < 4; q++) d[q] = r0 + r4 + q < re ? (double)rv[q] : 0.0;
                }
            }
            // With pending groups their solve, after two barriers, is the
            // first reader of these columns; without, the projection reads
            // them next.
            if (nraw > 0 && a.ngrp == 0) lds_barrier();
        }
        // The next tile's loads: by the waves that solve a group chunk in
        // this tile, and by all when no group is pending or no projection
        // follows the solve, at once; the other waves issue theirs while the
        // groups are solved, when they have nothing else to do (issuing
        // blocks a wave while the address path is busy: measured at the
        // widest benchmark pass, 11.2 ms against 12.3 ms with every wave
        // issuing at once, 12.0 ms with the non-solving waves issuing all
        // spans, 11.8 ms with the solving waves issuing theirs after the
        // solve, 12.5 ms with one span issued after every third step-1
        // MFMA; without a projection the late loads had too little time).
        const bool more = r0 + TS < re;
        const int jl = a.ngrp > 0 && a.na > 0 && !solver ? 0 : 1 << 20;
        if (more) {
            issue(r0 + TS, nxt, 0, jl);
            raw_issue(r0 + TS);
        }
        const int tq = opaque(tid), lq = opaque(lane);
        const int liq = lq & 15, lkq = lq >> 4;
        if (r0 + TS > re) {
            // Rows past the slab: loaded from clamped addresses, set to zero.
            const int nv = (int)(re - r0);
            for (int e = tq; e < a.wl * (TS - nv); e += kD) {
                const int c = e / (TS - nv), r = nv + e % (TS - nv);
                cur[dpos(c, r)] = 0.0;
            }
            lds_barrier();
        }

        if (a.ngrp > 0) {
            // Step 1: partial Z tile over this wave's k-steps (fully
            // unrolled so the fragments stay in registers; operand reads
            // bounded to two k-steps ahead).
            dbl4 z = dbl4{0.0, 0.0, 0.0, 0.0};
            if (zuse) {
                const double *za = cur + opaque(zab);
#pragma unroll
                for (int j = 0; j < kDKq; j++) {
                    if (klo + j < khi)
                        z = __builtin_amdgcn_mfma_f64_16x16x4f64(za[4 * LDc * j], cf[j], z, 0, 0, 0);
                    if ((j & 1) == 1) __builtin_amdgcn_sched_barrier(0);
                }
            }
            if (ksp > 0) {
#pragma unroll
                for (int v = 0; v < 4; v++) zs[((ksp - 1) * ZT + item) * 256 + v * 64 + lq] = z[v];
            }
            lds_barrier();
            if (more) issue(r0 + TS, nxt, jl, 1 << 20);
            if (solver) {
                // Chunk zm of the group in MFMA accumulator layout (lane:
                // column t0 + liq, rows 16 zm + lkq + 4 v), minus the
                // fixed-order sum of the KS partial Z tiles (columns below
                // the group width), then the blocked triangular solve
                // X <- X T^{-1} on MFMA (as in k_pass16b): block b's columns
                // go through LDS once to become an A operand.
                double *xb = cur + (gt0 + liq) * LDc + 16 * (zm ^ (liq & 1)) + lkq;
                const double *ab = cur + (gt0 + lkq) * LDc + 16 * (zm ^ (lkq & 1)) + liq;
                dbl4 x;
#pragma unroll
                for (int v = 0; v < 4; v++) x[v] = xb[4 * v];
                if (zuse) {
#pragma unroll
                    for (int k = 1; k < KS; k++)
#pragma unroll
                        for (int v = 0; v < 4; v++) z[v] += zs[((k - 1) * ZT + item) * 256 + v * 64 + lq];
                    if (liq < gw) {
#pragma unroll
                        for (int v = 0; v < 4; v++) x[v] -= z[v];
                    }
                }
                if (Tq) {
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        // Columns at or past the group width keep their tile
                        // contents (they only feed each other).
                        if ((liq >> 2) == b && liq < gw) {
#pragma unroll
                            for (int v = 0; v < 4; v++) xb[4 * v] = x[v];
                        }
                        // Other lanes' LDS writes become visible to this
                        // lane (the wave's LDS accesses execute in order);
                        // the fence stops the compiler from forwarding the
                        // lane's own earlier value instead of reloading.
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                        const double av = ab[4 * b * LDc];
                        if ((liq >> 2) == b) x = dbl4{0.0, 0.0, 0.0, 0.0};
                        x = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bw[b], x, 0, 0, 0);
                    }
                }
                if (liq < gw) {
#pragma unroll
                    for (int v = 0; v < 4; v++) xb[4 * v] = x[v];
                }
            }
            lds_barrier();
            // The updated groups back to X: thread -> (group, column, row).
            for (int e = tq; e < 2 * 16 * TS; e += kD) {
                const int g = e / (16 * TS), cc = (e / TS) % 16, r = e % TS;
                if (g < a.ngrp && cc < ORTH_G(a, g, w) && r0 + r < re) {
                    const int col = ORTH_G(a, g, t0) + cc;
                    __builtin_nontemporal_store(cur[dpos(col, r)],
                                                a.X + r0 + r + (int64_t)col * a.ldX);
                }
            }
        }

        // Step 2: projection partial over this tile's rows (the NB products
        // of a k-step share one A operand read).
        if (proj && !(anylow && lowt[0])) {
            const double *pa0 = cur + opaque(pab + sa), *pa1 = cur + opaque(pab - sa);
            const double *pb0 = cur + opaque(pbb + sb), *pb1 = cur + opaque(pbb - sb);
#pragma unroll
            for (int ks = 0; ks < TS / 4; ks++) {
                const double *pa = ks < 4 ? pa0 : pa1;
                const double *pb = ks < 4 ? pb0 : pb1;
                const double av = pa[4 * ks];
#pragma unroll
                for (int t = 0; t < NB; t++)
                    if (!lowt[t])
                        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, pb[4 * ks + 16 * LDc * t],
                                                                      acc[t], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if constexpr (LPP) {
            if (proj && anylow) {
                // Operands: rows 8 lk .. 8 lk + 7 of column a0 + 16 w + li
                // (A) and b0 + 16 t + li (B), eight consecutive doubles.
                const LpFrag<MXL> fa =
                    lp_frag<MXL>(cur + opaque(pab - lk + 8 * lk + (lk < 2 ? sa : -sa)), 1);
#pragma unroll
                for (int t = 0; t < NB; t++)
                    if (lowt[t]) {
                        const LpFrag<MXL> fb = lp_frag<MXL>(
                            cur + opaque(pbb - lk + 8 * lk + (lk < 2 ? sb : -sb)) + 16 * LDc * t, 1);
                        const f32x4 tt = lp_mma<MXL>(fa, fb, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                        for (int v = 0; v < 4; v++) acc[t][v] += (double)tt[v];
                    }
            }
        }
    }

    if (proj) {
        double *out = a.part + (int64_t)blockIdx.x * a.na * a.nb;
#pragma unroll
        for (int t = 0; t < NB; t++) {
            const int col = t * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = wave * 16 + (lowt[t] ? 4 * lk + v : lk + 4 * v);
                if (row < a.na && col < a.nb) out[row + (int64_t)col * a.na] = acc[t][v];
            }
        }
    }
    pass_tail(a, sm, kst, 64 * NW);
}

}  // namespace orth
