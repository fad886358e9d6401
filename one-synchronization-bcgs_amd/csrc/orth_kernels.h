// Device kernels of liborth_hip.so and their launch wrappers.
//
// Data layout: X is the caller's m x n column-major row slab (leading
// dimension ldX) and is updated in place, one block column at a time.  The
// small matrices (projections, Cholesky factors, R) live in library scratch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace orth {

constexpr int kMaxS = 32;  // widest block the kernels support
constexpr int kLdT = 32;   // leading dimension of every s x s factor
// Each factor slot holds the 32 x 32 upper factor T (ld kLdT), the
// reciprocals of its diagonal, and the row blocks of the blocked solve
// X T^{-1}: for block b (columns 4b..4b+3), W_b (4 x 32, row-major, at
// kSlotW + 128 b) holds D_b = inv(T(4b:4b+4, 4b:4b+4)) in columns 4b..4b+3,
// -D_b T(4b:4b+4, c) in the columns c right of the block, zero elsewhere and
// past the factor's width.  All are written by the Cholesky step.
constexpr int kSlotRecip = kLdT * kLdT;
constexpr int kSlotW = kSlotRecip + kLdT;
constexpr int kSlot = kSlotW + (kMaxS / 4) * 4 * kLdT;

// One row-local block update of a streaming pass:
//   X(:, col0:col0+w) <- (X(:, col0:col0+w) - X(:, 0:c) * W) * T^{-1}
// W is c x w (ld ldW; columns w..S-1 must be readable, their values unused),
// T is w x w upper with ld kLdT, or nullptr for no normalization.
struct Upd {
    int64_t col0;
    int w;
    int c;
    const double *W;
    int64_t ldW;
    const double *T;
};

// Small-matrix steps run by one workgroup between passes.
enum SmallOp {
    OP_OPEN_PYTH = 1,   // P1s prologue: Sigma = chol(G - S^T S), R(0:c,b) = S
    OP_CLOSE_OPEN = 2,  // finish block b (Omega, R) and open block b+1
    OP_PIPI1 = 3,       // PIPI first pass: Sigma, R(0:c,b) = S
    OP_PIPI2 = 4,       // PIPI second pass: Omega, R(0:c,b) += Y Sigma, R_bb
    OP_PLAIN_S = 5,     // R(0:c,b) = S (unnormalized first pass)
    OP_I2_W = 6,        // BCGSI+: R(0:c,b) += W Sigma
    OP_RBB = 7,         // R_bb = Omega Sigma
    OP_CQR1 = 8,        // shifted CholQR step 1
    OP_CQR2 = 9,        // CholQR step 2
    OP_CQR3 = 10        // CholQR step 3, writes R3 R2 R1
};
enum SmallFlag {
    F_PYTH_NEXT = 1,  // next block's Sigma from the Pythagorean Gram (P1s)
    F_1S = 2,         // R(0:c,b) = S + Y and R_bb = Omega (BCGSI+A-1S)
    F_SEQ = 4         // emit the next pass for sequential mode (no folding)
};

struct SmallArgs {
    int op;
    int flags;
    int blk;  // 0-based block index b
    int c;    // columns final before block b (= first column of block b)
    int wb;   // width of block b
    int wn;   // width of block b+1 (0 if none)
    int na;   // leading dimension of P
    int64_t m;
    double *P;
    double *R;
    int64_t ldR;
    double *Sig, *Om, *Cq, *Ctmp;  // factor slots (kSlot apart; Cq holds 3)
    double *Tbig;                  // (n + kMaxS) x kMaxS scratch
    double *Rdst;
    int64_t ldRdst;
    int *status;  // [0] = first breakdown block (1-based), [1] = pivot
    // Coefficients of the next fused pass (row-major, ld n1), written by
    // steps that open or close a block when Cout != nullptr.
    double *Cout;
    int n1;
    int S;  // block template width: column offset of the second Z group
    // Pythagorean Cholesky steps factor G - S^T S + pshift * u * tr(G) * I
    // (u = 2^-53, G the Gram before the subtraction); 0 = no shift.
    double pshift;
};

hipError_t launch_update(int S, double *X, int64_t ldX, int64_t m,
                         const Upd *u, int nupd, hipStream_t st);

// Fixed summation order of the projection partials, shared by every
// reduction (launch_reduce, launch_reduce_small and the in-pass tail): the
// nwg partials form groups of tail_group(nwg) consecutive workgroups; each
// group is summed in workgroup order starting from its first partial, and
// the group sums are added in group order starting from the first.  At most
// kTailMaxGroups groups.
constexpr int kTailMaxGroups = 32;
inline int tail_group(int nwg)
{
    int g = 1;
    while (g * g < nwg) g++;
    return g * kTailMaxGroups < nwg ? (nwg + kTailMaxGroups - 1) / kTailMaxGroups : g;
}

// The synchronization at the end of a fused pass (the reference's
// MPI_Allreduce and the small step after it, SRC/orth.c:114, :133, :188,
// :197), run inside the pass launch by tickets, without waiting: the last
// workgroup of each group to finish sums that group's partials into lvl, the
// last of those sums the group sums into out and runs the small step.  The
// counters are set back to zero by the workgroups that read them last.
struct TailArgs {
    int small;     // 1: after the sum, run `sm` (sm.P = out) in the last workgroup
    int nwg;       // workgroups of the pass
    int grp;       // tail_group(nwg)
    int pad;
    int *tickets;  // [0, kTailMaxGroups) group counters, [kTailTop] the top counter
    double *lvl;   // group sums: ceil(nwg / grp) x (na nb) doubles
    double *out;   // the sum, na x nb, ld na
    SmallArgs sm;
};
constexpr int kTailTop = kTailMaxGroups;

// Projection partials: part[x][i + j*na] = sum over the rows of workgroup x
// of A(:, i) * B(:, j), A = columns [a0, a0+na), B = columns [b0, b0+nb).
struct ProjPlan {
    int Ma, Nb;        // 16-wide tiles of A and B
    int ga, gb;        // tiles per group along A and B
    int ngy;           // number of tile groups (gridDim.y)
    int nwg;           // row chunks (gridDim.x)
    int64_t rows_per;  // rows per chunk (multiple of 16)
    size_t lds_bytes;
};
ProjPlan plan_proj(int64_t m, int64_t na, int64_t nb, int target_wg);
hipError_t launch_proj(const ProjPlan &pl, const double *X, int64_t ldX,
                       int64_t m, int64_t a0, int64_t na, int64_t b0,
                       int64_t nb, double *part, hipStream_t st);
// out[i + j*ldo] = sum_x part[x][i + j*na], fixed order.
hipError_t launch_reduce(const double *part, int nwg, int64_t na, int64_t nb,
                         double *out, int64_t ldo, hipStream_t st);

hipError_t launch_small(const SmallArgs &a, hipStream_t st);

// The synchronization of a pass in one launch (the reference's
// MPI_Allreduce, SRC/orth.c:114, :133, :188, :197, and the small step after
// it): workgroups sum slices of the partials in the fixed order of
// launch_reduce, then take a ticket; the last one to arrive runs the small
// step `sm` (sm.P = out) and sets the ticket counter back to zero.
// ticket: one int, zero on entry.
hipError_t launch_reduce_small(const double *part, int nwg, int64_t na, int64_t nb, double *out,
                               int *ticket, const SmallArgs &sm, hipStream_t st);

// ---------------------------------------------------------------------------
// Fused streaming pass: one sweep over the row slab that applies the
// row-local block updates and accumulates the next projection.
//
// Per 32-row tile staged in LDS (columns [0, ncols) of X):
//   step 1  Z = tile(:, 0:k1) * C on FP64 MFMA (C is k1 x n1, row-major);
//   groups  tile(:, t0:t0+w) -= Z(:, zoff:zoff+w), then a row-wise forward
//           substitution with the upper factor T (if any); the groups are
//           written back to X;
//   step 2  partial += tile(:, a0:a0+na)^T tile(:, b0:b0+nb) on FP64 MFMA.
// ---------------------------------------------------------------------------
struct PassGroup {
    int t0;          // first column of the target block
    int w;           // its width (<= S)
    int zoff;        // first column of Z subtracted from it
    int k;           // inner dimension of its product in sequential mode
    const double *T; // w x w upper factor (ld kLdT) or nullptr
    int lp;          // mixed form: its product runs in the reduced precision
};

struct PassArgs {
    double *X;
    int64_t ldX;
    int64_t m;
    // 1: the pass synchronizes itself (tail below).  Beside the fields every
    // workgroup reads at its start, so that the check at its end finds the
    // line cached.
    int tail_on;
    int ncols;  // columns of X staged per tile
    int wl;     // LDS tile width: multiple of 16 covering every range
    int k1;     // inner dimension of step 1 (0: no product)
    int n1;     // columns of Z (multiple of 16)
    const double *C;
    int ngrp;
    int seq;  // 1: group 1's product uses group 0's finished columns
    PassGroup g[2];
    int a0, na, b0, nb;  // step 2 ranges (na == 0: no projection)
    double *part;
    int64_t rows_per;
    // Mixed form only: columns >= raw0 (never written yet) are read from the
    // FP32 input X32 instead of X; projection tiles whose B columns start at
    // or after lpb run in the reduced precision.
    const float *X32;
    int64_t ldX32;
    int raw0;
    int lpb;
    TailArgs tail;
};

// LDS bytes a pass needs at least when its tail runs a small step (the
// small step's region and the ticket flag); the launchers raise their
// dynamic LDS to it.
size_t pass_tail_lds(const PassArgs &a, size_t lds);

constexpr int kPassRows = 32;
constexpr int kPassMaxCols = 288;

// LDS bytes the fused pass needs, or 0 when the shape does not fit it.
size_t pass_lds_bytes(const PassArgs &a, int F = 1);
hipError_t launch_pass(int S, const PassArgs &a, int nwg, hipStream_t st);
// Workgroups per CU the generic pass for this shape can keep resident (tail:
// the pass runs its synchronization, whose LDS it then holds).
int pass_wg_per_cu(int S, const PassArgs &a, bool tail);
// Rows per tile (the row-chunk unit) of the 256-thread passes for a.
int pass_rows(int S, const PassArgs &a);

// Block width 16 with 512 threads per workgroup: super-tile rows for the
// pass (multiple of 32, 0 when it does not apply); rows_per must be a
// multiple of it.
int pass16b_plan(const PassArgs &a, int *mt, int *nb);
size_t pass16b_lds_bytes(const PassArgs &a, int ts);
hipError_t launch_pass16b(const PassArgs &a, int nwg, hipStream_t st);

// Block width 16, FP64, double-buffered direct-to-LDS loads in 32-row tiles
// (rows_per must be a multiple of 32): waves per workgroup (8 or 16), or 0
// when it does not apply.
// any_width: also below the column bound (the mixed form with reduced
// precisions, whose 512-thread pass runs one workgroup per CU).  min_cols:
// the column bound (0: the library's default; the batched form's tuner
// picks it per batch).
int pass16d_plan(const PassArgs &a, int *mt, int *nb, bool any_width = false, int min_cols = 0);
// Launch with the plan's waves per workgroup (nw) and N-tiles (nb).
hipError_t launch_pass16d(const PassArgs &a, int nw, int nb, int nwg, hipStream_t st);
// Its mixed form (a.X32 set): mx = kMxFP64 (FP32 storage, FP64 products) or
// a reduced precision for the projection tiles from lpb on.
hipError_t launch_pass16d_mixed(const PassArgs &a, int mx, int nw, int nb, int nwg,
                                hipStream_t st);

// Mixed-precision s = 16 pass (32-row tiles, 512 threads): mx selects the
// reduced precision of the marked products (1 bf16, 2 fp16, 3 bf16x3 split,
// 4 none: FP64 arithmetic on FP32-stored input).  pass16m_lds_bytes returns
// 0 when the shape does not fit it.
enum { kMxBF16 = 1, kMxFP16 = 2, kMxBF16X3 = 3, kMxFP64 = 4 };
// Host flag beside a reduced precision: only the first-pass projection and
// Gram are reduced, the first-pass update stays FP64.
constexpr int kMxProjOnly = 8;
size_t pass16m_lds_bytes(const PassArgs &a);
hipError_t launch_pass16m(const PassArgs &a, int mx, int nwg, hipStream_t st);

// The mixed form on the 512-thread s = 16 pass (super-tiles, two workgroups
// per CU): 0 when the shape does not fit it (then launch_pass16m applies).
int pass16bm_plan(const PassArgs &a);
hipError_t launch_pass16bm(const PassArgs &a, int mx, int nwg, hipStream_t st);

// Streaming probe: mode 0 copies `bytes` from src to dst, mode 1 reads
// `bytes` of src (one partial sum per workgroup to part[nwg]); 16-byte
// non-temporal accesses, nwg workgroups of 256 threads.
hipError_t launch_stream_probe(int mode, double *dst, const double *src, int64_t bytes,
                               double *part, int nwg, hipStream_t st);

// Q = X32 widened exactly to FP64 (m x n, column-major).
hipError_t launch_widen(const float *x, int64_t ldx, double *q, int64_t ldq, int64_t m, int64_t n,
                        hipStream_t st);

// Zero n entries with a kernel (no memset nodes, so the sequence can be
// captured in a graph).
hipError_t launch_fill(double *p, size_t n, hipStream_t st);
hipError_t launch_fill_int(int *p, size_t n, hipStream_t st);

// Accuracy harness: Q^T Q with compensated (double-double) accumulation.
// part receives, per row chunk (gram_dd_chunks(m) of them) and 64 x 64
// block of the upper block triangle (nb (nb + 1) / 2, nb = ceil(n / 64)), 256
// threads x 16 pairs x (hi, lo): pair k = 4 di + dj of thread t is
// (64 I + 4 (t % 16) + di, 64 J + 4 (t / 16) + dj).
int gram_dd_chunks(int64_t m);
hipError_t launch_gram_dd(const double *Q, int64_t ldQ, int64_t m, int64_t n, double *part,
                          hipStream_t st);

// Residual harness: per-workgroup partial sums of ||Q R - X||^2 and ||X||^2
// into part2[2*x], part2[2*x+1]; returns the number of workgroups used.
int resid_wg_count(int64_t m);
hipError_t launch_resid(int64_t m, int64_t n, const double *X, int64_t ldX,
                        const double *Q, int64_t ldQ, const double *R,
                        int64_t ldR, double *part2, hipStream_t st);

}  // namespace orth
