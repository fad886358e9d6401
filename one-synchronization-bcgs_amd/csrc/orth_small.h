// The small-matrix steps between passes (FP64, one workgroup of any size;
// nt, the first argument of each, is its thread count):
// Pythagorean Gram, Cholesky factors with their solve blocks, lag
// correction, R assembly and the next pass's coefficients.  Run by k_small
// on its own, or by the last workgroup of a fused pass after the in-kernel
// reduction.  Their scratch is an LDS region of kSmallLds doubles passed in
// (static in k_small, the tile region in a pass, which is free by then).
#pragma once

#include <hip/hip_runtime.h>

#include "orth_kernels.h"

namespace orth {

// LDS layout of the small steps: two 32 x 33 matrices, the inverted 4 x 4
// diagonal blocks, and two scalar slots (breakdown pivot, trace).
constexpr int kSmallMat = kMaxS * (kMaxS + 1);
constexpr int kSmallDsh = 2 * kSmallMat;
constexpr int kSmallInfo = kSmallDsh + kMaxS * 4;
constexpr int kSmallTr = kSmallInfo + 1;
constexpr int kSmallLds = kSmallTr + 7;

// C(ma x nb) = beta*C + alpha * A^T B, A: k x ma, B: k x nb.
static __device__ void d_tn(int nt, int k, int ma, const double *A, int64_t lda, int nb, const double *B,
                            int64_t ldb, double *C, int64_t ldc, double alpha, double beta)
{
    for (int e = threadIdx.x; e < ma * nb; e += nt) {
        const int i = e % ma, j = e / ma;
        double s = 0.0;
        // Unrolled so that the loads of several terms are in flight; the
        // summation order stays sequential.
#pragma unroll 8
        for (int l = 0; l < k; l++) s = fma(A[l + i * lda], B[l + j * ldb], s);
        C[i + j * ldc] = (beta == 0.0 ? 0.0 : beta * C[i + j * ldc]) + alpha * s;
    }
    __syncthreads();
}

// C(ma x nb) = beta*C + alpha * A B, A: ma x k, B: k x nb.
static __device__ void d_nn(int nt, int ma, int k, const double *A, int64_t lda, int nb, const double *B,
                            int64_t ldb, double *C, int64_t ldc, double alpha, double beta)
{
    for (int e = threadIdx.x; e < ma * nb; e += nt) {
        const int i = e % ma, j = e / ma;
        double s = 0.0;
#pragma unroll 8
        for (int l = 0; l < k; l++) s = fma(A[i + l * lda], B[l + j * ldb], s);
        C[i + j * ldc] = (beta == 0.0 ? 0.0 : beta * C[i + j * ldc]) + alpha * s;
    }
    __syncthreads();
}

static __device__ void d_copy(int nt, int rows, int cols, const double *A, int64_t lda, double *B,
                              int64_t ldb)
{
    for (int e = threadIdx.x; e < rows * cols; e += nt) {
        const int i = e % rows, j = e / rows;
        B[i + j * ldb] = A[i + j * lda];
    }
    __syncthreads();
}

// T (w x w, ld kLdT) = chol(upper(G) + shift*I), strict lower zeroed.
// Returns 0 or the 1-based first non-positive pivot.
static __device__ int d_chol(int nt, int w, const double *G, int64_t ldg, double shift, double *T,
                             double *lds)
{
    double *sh = lds;  // sh[i * (kMaxS + 1) + j] = upper factor entry (i, j)
    double *dsh = lds + kSmallDsh;
    int *info_s = (int *)(lds + kSmallInfo);
    constexpr int L = kMaxS + 1;
    for (int e = threadIdx.x; e < w * w; e += nt) {
        const int i = e % w, j = e / w;
        sh[i * L + j] = (i <= j) ? G[i + j * ldg] + (i == j ? shift : 0.0) : 0.0;
    }
    if (threadIdx.x == 0) *info_s = 0;
    __syncthreads();
    for (int j = 0; j < w; j++) {
        if (threadIdx.x == 0) {
            double d = sh[j * L + j];
            for (int k = 0; k < j; k++) d = fma(-sh[k * L + j], sh[k * L + j], d);
            if (!(d > 0.0) && *info_s == 0) *info_s = j + 1;
            sh[j * L + j] = sqrt(d);
        }
        __syncthreads();
        const double rinv = 1.0 / sh[j * L + j];
        for (int i = j + 1 + threadIdx.x; i < w; i += nt) {
            double v = sh[j * L + i];
            for (int k = 0; k < j; k++) v = fma(-sh[k * L + j], sh[k * L + i], v);
            sh[j * L + i] = v * rinv;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < w * w; e += nt) {
        const int i = e % w, j = e / w;
        T[i + j * kLdT] = (i <= j) ? sh[i * L + j] : 0.0;
    }
    // Reciprocal pivots after the factor (the dtrsm_ reference scales by
    // 1 / T(j,j)); zero past w.
    for (int i = threadIdx.x; i < kLdT; i += nt)
        T[kSlotRecip + i] = i < w ? 1.0 / sh[i * L + i] : 0.0;
    // Inverses of the 4 x 4 diagonal blocks, one row per thread:
    // D(i,i) = 1/u(i,i), D(i,j) = -(sum_{k=i}^{j-1} D(i,k) u(k,j)) / u(j,j).
    for (int e = threadIdx.x; e < kMaxS; e += nt) {
        const int b = e >> 2, i = e & 3, o = 4 * b;
        double d[4] = {0.0, 0.0, 0.0, 0.0};
        if (o + i < w) {
            d[i] = 1.0 / sh[(o + i) * L + o + i];
            for (int j = i + 1; j < 4 && o + j < w; j++) {
                double v = 0.0;
                for (int k = i; k < j; k++) v = fma(d[k], sh[(o + k) * L + o + j], v);
                d[j] = -v * (1.0 / sh[(o + j) * L + o + j]);
            }
        }
        for (int j = 0; j < 4; j++) dsh[e * 4 + j] = d[j];
    }
    __syncthreads();
    // Row blocks W_b of the blocked solve (see orth_kernels.h).
    for (int e = threadIdx.x; e < (kMaxS / 4) * 4 * kLdT; e += nt) {
        const int b = e / (4 * kLdT), k = (e / kLdT) & 3, c = e % kLdT, o = 4 * b;
        double v = 0.0;
        if (o + k < w && c < w) {
            if (c >= o && c < o + 4) {
                v = dsh[(o + k) * 4 + c - o];
            } else if (c >= o + 4) {
                for (int j = 0; j < 4 && o + j < w; j++)
                    v = fma(-dsh[(o + k) * 4 + j], sh[(o + j) * L + c], v);
            }
        }
        T[kSlotW + e] = v;
    }
    const int info = *info_s;
    __syncthreads();
    return info;
}

// Stage the upper factor U (w x w, ld kLdT) into LDS (ld kMaxS + 1).
static __device__ void d_stage_upper(int nt, int w, const double *U, double *sU)
{
    for (int e = threadIdx.x; e < w * w; e += nt) {
        const int i = e % w, j = e / w;
        sU[i + j * (kMaxS + 1)] = U[i + j * kLdT];
    }
}

// B (w x k) = U^{-T} B with U upper (ld kLdT): dtrsm_('L','U','T','N').  The
// substitution runs in LDS, one column of B per thread.
static __device__ void d_trsm_lut(int nt, int w, int k, const double *U, double *B, int64_t ldb,
                                  double *lds)
{
    constexpr int L = kMaxS + 1;
    double *sU = lds, *sB = lds + kSmallMat;
    d_stage_upper(nt, w, U, sU);
    for (int e = threadIdx.x; e < w * k; e += nt) {
        const int i = e % w, j = e / w;
        sB[i + j * L] = B[i + j * ldb];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += nt) {
        for (int i = 0; i < w; i++) {
            double t = sB[i + j * L];
            for (int l = 0; l < i; l++) t = fma(-sU[l + i * L], sB[l + j * L], t);
            sB[i + j * L] = t / sU[i + i * L];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < w * k; e += nt) {
        const int i = e % w, j = e / w;
        B[i + j * ldb] = sB[i + j * L];
    }
    __syncthreads();
}

// C = A B for upper-triangular A, B (ld kLdT); C strict lower zeroed.
static __device__ void d_trmm(int nt, int w, const double *A, const double *B, double *C, int64_t ldc)
{
    for (int e = threadIdx.x; e < w * w; e += nt) {
        const int i = e % w, j = e / w;
        double s = 0.0;
        for (int l = i; l <= j; l++) s = fma(A[i + l * kLdT], B[l + j * kLdT], s);
        C[i + j * ldc] = (i <= j) ? s : 0.0;
    }
    __syncthreads();
}

// Ui = U^{-1} for upper-triangular U (both ld kLdT), one column per thread,
// in LDS.
static __device__ void d_trinv(int nt, int w, const double *U, double *Ui, double *lds)
{
    constexpr int L = kMaxS + 1;
    double *sU = lds, *sI = lds + kSmallMat;
    d_stage_upper(nt, w, U, sU);
    __syncthreads();
    for (int j = threadIdx.x; j < w; j += nt) {
        for (int i = w - 1; i >= 0; i--) {
            if (i > j) {
                sI[i + j * L] = 0.0;
                continue;
            }
            double t = (i == j) ? 1.0 : 0.0;
            for (int l = i + 1; l <= j; l++) t = fma(-sU[i + l * L], sI[l + j * L], t);
            sI[i + j * L] = t / sU[i + i * L];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < w * w; e += nt) {
        const int i = e % w, j = e / w;
        Ui[i + j * kLdT] = sI[i + j * L];
    }
    __syncthreads();
}

// Coefficients of the next fused pass, row-major C[k][n] (ld n1) for
// k < roundup4(k1): the first Z group is A(:, 0:w0) and comes from G0
// (column-major, ld ld0, k0 rows, zero below); the optional second group
// starts at column S and is given by G1 (k1 rows, ld ld1).
static __device__ void d_emit(int nt, double *C, int n1, int S, int k1, int k0, int w0, const double *G0,
                              int64_t ld0, int w1, const double *G1, int64_t ld1)
{
    const int k1p = (k1 + 3) & ~3;
    for (int e = threadIdx.x; e < k1p * n1; e += nt) {
        const int k = e / n1, n = e - k * n1;
        double v = 0.0;
        if (n < w0) {
            if (k < k0) v = G0[k + n * ld0];
        } else if (n >= S && n < S + w1 && G1) {
            if (k < k1) v = G1[k + (n - S) * ld1];
        }
        C[e] = v;
    }
    __syncthreads();
}

static __device__ void note_info(int *status, int blk, int info)
{
    if (threadIdx.x == 0 && info != 0 && status[0] == 0) {
        status[0] = blk + 1;
        status[1] = info;
    }
}

// Sum of the diagonal of the w x w matrix G (ld ldg), broadcast to every
// thread through LDS.
static __device__ double d_trace(int nt, int w, const double *G, int64_t ldg, double *lds)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < w; i++) t += G[i + i * ldg];
        lds[kSmallTr] = t;
    }
    __syncthreads();
    const double t = lds[kSmallTr];
    __syncthreads();
    return t;
}

// Shift of a Pythagorean Cholesky step: pshift * u * tr(G) over the w x w
// Gram G (ld ldg) before S^T S is subtracted; block-uniform, 0 when off.
static __device__ __forceinline__ double d_pyth_shift(int nt, const SmallArgs &a, int w, const double *G, int64_t ldg,
                                      double *lds)
{
    if (a.pshift == 0.0) return 0.0;
    return a.pshift * 1.1102230246251565e-16 * d_trace(nt, w, G, ldg, lds);
}

// One small step (SmallArgs::op); lds: kSmallLds doubles of LDS.
static __device__ __forceinline__ void small_step(int nt, const SmallArgs &a, double *lds)
{
    double *P = a.P;
    const int c = a.c, wb = a.wb, wn = a.wn, na = a.na;
    double *Rb = a.R + (int64_t)c * a.ldR;  // R(:, block b)
    double *Rbb = Rb + c;                   // R(b, b)
    switch (a.op) {
    case OP_OPEN_PYTH: {
        // G -= S^T S; R(0:c, b) = S; Sigma = chol(G)  (SRC/orth.c:115-119).
        const double sh = d_pyth_shift(nt, a, wb, P + c, na, lds);
        d_tn(nt, c, wb, P, na, wb, P, na, P + c, na, -1.0, 1.0);
        d_copy(nt, c, wb, P, na, Rb, a.ldR);
        note_info(a.status, a.blk, d_chol(nt, wb, P + c, na, sh, a.Sig, lds));
        if (a.Cout) d_emit(nt, a.Cout, a.n1, a.S, c, c, wb, P, na, 0, nullptr, 0);
        break;
    }
    case OP_CLOSE_OPEN: {
        // Omega = chol(Q^T Q^ - Y^T Y) (SRC/orth.c:134-137, :200-204).
        const double sho = d_pyth_shift(nt, a, wb, P + c, na, lds);
        d_tn(nt, c, wb, P, na, wb, P, na, P + c, na, -1.0, 1.0);
        note_info(a.status, a.blk, d_chol(nt, wb, P + c, na, sho, a.Om, lds));
        if (a.flags & F_1S) {
            // R(0:c, b) = S + Y; R_bb = Omega (SRC/orth.c:356-360).
            for (int e = threadIdx.x; e < c * wb; e += nt) {
                const int i = e % c, j = e / c;
                Rb[i + j * a.ldR] += P[i + j * na];
            }
            __syncthreads();
            d_copy(nt, wb, wb, a.Om, kLdT, Rbb, a.ldR);
        } else {
            // R(0:c, b) = S + Y Sigma; R_bb = Omega Sigma (SRC/orth.c:211-215).
            d_nn(nt, c, wb, P, na, wb, a.Sig, kLdT, Rb, a.ldR, 1.0, 1.0);
            d_trmm(nt, wb, a.Om, a.Sig, Rbb, a.ldR);
        }
        if (wn > 0) {
            // Lag correction of the next block's projection (SRC/orth.c:155-159).
            double *Pn = P + (int64_t)wb * na;
            d_tn(nt, c, wb, P, na, wn, Pn, na, Pn + c, na, -1.0, 1.0);
            d_trsm_lut(nt, wb, wn, a.Om, Pn + c, na, lds);
            d_copy(nt, c + wb, wn, Pn, na, Rb + (int64_t)wb * a.ldR, a.ldR);
            if (a.flags & F_PYTH_NEXT) {
                // Sigma_n = chol(X_n^T X_n - S^T S) (SRC/orth.c:160-171).
                double *H = Pn + c + wb;
                const double shn = d_pyth_shift(nt, a, wn, H, na, lds);
                d_tn(nt, c + wb, wn, Pn, na, wn, Pn, na, H, na, -1.0, 1.0);
                note_info(a.status, a.blk + 1, d_chol(nt, wn, H, na, shn, a.Sig, lds));
            }
        }
        if (a.Cout && (a.flags & F_SEQ)) {
            // Sequential next pass: group 1 uses S_n on the finished Q_b.
            d_emit(nt, a.Cout, a.n1, a.S, c + wb, c, wb, P, na, wn, P + (int64_t)wb * na, na);
        } else if (a.Cout) {
            // Next pass: group 0 finalizes block b with [Y; 0] and Omega; group 1
            // is the first pass of block b+1 with the prefix [Q Q^_b] and
            // [S_top - Y Omega^{-1} S_bot; Omega^{-1} S_bot], which equals
            // Q S_top + Q_b S_bot without waiting for Q_b.
            double *G1 = nullptr;
            if (wn > 0) {
                double *Oi = a.Ctmp, *T2 = a.Tbig;  // T2: (c+wb) x wn, ld c+wb
                double *Pn = P + (int64_t)wb * na;
                d_trinv(nt, wb, a.Om, Oi, lds);
                // T2 rows c..c+wb-1 = Omega^{-1} S_bot.
                d_nn(nt, wb, wb, Oi, kLdT, wn, Pn + c, na, T2 + c, c + wb, 1.0, 0.0);
                // T2 rows 0..c-1 = S_top - Y (Omega^{-1} S_bot).
                d_copy(nt, c, wn, Pn, na, T2, c + wb);
                d_nn(nt, c, wb, P, na, wn, T2 + c, c + wb, T2, c + wb, -1.0, 1.0);
                G1 = T2;
            }
            d_emit(nt, a.Cout, a.n1, a.S, c + wb, c, wb, P, na, wn, G1, c + wb);
        }
        break;
    }
    case OP_PIPI1: {
        // Sigma = chol(X^T X - S^T S); R(0:c, b) = S (SRC/orth.c:425-429).
        const double sh = d_pyth_shift(nt, a, wb, P + c, na, lds);
        d_tn(nt, c, wb, P, na, wb, P, na, P + c, na, -1.0, 1.0);
        note_info(a.status, a.blk, d_chol(nt, wb, P + c, na, sh, a.Sig, lds));
        d_copy(nt, c, wb, P, na, Rb, a.ldR);
        if (a.Cout) d_emit(nt, a.Cout, a.n1, a.S, c, c, wb, P, na, 0, nullptr, 0);
        break;
    }
    case OP_PIPI2: {
        // Omega; R(0:c, b) += Y Sigma; R_bb = Omega Sigma (SRC/orth.c:440-455).
        const double sh = d_pyth_shift(nt, a, wb, P + c, na, lds);
        d_tn(nt, c, wb, P, na, wb, P, na, P + c, na, -1.0, 1.0);
        note_info(a.status, a.blk, d_chol(nt, wb, P + c, na, sh, a.Om, lds));
        d_nn(nt, c, wb, P, na, wb, a.Sig, kLdT, Rb, a.ldR, 1.0, 1.0);
        d_trmm(nt, wb, a.Om, a.Sig, Rbb, a.ldR);
        if (a.Cout) d_emit(nt, a.Cout, a.n1, a.S, c, c, wb, P, na, 0, nullptr, 0);
        break;
    }
    case OP_PLAIN_S: {
        d_copy(nt, c, wb, P, na, Rb, a.ldR);
        if (a.Cout) d_emit(nt, a.Cout, a.n1, a.S, c, c, wb, P, na, 0, nullptr, 0);
        break;
    }
    case OP_I2_W: {
        d_nn(nt, c, wb, P, na, wb, a.Sig, kLdT, Rb, a.ldR, 1.0, 1.0);
        if (a.Cout) d_emit(nt, a.Cout, a.n1, a.S, c, c, wb, P, na, 0, nullptr, 0);
        break;
    }
    case OP_RBB: {
        d_trmm(nt, wb, a.Om, a.Sig, Rbb, a.ldR);
        break;
    }
    case OP_CQR1: {
        // Shift of Fukaya et al. (2020): 11 (m w + w (w+1)) u ||X||_F^2.
        const double tr = d_trace(nt, wb, P, na, lds);
        const double u = 1.1102230246251565e-16;
        const double shift = 11.0 * ((double)a.m * wb + (double)wb * (wb + 1)) * u * tr;
        note_info(a.status, a.blk, d_chol(nt, wb, P, na, shift, a.Cq, lds));
        break;
    }
    case OP_CQR2: {
        note_info(a.status, a.blk, d_chol(nt, wb, P, na, 0.0, a.Cq + kSlot, lds));
        break;
    }
    case OP_CQR3: {
        double *C2 = a.Cq + 2 * kSlot;
        note_info(a.status, a.blk, d_chol(nt, wb, P, na, 0.0, C2, lds));
        d_trmm(nt, wb, C2, a.Cq + kSlot, a.Ctmp, kLdT);
        d_trmm(nt, wb, a.Ctmp, a.Cq, a.Rdst, a.ldRdst);
        break;
    }
    default: break;
    }
}

}  // namespace orth
