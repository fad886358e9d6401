/*
 * CPU oracle for the reorthogonalized block classical Gram-Schmidt path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, by the smoke check and by the
 * benchmark's CPU-baseline leg, always as the checker.  The product library
 * never links or calls it.
 *
 * A plain-C restatement of SRC/orth.c (five variants) and SRC/util.c:36-75
 * (accuracy) of the reference, in FP64 with 64-bit sizes.  Every BLAS/LAPACK
 * call of the reference is written out as a loop nest below:
 *   dgemm_('T','N') row reductions  -> oracle_gram   (fixed slab order)
 *   dgemm_('N','N') + dtrsm_('R','U','N','N') -> oracle_update
 *   dpotrf_('U')                    -> oracle_potrf
 *   dtrsm_('L','U','T','N')         -> trsm_lut
 *   Trilinos TSQR factorExplicit    -> oracle_houseqr (Householder, np=1)
 *
 * Deviations from the reference, all deliberate:
 *  - sizes are 64-bit (the reference's int products overflow at m*n >= 2^31);
 *  - the Gram of the block after next is formed with that block's own width
 *    (SRC/orth.c:129-131 uses s and reads past X when n < 3s);
 *  - diagonal blocks of R are formed as triangular products with exact zeros
 *    below the diagonal (SRC/orth.c:146-147 multiplies full squares whose
 *    strict-lower parts can hold stale workspace);
 *  - the intra-block QR returns R with a non-negative diagonal (Trilinos TSQR
 *    does not force signs; Q and R differ from it only by D = diag(+-1));
 *  - Cholesky breakdown is reported through info[] instead of ignored; the
 *    non-positive pivot still yields NaN/Inf downstream as LAPACK would leave
 *    garbage.
 * Row sums are split into NSLAB fixed slabs summed in slab order, so results
 * are bitwise independent of the thread count.
 */
#include "bcgs_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t idx;

#define NSLAB 64
#define CH 256
#define MINI(a, b) ((a) < (b) ? (a) : (b))
#define MAXI(a, b) ((a) > (b) ? (a) : (b))

static int g_threads = 1;

/* Mistake planted on purpose by the tests (0 = none), to show that their
 * checks reject a wrong restatement:
 *   1 drop the Y*Sigma term of R(0:c, b)
 *   2 add Y^T Y in the Pythagorean Gram of the second pass
 *   3 skip Omega^{-T} in the lag correction of the next block's projection
 *   4 leave the last prefix column out of the second-pass update
 *   5 multiply R_bb = Sigma Omega instead of Omega Sigma */
static int g_plant = 0;
void oracle_plant(int mistake) { g_plant = mistake; }

/* Relative perturbation of every first-pass projection S (0 = none).  The
 * second pass must remove it; with it the reorthogonalization terms are no
 * longer rounding-level, so the tests can pin them.  It also models a
 * first pass computed in reduced precision. */
static double g_noise = 0.0;
static int g_noise_abs = 0;
void oracle_set_first_pass_noise(double eta) { g_noise = eta; }
/* abs != 0: the perturbation of entry (i, j) is eta * ||S(:, j)||_2 * r
 * (the error model of a dot product in reduced precision) instead of
 * eta * |S(i, j)| * r. */
void oracle_set_noise_mode(int abs_mode) { g_noise_abs = abs_mode; }

static void perturb(idx rows, idx cols, double *S, idx ld, idx blk)
{
    if (g_noise == 0.0) return;
    for (idx j = 0; j < cols; j++) {
        double cn = 0.0;
        for (idx i = 0; i < rows; i++) cn += S[i + j * ld] * S[i + j * ld];
        cn = sqrt(cn);
        for (idx i = 0; i < rows; i++) {
            uint64_t z = (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull ^
                         (uint64_t)(j + 7) * 0xBF58476D1CE4E5B9ull ^
                         (uint64_t)(blk + 13) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            z *= 0xD6E8FEB86648D5A5ull;
            z ^= z >> 29;
            const double r = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
            if (g_noise_abs)
                S[i + j * ld] += g_noise * cn * r;
            else
                S[i + j * ld] *= 1.0 + g_noise * r;
        }
    }
}

void oracle_set_threads(int nthreads) { g_threads = nthreads < 1 ? 1 : nthreads; }
int oracle_get_threads(void) { return g_threads; }

static void slab(idx m, int k, idx *r0, idx *r1)
{
    *r0 = m * k / NSLAB;
    *r1 = m * (k + 1) / NSLAB;
}

/* ------------------------------------------------------------------ */
/* Row-parallel primitives                                             */
/* ------------------------------------------------------------------ */

/* P(na x nb, ld na) += A(r0:r1, 0:na)^T B(r0:r1, 0:nb). */
static void gram_rows(idx r0, idx r1, const double *A, idx lda, idx na,
                      const double *B, idx ldb, idx nb, double *P)
{
    for (idx c0 = r0; c0 < r1; c0 += CH) {
        idx len = MINI((idx)CH, r1 - c0);
        for (idx j = 0; j < nb; j++) {
            const double *b = B + j * ldb + c0;
            idx i = 0;
            for (; i + 4 <= na; i += 4) {
                const double *a0 = A + i * lda + c0;
                const double *a1 = a0 + lda, *a2 = a1 + lda, *a3 = a2 + lda;
                double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma omp simd reduction(+ : s0, s1, s2, s3)
                for (idx r = 0; r < len; r++) {
                    double bv = b[r];
                    s0 += a0[r] * bv;
                    s1 += a1[r] * bv;
                    s2 += a2[r] * bv;
                    s3 += a3[r] * bv;
                }
                P[i + j * na] += s0;
                P[i + 1 + j * na] += s1;
                P[i + 2 + j * na] += s2;
                P[i + 3 + j * na] += s3;
            }
            for (; i < na; i++) {
                const double *a = A + i * lda + c0;
                double s0 = 0;
#pragma omp simd reduction(+ : s0)
                for (idx r = 0; r < len; r++) s0 += a[r] * b[r];
                P[i + j * na] += s0;
            }
        }
    }
}

/* C(na x nb, ld ldc) = A^T B over m rows (the dgemm_('T','N') + Allreduce
 * pairs of SRC/orth.c, e.g. :111-114, :182-188). */
void oracle_gram(idx m, idx na, const double *A, idx lda, idx nb,
                 const double *B, idx ldb, double *C, idx ldc)
{
    if (na <= 0 || nb <= 0) return;
    double *part = (double *)calloc((size_t)(NSLAB * na * nb), sizeof(double));
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (int k = 0; k < NSLAB; k++) {
        idx r0, r1;
        slab(m, k, &r0, &r1);
        gram_rows(r0, r1, A, lda, na, B, ldb, nb, part + (idx)k * na * nb);
    }
    for (idx j = 0; j < nb; j++)
        for (idx i = 0; i < na; i++) {
            double s0 = 0;
            for (int k = 0; k < NSLAB; k++) s0 += part[(idx)k * na * nb + i + j * na];
            C[i + j * ldc] = s0;
        }
    free(part);
}

/* X(:, 0:w) = (X(:, 0:w) - Q(:, 0:c) W) T^{-1}; W is c x w (ld ldW), T is
 * w x w upper (ld ldT) or NULL for no normalization.  This is the
 * dgemm_('N','N',..,-1,..,1) + dtrsm_('R','U','N','N') pair of SRC/orth.c
 * (e.g. :121-124, :173-176, :206-209); the solve multiplies by the
 * reciprocal pivot as the reference BLAS dtrsm does. */
void oracle_update(idx m, idx w, double *X, idx ldX, idx c, const double *Q,
                   idx ldQ, const double *W, idx ldW, const double *T,
                   idx ldT)
{
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (int k = 0; k < NSLAB; k++) {
        idx r0, r1;
        slab(m, k, &r0, &r1);
        for (idx c0 = r0; c0 < r1; c0 += 2 * CH) {
            idx len = MINI((idx)(2 * CH), r1 - c0);
            for (idx j = 0; j < w; j++) {
                double *x = X + j * ldX + c0;
                for (idx i = 0; i < c; i++) {
                    const double wij = W[i + j * ldW];
                    const double *q = Q + i * ldQ + c0;
#pragma omp simd
                    for (idx r = 0; r < len; r++) x[r] -= wij * q[r];
                }
            }
            if (T) {
                for (idx j = 0; j < w; j++) {
                    double *x = X + j * ldX + c0;
                    for (idx l = 0; l < j; l++) {
                        const double t = T[l + j * ldT];
                        const double *xl = X + l * ldX + c0;
#pragma omp simd
                        for (idx r = 0; r < len; r++) x[r] -= t * xl[r];
                    }
                    const double rinv = 1.0 / T[j + j * ldT];
#pragma omp simd
                    for (idx r = 0; r < len; r++) x[r] *= rinv;
                }
            }
        }
    }
}

/* d[k] = sum_{r=lo}^{m-1} v[r] * A[r + k*lda], k < nk (fixed slab order). */
static void coldots(idx m, idx lo, const double *v, const double *A, idx lda,
                    idx nk, double *d)
{
    if (nk <= 0) return;
    double *part = (double *)calloc((size_t)(NSLAB * nk), sizeof(double));
#pragma omp parallel for num_threads(g_threads) schedule(static) if (m - lo > 32768)
    for (int k = 0; k < NSLAB; k++) {
        idx r0, r1;
        slab(m, k, &r0, &r1);
        r0 = MAXI(r0, lo);
        for (idx kk = 0; kk < nk; kk++) {
            const double *a = A + kk * lda;
            double s0 = 0;
            for (idx r = r0; r < r1; r++) s0 += v[r] * a[r];
            part[k * nk + kk] = s0;
        }
    }
    for (idx kk = 0; kk < nk; kk++) {
        double s0 = 0;
        for (int k = 0; k < NSLAB; k++) s0 += part[k * nk + kk];
        d[kk] = s0;
    }
    free(part);
}

/* A(lo:m, k) -= f[k] * v(lo:m) for k < nk. */
static void colaxpy(idx m, idx lo, const double *v, double *A, idx lda,
                    idx nk, const double *f)
{
#pragma omp parallel for num_threads(g_threads) schedule(static) if (m - lo > 32768)
    for (int k = 0; k < NSLAB; k++) {
        idx r0, r1;
        slab(m, k, &r0, &r1);
        r0 = MAXI(r0, lo);
        for (idx kk = 0; kk < nk; kk++) {
            double *a = A + kk * lda;
            const double fk = f[kk];
            for (idx r = r0; r < r1; r++) a[r] -= fk * v[r];
        }
    }
}

/* Householder QR with explicit Q (LAPACK dgeqr2 + dorg2r), replacing the
 * Trilinos TSQR factorExplicit calls of SRC/orth.c:105-107 (and :269-271,
 * :411-413, :503-505, :520-523, :532-535, :599-601, :612-614, :649-652). */
void oracle_houseqr(idx m, idx w, double *A, idx ldA, double *R, idx ldR)
{
    double *tau = (double *)calloc((size_t)w, sizeof(double));
    double *d = (double *)calloc((size_t)(w + 1), sizeof(double));
    for (idx j = 0; j < w; j++) {
        double *aj = A + j * ldA;
        const double alpha = aj[j];
        double xn2;
        coldots(m, j + 1, aj, aj, ldA, 1, &xn2);
        if (xn2 == 0.0) {
            tau[j] = 0.0;
        } else {
            const double beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
            tau[j] = (beta - alpha) / beta;
            const double scal = 1.0 / (alpha - beta);
            for (idx r = j + 1; r < m; r++) aj[r] *= scal;
            aj[j] = beta;
        }
        if (tau[j] != 0.0 && j + 1 < w) {
            idx nk = w - j - 1;
            double *rest = A + (j + 1) * ldA;
            coldots(m, j + 1, aj, rest, ldA, nk, d);
            for (idx k = 0; k < nk; k++) {
                d[k] = tau[j] * (rest[j + k * ldA] + d[k]);
                rest[j + k * ldA] -= d[k];
            }
            colaxpy(m, j + 1, aj, rest, ldA, nk, d);
        }
    }
    for (idx j = 0; j < w; j++)
        for (idx i = 0; i < w; i++) R[i + j * ldR] = (i <= j) ? A[i + j * ldA] : 0.0;
    /* dorg2r: accumulate the reflectors backwards into the explicit Q. */
    for (idx j = w - 1; j >= 0; j--) {
        double *aj = A + j * ldA;
        if (j + 1 < w) {
            idx nk = w - j - 1;
            double *rest = A + (j + 1) * ldA;
            aj[j] = 1.0;
            coldots(m, j + 1, aj, rest, ldA, nk, d);
            for (idx k = 0; k < nk; k++) {
                d[k] = tau[j] * (rest[j + k * ldA] + d[k]);
                rest[j + k * ldA] -= d[k];
            }
            colaxpy(m, j + 1, aj, rest, ldA, nk, d);
        }
        for (idx r = j + 1; r < m; r++) aj[r] *= -tau[j];
        aj[j] = 1.0 - tau[j];
        for (idx l = 0; l < j; l++) aj[l] = 0.0;
    }
    /* Non-negative diagonal of R. */
    for (idx j = 0; j < w; j++) {
        if (R[j + j * ldR] < 0.0) {
            for (idx k = j; k < w; k++) R[j + k * ldR] = -R[j + k * ldR];
            double *aj = A + j * ldA;
            for (idx r = 0; r < m; r++) aj[r] = -aj[r];
        }
    }
    free(tau);
    free(d);
}

/* ------------------------------------------------------------------ */
/* Small dense kernels (the s x s / c x s work of SRC/orth.c)          */
/* ------------------------------------------------------------------ */

/* dpotrf_('U') (SRC/orth.c:119 etc.): A = U^T U in place, strict lower
 * zeroed.  Returns 0 or the 1-based index of the first non-positive pivot. */
int oracle_potrf(idx n, double *A, idx lda)
{
    int info = 0;
    for (idx j = 0; j < n; j++) {
        double dj = A[j + j * lda];
        for (idx k = 0; k < j; k++) dj -= A[k + j * lda] * A[k + j * lda];
        if (!(dj > 0.0) && info == 0) info = (int)(j + 1);
        const double ujj = sqrt(dj);
        A[j + j * lda] = ujj;
        const double rinv = 1.0 / ujj;
        for (idx i = j + 1; i < n; i++) {
            double v = A[j + i * lda];
            for (idx k = 0; k < j; k++) v -= A[k + j * lda] * A[k + i * lda];
            A[j + i * lda] = v * rinv;
        }
    }
    for (idx j = 0; j < n; j++)
        for (idx i = j + 1; i < n; i++) A[i + j * lda] = 0.0;
    return info;
}

/* C(ma x nb) = beta*C + alpha*A^T B, inner dimension k. */
static void tn_small(idx k, idx ma, const double *A, idx lda, idx nb,
                     const double *B, idx ldb, double *C, idx ldc,
                     double alpha, double beta)
{
    for (idx j = 0; j < nb; j++)
        for (idx i = 0; i < ma; i++) {
            double s0 = 0;
            for (idx l = 0; l < k; l++) s0 += A[l + i * lda] * B[l + j * ldb];
            C[i + j * ldc] = beta * C[i + j * ldc] + alpha * s0;
        }
}

/* C(ma x nb) = beta*C + alpha*A B, inner dimension k. */
static void nn_small(idx ma, idx k, const double *A, idx lda, idx nb,
                     const double *B, idx ldb, double *C, idx ldc,
                     double alpha, double beta)
{
    for (idx j = 0; j < nb; j++)
        for (idx i = 0; i < ma; i++) {
            double s0 = 0;
            for (idx l = 0; l < k; l++) s0 += A[i + l * lda] * B[l + j * ldb];
            C[i + j * ldc] = beta * C[i + j * ldc] + alpha * s0;
        }
}

/* B(w x k) = U^{-T} B, U upper w x w: dtrsm_('L','U','T','N')
 * (SRC/orth.c:158-159). */
static void trsm_lut(idx w, idx k, const double *U, idx ldU, double *B,
                     idx ldB)
{
    for (idx j = 0; j < k; j++)
        for (idx i = 0; i < w; i++) {
            double t = B[i + j * ldB];
            for (idx l = 0; l < i; l++) t -= U[l + i * ldU] * B[l + j * ldB];
            B[i + j * ldB] = t / U[i + i * ldU];
        }
}

/* C = A B for upper-triangular w x w A, B; C written with zero strict lower. */
static void trmm_uu(idx w, const double *A, idx lda, const double *B, idx ldb,
                    double *C, idx ldc)
{
    for (idx j = 0; j < w; j++)
        for (idx i = 0; i < w; i++) {
            double s0 = 0;
            for (idx l = i; l <= j; l++) s0 += A[i + l * lda] * B[l + j * ldb];
            C[i + j * ldc] = (i <= j) ? s0 : 0.0;
        }
}

static void copy_block(idx rows, idx cols, const double *A, idx lda, double *B,
                       idx ldb)
{
    for (idx j = 0; j < cols; j++)
        for (idx i = 0; i < rows; i++) B[i + j * ldb] = A[i + j * lda];
}

/* ------------------------------------------------------------------ */
/* Variants                                                            */
/* ------------------------------------------------------------------ */

typedef struct {
    idx m, n, s, N;
    double *X;
    idx ldX;
    double *R;
    idx ldR;
    int *info;
} ctx;

static idx bw(const ctx *c, idx b)
{
    return b == 0 ? c->s : MINI(c->s, c->n - b * c->s);
}
static double *colp(const ctx *c, idx j) { return c->X + j * c->ldX; }
static double *Rp(const ctx *c, idx i, idx j) { return c->R + i + j * c->ldR; }
static void note(const ctx *c, idx b, int inf)
{
    if (inf && c->info[0] == 0) {
        c->info[0] = (int)(b + 1);
        c->info[1] = inf;
    }
}

/* T = chol(G + shift I) for a w x w block (upper part of G used). */
static int chol_shift(idx w, const double *G, idx ldg, double shift, double *T)
{
    for (idx j = 0; j < w; j++)
        for (idx i = 0; i < w; i++)
            T[i + j * w] = (i <= j) ? G[i + j * ldg] + (i == j ? shift : 0.0) : 0.0;
    return oracle_potrf(w, T, w);
}

/* Optional shift of the Pythagorean Cholesky steps (not in the reference,
 * which factors G - S^T S as is): G - S^T S + c u tr(G) I, with u = 2^-53
 * and G the Gram before the subtraction.  c = 0 (default) is the reference. */
static double g_pshift = 0.0;
void oracle_set_pythagorean_shift(double c) { g_pshift = c; }

static double pyth_shift(idx w, const double *G, idx ldg)
{
    if (g_pshift == 0.0) return 0.0;
    double t = 0.0;
    for (idx i = 0; i < w; i++) t += G[i + i * ldg];
    return g_pshift * 1.1102230246251565e-16 * t;
}

/* Second pass shared by P1s, 1s and P2s after the sync on block b:
 * P = [Q_{0:c} X_b]^T [X_b X_next] (ld na).  Finalizes block b and, when a
 * next block exists, forms its lag-corrected projection S_n into R(0:c+wb,
 * b+1).  Returns Y via Yout (c x wb, ld c) and Omega (wb x wb). */
static void second_pass(const ctx *c, idx b, idx cc, idx wb, idx wn,
                        double *P, idx na, double *Om)
{
    /* Omega = chol(Q^_b^T Q^_b - Y^T Y) (:134-137, :200-204, :288-291,
     * :345-349, :621-624, :676-680). */
    double *G = P + cc; /* rows cc.., cols 0..wb */
    const double sho = pyth_shift(wb, G, na);
    tn_small(cc, wb, P, na, wb, P, na, G, na, g_plant == 2 ? 1.0 : -1.0, 1.0);
    note(c, b, chol_shift(wb, G, na, sho, Om));
    /* Q_b = (Q^_b - Q Y) Omega^{-1} (:139-142, :206-209, :293-296, :351-354,
     * :625-628, :681-684). */
    oracle_update(c->m, wb, colp(c, cc), c->ldX, g_plant == 4 ? cc - 1 : cc,
                  colp(c, 0), c->ldX, P, na, Om, wb);
    if (wn > 0) {
        /* S-correction (:155-159, :310-314, :641-645): rows cc..cc+wb of
         * [Q Q^_b]^T X_next become Omega^{-T}(Q^_b^T X_n - Y^T Q^T X_n). */
        double *Pn = P + wb * na; /* column wb of P */
        tn_small(cc, wb, P, na, wn, Pn, na, Pn + cc, na, -1.0, 1.0);
        if (g_plant != 3) trsm_lut(wb, wn, Om, wb, Pn + cc, na);
        perturb(cc + wb, wn, Pn, na, b + 1);
        copy_block(cc + wb, wn, Pn, na, Rp(c, 0, cc + wb), c->ldR);
    }
}

/* BCGSI+P-1S, SRC/orth.c:48-222. */
static void run_p1s(const ctx *c)
{
    const idx s = c->s, N = c->N, m = c->m;
    oracle_houseqr(m, s, colp(c, 0), c->ldX, Rp(c, 0, 0), c->ldR); /* :105-107 */
    if (N == 1) return;
    double *P = (double *)calloc((size_t)((c->n + 2 * s) * 2 * s), sizeof(double));
    double *Sig = (double *)calloc((size_t)(s * s), sizeof(double));
    double *Om = (double *)calloc((size_t)(s * s), sizeof(double));
    double *Tmp = (double *)calloc((size_t)(s * s), sizeof(double));
    idx w1 = bw(c, 1);
    {
        /* [Q1 X2]^T X2 (:111-114), Pythagorean Sigma (:115-119). */
        idx na = s + w1;
        oracle_gram(m, na, colp(c, 0), c->ldX, w1, colp(c, s), c->ldX, P, na);
        perturb(s, w1, P, na, 1);
        const double sh = pyth_shift(w1, P + s, na);
        tn_small(s, w1, P, na, w1, P, na, P + s, na, -1.0, 1.0);
        copy_block(s, w1, P, na, Rp(c, 0, s), c->ldR); /* :117 */
        note(c, 1, chol_shift(w1, P + s, na, sh, Sig));
        /* X2 = (X2 - Q1 S) Sigma^{-1} (:121-124). */
        oracle_update(m, w1, colp(c, s), c->ldX, s, colp(c, 0), c->ldX, P, na, Sig, w1);
    }
    idx cc = s;
    for (idx b = 1; b < N; b++) {
        const idx wb = bw(c, b), wn = (b + 1 < N) ? bw(c, b + 1) : 0;
        /* One sync: [Q Q^_b]^T [Q^_b X_n] and X_n^T X_n (:127-133, :182-197). */
        const idx na = cc + wb + wn;
        oracle_gram(m, na, colp(c, 0), c->ldX, wb + wn, colp(c, cc), c->ldX, P, na);
        second_pass(c, b, cc, wb, wn, P, na, Om);
        /* R(0:c, b) = S + Y Sigma; R_bb = Omega Sigma (:144-147, :211-215). */
        if (g_plant != 1)
            nn_small(cc, wb, P, na, wb, Sig, wb, Rp(c, 0, cc), c->ldR, 1.0, 1.0);
        if (g_plant == 5)
            trmm_uu(wb, Sig, wb, Om, wb, Rp(c, cc, cc), c->ldR);
        else
            trmm_uu(wb, Om, wb, Sig, wb, Rp(c, cc, cc), c->ldR);
        if (wn > 0) {
            /* Sigma_n = chol(X_n^T X_n - S^T S) (:160-171). */
            double *Sn = Rp(c, 0, cc + wb);
            double *H = P + wb * na + cc + wb;
            const double shn = pyth_shift(wn, H, na);
            tn_small(cc + wb, wn, Sn, c->ldR, wn, Sn, c->ldR, H, na, -1.0, 1.0);
            note(c, b + 1, chol_shift(wn, H, na, shn, Sig));
            /* X_n = (X_n - Q S) Sigma_n^{-1} (:173-176). */
            oracle_update(m, wn, colp(c, cc + wb), c->ldX, cc + wb, colp(c, 0),
                          c->ldX, Sn, c->ldR, Sig, wn);
        }
        cc += wb;
    }
    free(P);
    free(Sig);
    free(Om);
    free(Tmp);
}

/* BCGSI+A-1S, SRC/orth.c:224-367. */
static void run_1s(const ctx *c)
{
    const idx s = c->s, N = c->N, m = c->m;
    oracle_houseqr(m, s, colp(c, 0), c->ldX, Rp(c, 0, 0), c->ldR); /* :269-271 */
    if (N == 1) return;
    double *P = (double *)calloc((size_t)((c->n + 2 * s) * 2 * s), sizeof(double));
    double *Om = (double *)calloc((size_t)(s * s), sizeof(double));
    idx w1 = bw(c, 1);
    /* S = Q1^T X2 (:275-278); X2 -= Q1 S without normalization (:280-281). */
    oracle_gram(m, s, colp(c, 0), c->ldX, w1, colp(c, s), c->ldX, Rp(c, 0, s), c->ldR);
    perturb(s, w1, Rp(c, 0, s), c->ldR, 1);
    oracle_update(m, w1, colp(c, s), c->ldX, s, colp(c, 0), c->ldX, Rp(c, 0, s), c->ldR, NULL, 0);
    idx cc = s;
    for (idx b = 1; b < N; b++) {
        const idx wb = bw(c, b), wn = (b + 1 < N) ? bw(c, b + 1) : 0;
        /* [Q V_b]^T [V_b X_n] (:284-287, :330-342); no Gram of X_n. */
        const idx na = cc + wb;
        oracle_gram(m, na, colp(c, 0), c->ldX, wb + wn, colp(c, cc), c->ldX, P, na);
        second_pass(c, b, cc, wb, wn, P, na, Om);
        /* R(0:c, b) = S + Y; R_bb = Omega (:298-302, :356-360). */
        for (idx j = 0; j < wb; j++)
            for (idx i = 0; i < cc; i++) *Rp(c, i, cc + j) += P[i + j * na];
        copy_block(wb, wb, Om, wb, Rp(c, cc, cc), c->ldR);
        if (wn > 0) {
            /* V_n = X_n - Q S (:323-324). */
            double *Sn = Rp(c, 0, cc + wb);
            oracle_update(m, wn, colp(c, cc + wb), c->ldX, cc + wb, colp(c, 0),
                          c->ldX, Sn, c->ldR, NULL, 0);
        }
        cc += wb;
    }
    free(P);
    free(Om);
}

/* BCGS-PIPI+, SRC/orth.c:369-462. */
static void run_pipi(const ctx *c)
{
    const idx s = c->s, N = c->N, m = c->m;
    oracle_houseqr(m, s, colp(c, 0), c->ldX, Rp(c, 0, 0), c->ldR); /* :411-413 */
    double *P = (double *)calloc((size_t)((c->n + s) * s), sizeof(double));
    double *P2 = (double *)calloc((size_t)((c->n + s) * s), sizeof(double));
    double *Sig = (double *)calloc((size_t)(s * s), sizeof(double));
    double *Om = (double *)calloc((size_t)(s * s), sizeof(double));
    idx cc = s;
    for (idx b = 1; b < N; b++) {
        const idx wb = bw(c, b), na = cc + wb;
        /* [Q X]^T X (:421-424); Sigma = chol(X^T X - S^T S) (:425-429). */
        oracle_gram(m, na, colp(c, 0), c->ldX, wb, colp(c, cc), c->ldX, P, na);
        perturb(cc, wb, P, na, b);
        const double sh1 = pyth_shift(wb, P + cc, na);
        tn_small(cc, wb, P, na, wb, P, na, P + cc, na, -1.0, 1.0);
        note(c, b, chol_shift(wb, P + cc, na, sh1, Sig));
        /* X = (X - Q S) Sigma^{-1} (:431-434). */
        oracle_update(m, wb, colp(c, cc), c->ldX, cc, colp(c, 0), c->ldX, P, na, Sig, wb);
        /* [Q X]^T X (:435-438); Omega = chol(X^T X - Y^T Y) (:440-444). */
        oracle_gram(m, na, colp(c, 0), c->ldX, wb, colp(c, cc), c->ldX, P2, na);
        const double sh2 = pyth_shift(wb, P2 + cc, na);
        tn_small(cc, wb, P2, na, wb, P2, na, P2 + cc, na, -1.0, 1.0);
        note(c, b, chol_shift(wb, P2 + cc, na, sh2, Om));
        /* X = (X - Q Y) Omega^{-1} (:446-449). */
        oracle_update(m, wb, colp(c, cc), c->ldX, cc, colp(c, 0), c->ldX, P2, na, Om, wb);
        /* R(0:c, b) = S + Y Sigma; R_bb = Omega Sigma (:451-455). */
        copy_block(cc, wb, P, na, Rp(c, 0, cc), c->ldR);
        nn_small(cc, wb, P2, na, wb, Sig, wb, Rp(c, 0, cc), c->ldR, 1.0, 1.0);
        trmm_uu(wb, Om, wb, Sig, wb, Rp(c, cc, cc), c->ldR);
        cc += wb;
    }
    free(P);
    free(P2);
    free(Sig);
    free(Om);
}

/* BCGSI+ (BCGS2), SRC/orth.c:464-548. */
static void run_i2(const ctx *c)
{
    const idx s = c->s, N = c->N, m = c->m;
    oracle_houseqr(m, s, colp(c, 0), c->ldX, Rp(c, 0, 0), c->ldR); /* :503-505 */
    double *S = (double *)calloc((size_t)(c->n * s), sizeof(double));
    double *W = (double *)calloc((size_t)(c->n * s), sizeof(double));
    double *Sig = (double *)calloc((size_t)(s * s), sizeof(double));
    double *Om = (double *)calloc((size_t)(s * s), sizeof(double));
    idx cc = s;
    for (idx b = 1; b < N; b++) {
        const idx wb = bw(c, b);
        /* S = Q^T X (:512-515); X -= Q S (:518); TSQR (:520-523). */
        oracle_gram(m, cc, colp(c, 0), c->ldX, wb, colp(c, cc), c->ldX, S, cc);
        perturb(cc, wb, S, cc, b);
        oracle_update(m, wb, colp(c, cc), c->ldX, cc, colp(c, 0), c->ldX, S, cc, NULL, 0);
        oracle_houseqr(m, wb, colp(c, cc), c->ldX, Sig, wb);
        /* W = Q^T Q^ (:524-527); Q^ -= Q W (:530); TSQR (:532-535). */
        oracle_gram(m, cc, colp(c, 0), c->ldX, wb, colp(c, cc), c->ldX, W, cc);
        oracle_update(m, wb, colp(c, cc), c->ldX, cc, colp(c, 0), c->ldX, W, cc, NULL, 0);
        oracle_houseqr(m, wb, colp(c, cc), c->ldX, Om, wb);
        /* R(0:c, b) = S + W Sigma; R_bb = Omega Sigma (:537-541). */
        copy_block(cc, wb, S, cc, Rp(c, 0, cc), c->ldR);
        nn_small(cc, wb, W, cc, wb, Sig, wb, Rp(c, 0, cc), c->ldR, 1.0, 1.0);
        trmm_uu(wb, Om, wb, Sig, wb, Rp(c, cc, cc), c->ldR);
        cc += wb;
    }
    free(S);
    free(W);
    free(Sig);
    free(Om);
}

/* BCGSI+P-2S, SRC/orth.c:550-697. */
static void run_p2s(const ctx *c)
{
    const idx s = c->s, N = c->N, m = c->m;
    oracle_houseqr(m, s, colp(c, 0), c->ldX, Rp(c, 0, 0), c->ldR); /* :599-601 */
    if (N == 1) return;
    double *P = (double *)calloc((size_t)((c->n + 2 * s) * 2 * s), sizeof(double));
    double *Sig = (double *)calloc((size_t)(s * s), sizeof(double));
    double *Om = (double *)calloc((size_t)(s * s), sizeof(double));
    idx w1 = bw(c, 1);
    /* S = Q1^T X2 (:605-608), R12 = S (:609), X2 -= Q1 S (:610-611),
     * TSQR of X2 (:612-614). */
    oracle_gram(m, s, colp(c, 0), c->ldX, w1, colp(c, s), c->ldX, Rp(c, 0, s), c->ldR);
    perturb(s, w1, Rp(c, 0, s), c->ldR, 1);
    oracle_update(m, w1, colp(c, s), c->ldX, s, colp(c, 0), c->ldX, Rp(c, 0, s), c->ldR, NULL, 0);
    oracle_houseqr(m, w1, colp(c, s), c->ldX, Sig, w1);
    idx cc = s;
    for (idx b = 1; b < N; b++) {
        const idx wb = bw(c, b), wn = (b + 1 < N) ? bw(c, b + 1) : 0;
        /* [Q Q^_b]^T [Q^_b X_n] (:617-620, :658-675).  The X_n^T X_n Gram
         * reduced at :660-664 is never read and is not formed. */
        const idx na = cc + wb;
        oracle_gram(m, na, colp(c, 0), c->ldX, wb + wn, colp(c, cc), c->ldX, P, na);
        second_pass(c, b, cc, wb, wn, P, na, Om);
        /* R(0:c, b) = S + Y Sigma; R_bb = Omega Sigma (:630-633, :686-690). */
        nn_small(cc, wb, P, na, wb, Sig, wb, Rp(c, 0, cc), c->ldR, 1.0, 1.0);
        trmm_uu(wb, Om, wb, Sig, wb, Rp(c, cc, cc), c->ldR);
        if (wn > 0) {
            /* V = X_n - Q S (:647-648); TSQR (:649-652). */
            double *Sn = Rp(c, 0, cc + wb);
            oracle_update(m, wn, colp(c, cc + wb), c->ldX, cc + wb, colp(c, 0),
                          c->ldX, Sn, c->ldR, NULL, 0);
            oracle_houseqr(m, wn, colp(c, cc + wb), c->ldX, Sig, wn);
        }
        cc += wb;
    }
    free(P);
    free(Sig);
    free(Om);
}

int oracle_bcgs(int variant, idx m, idx n, idx s, double *X, idx ldX,
                double *R, idx ldR, int *info)
{
    int dummy[2] = {0, 0};
    if (!info) info = dummy;
    info[0] = info[1] = 0;
    if (s < 1 || n < s || m < n || ldX < m || ldR < n || !X || !R) return 2;
    ctx c;
    c.m = m;
    c.n = n;
    c.s = s;
    c.N = 1 + (n - s + s - 1) / s;
    c.X = X;
    c.ldX = ldX;
    c.R = R;
    c.ldR = ldR;
    c.info = info;
    switch (variant) {
    case ORC_P1S: run_p1s(&c); break;
    case ORC_1S: run_1s(&c); break;
    case ORC_PIPI: run_pipi(&c); break;
    case ORC_I2: run_i2(&c); break;
    case ORC_P2S: run_p2s(&c); break;
    default: return 2;
    }
    return 0;
}

/* Workspace rule of SRC/orth.c:82-86 (and :255, :397, :489, :584), in
 * 64-bit arithmetic.  work == NULL means the library manages scratch. */
static int dropin(int variant, int m, int n, int s, double *X, int ldX,
                  double *R, int ldR, double *work, int lwork)
{
    const idx need = 5 * (idx)n * s + 9 * (idx)s * s +
                     (idx)m * (variant == ORC_I2 ? s : n);
    if (work != NULL && (idx)lwork < need) {
        printf("%% Work space is too small. It should be larger than 6*n*s + 8*s*s.\n");
        return 1;
    }
    int info[2];
    return oracle_bcgs(variant, m, n, s, X, ldX, R, ldR, info);
}

int oracle_bcgsi2P1s(int m, int n, int s, double *X, int ldX, double *R,
                     int ldR, double *work, int lwork)
{
    return dropin(ORC_P1S, m, n, s, X, ldX, R, ldR, work, lwork);
}
int oracle_bcgsi21s(int m, int n, int s, double *X, int ldX, double *R,
                    int ldR, double *work, int lwork)
{
    return dropin(ORC_1S, m, n, s, X, ldX, R, ldR, work, lwork);
}
int oracle_bcgspipi2(int m, int n, int s, double *X, int ldX, double *R,
                     int ldR, double *work, int lwork)
{
    return dropin(ORC_PIPI, m, n, s, X, ldX, R, ldR, work, lwork);
}
int oracle_bcgsi2(int m, int n, int s, double *X, int ldX, double *R, int ldR,
                  double *work, int lwork)
{
    return dropin(ORC_I2, m, n, s, X, ldX, R, ldR, work, lwork);
}
int oracle_bcgsi2P2s(int m, int n, int s, double *X, int ldX, double *R,
                     int ldR, double *work, int lwork)
{
    return dropin(ORC_P2S, m, n, s, X, ldX, R, ldR, work, lwork);
}

/* ------------------------------------------------------------------ */
/* Accuracy harness, SRC/util.c:36-75                                  */
/* ------------------------------------------------------------------ */

void oracle_accuracy(idx m, idx n, const double *X, idx ldX, const double *Q,
                     idx ldQ, const double *R, idx ldR, double *out)
{
    /* ||Q^T Q - I||_F (:50-58). */
    double *G = (double *)calloc((size_t)(n * n), sizeof(double));
    oracle_gram(m, n, Q, ldQ, n, Q, ldQ, G, n);
    double loo = 0;
    for (idx j = 0; j < n; j++)
        for (idx i = 0; i < n; i++) {
            double v = G[i + j * n] - (i == j ? 1.0 : 0.0);
            loo += v * v;
        }
    free(G);
    /* ||Q R - X||_F (:61-71) and ||X||_F. */
    double part_r[NSLAB], part_x[NSLAB];
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (int k = 0; k < NSLAB; k++) {
        idx r0, r1;
        slab(m, k, &r0, &r1);
        double sr = 0, sx = 0;
        double *t = (double *)malloc(sizeof(double) * CH);
        for (idx c0 = r0; c0 < r1; c0 += CH) {
            idx len = MINI((idx)CH, r1 - c0);
            for (idx j = 0; j < n; j++) {
                const double *x = X + j * ldX + c0;
                for (idx r = 0; r < len; r++) t[r] = -x[r];
                for (idx i = 0; i < n; i++) {
                    const double rij = R[i + j * ldR];
                    if (rij == 0.0) continue;
                    const double *q = Q + i * ldQ + c0;
                    for (idx r = 0; r < len; r++) t[r] += q[r] * rij;
                }
                for (idx r = 0; r < len; r++) {
                    sr += t[r] * t[r];
                    sx += x[r] * x[r];
                }
            }
        }
        free(t);
        part_r[k] = sr;
        part_x[k] = sx;
    }
    double sr = 0, sx = 0;
    for (int k = 0; k < NSLAB; k++) {
        sr += part_r[k];
        sx += part_x[k];
    }
    out[0] = sqrt(loo);
    out[1] = sqrt(sr);
    out[2] = sqrt(sx);
}
