/*
 * CPU oracle for the reorthogonalized block classical Gram-Schmidt path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark may load this library, and only as the
 * checker.  The product path (liborth_hip.so) never links or calls it.
 *
 * Plain-C restatement (FP64, column-major, 64-bit sizes) of the five
 * variants in the reference's SRC/orth.c and of the accuracy harness in
 * SRC/util.c:36-75.  Each function cites the reference lines it follows.
 * BLAS/LAPACK primitives are written out here (no external library), so
 * the oracle builds with gcc alone.
 */
#ifndef BCGS_ORACLE_H
#define BCGS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Variant ids (same order as the extended device API). */
enum {
    ORC_P1S = 0,   /* bcgsi2P1s  SRC/orth.c:48-222  */
    ORC_1S = 1,    /* bcgsi21s   SRC/orth.c:224-367 */
    ORC_PIPI = 2,  /* bcgspipi2  SRC/orth.c:369-462 */
    ORC_I2 = 3,    /* bcgsi2     SRC/orth.c:464-548 */
    ORC_P2S = 4    /* bcgsi2P2s  SRC/orth.c:550-697 */
};

/* Number of worker threads used by the row-parallel primitives (>=1).
 * Results do not depend on it: row sums use a fixed slab partition. */
void oracle_set_threads(int nthreads);

/* Optional shift of the Pythagorean Cholesky steps (P-1S, A-1S, PIPI+,
 * P-2S second passes): chol(G - S^T S + c u tr(G) I), u = 2^-53, G the Gram
 * before the subtraction.  c = 0 (the default) is the reference's
 * arithmetic.  The same rule as the device library's
 * orth_set_pythagorean_shift. */
void oracle_set_pythagorean_shift(double c);
int oracle_get_threads(void);

/* Full orthogonalization: X (m x n, ld ldX) is overwritten by Q, R (ld ldR)
 * receives the block upper-triangular factor (rows below each diagonal block
 * are not written; the caller zeroes R).  info[0] = first block whose
 * Cholesky broke down (1-based, 0 = none), info[1] = its failing pivot.
 * Returns 0, or 2 for an invalid shape (s < 1, n < s, m < n). */
int oracle_bcgs(int variant, int64_t m, int64_t n, int64_t s, double *X,
                int64_t ldX, double *R, int64_t ldR, int *info);

/* Reference-signature wrappers (include/orth.h of the reference): same
 * workspace check and return codes (1 = workspace too small). */
int oracle_bcgsi2P1s(int m, int n, int s, double *Xsub, int ldXsub, double *R,
                     int ldR, double *work, int lwork);
int oracle_bcgsi21s(int m, int n, int s, double *Xsub, int ldXsub, double *R,
                    int ldR, double *work, int lwork);
int oracle_bcgspipi2(int m, int n, int s, double *Xsub, int ldXsub, double *R,
                     int ldR, double *work, int lwork);
int oracle_bcgsi2(int m, int n, int s, double *Xsub, int ldXsub, double *R,
                  int ldR, double *work, int lwork);
int oracle_bcgsi2P2s(int m, int n, int s, double *Xsub, int ldXsub, double *R,
                     int ldR, double *work, int lwork);

/* Intra-block QR (stands in for Trilinos TSQR factorExplicit at one rank):
 * Householder QR of A (m x w), A overwritten by the explicit Q, R (w x w,
 * upper, zero strict-lower) written with ld ldR.  Diagonal of R made
 * non-negative (column signs of Q flipped accordingly). */
void oracle_houseqr(int64_t m, int64_t w, double *A, int64_t ldA, double *R,
                    int64_t ldR);

/* Accuracy harness (SRC/util.c:36-75): out[0] = ||Q^T Q - I||_F,
 * out[1] = ||Q R - X||_F (absolute), out[2] = ||X||_F. */
void oracle_accuracy(int64_t m, int64_t n, const double *X, int64_t ldX,
                     const double *Q, int64_t ldQ, const double *R,
                     int64_t ldR, double *out);

/* Plant a numbered mistake (tests only, 0 = none); see bcgs_oracle.c. */
void oracle_plant(int mistake);

/* Relative perturbation of every first-pass projection (tests only). */
void oracle_set_first_pass_noise(double eta);
/* 1: perturbation proportional to the column norm of the projection (a
 * reduced-precision dot product); 0: relative to each entry. */
void oracle_set_noise_mode(int abs_mode);

/* Primitives, exposed so the tests can pin them against numpy. */
void oracle_gram(int64_t m, int64_t na, const double *A, int64_t lda,
                 int64_t nb, const double *B, int64_t ldb, double *C,
                 int64_t ldc);
int oracle_potrf(int64_t n, double *A, int64_t lda);
void oracle_update(int64_t m, int64_t w, double *X, int64_t ldX, int64_t c,
                   const double *Q, int64_t ldQ, const double *W, int64_t ldW,
                   const double *T, int64_t ldT);

#ifdef __cplusplus
}
#endif

#endif
