/*
 * Drop-in C interface of the reorthogonalized block classical Gram-Schmidt
 * (BCGS) family, served by liborth_hip.so on an AMD Instinct MI355X (gfx950).
 *
 * Each entry point keeps the exact signature, argument meaning and return
 * codes of the reference's include/orth.h:8-17:
 *
 *   m       global row count of X (one process: the local row count msub = m,
 *           the row split of SRC/orth.c:69-76 with one rank)
 *   n, s    number of columns and block width: blocks of s columns from the
 *           left, the last block holding the n mod s remaining columns when
 *           s does not divide n
 *   Xsub    m x n column-major, leading dimension ldXsub; overwritten by Q
 *   R       n x n column-major, leading dimension ldR; for block column k the
 *           rows 0 .. (end of block k) are written, rows below are never
 *           touched, so the caller zeroes R first (as TESTS/test_orth.c:17)
 *   work    scratch; only its size is checked (lwork doubles, 5ns + 9s^2 +
 *           m*n, or 5ns + 9s^2 + m*s for bcgsi2).  work == NULL selects
 *           library-managed device scratch, which is the only way to call at
 *           sizes where m*n does not fit in an int.
 *
 * Return value: 0 on success, 1 when work != NULL and lwork is too small
 * (SRC/orth.c:82-86), 2 for an invalid shape (s < 1, n < s, m < n, ld too
 * small) or a device failure.  Cholesky breakdown does not change the return
 * value (the reference ignores dpotrf_'s info); orth_last_status() in
 * orth_ext.h reports it.
 *
 * Xsub and R may each be host or device (hipMalloc) memory; host buffers are
 * staged through device memory.  Calls are synchronous.
 */
#ifndef ORTH_H
#define ORTH_H

#ifdef __cplusplus
extern "C" {
#endif

/* BCGSI+P-1S, one synchronization per block.  Replaces SRC/orth.c:48-222
 * (declared at include/orth.h:8-9 of the reference). */
int bcgsi2P1s(int m, int n, int s, double *Xsub, int ldXsub, double *R,
              int ldR, double *work, int lwork);

/* BCGSI+A-1S (Yamazaki et al.).  Replaces SRC/orth.c:224-367
 * (reference include/orth.h:10-11). */
int bcgsi21s(int m, int n, int s, double *Xsub, int ldXsub, double *R,
             int ldR, double *work, int lwork);

/* BCGS-PIPI+, two synchronizations per block.  Replaces SRC/orth.c:369-462
 * (reference include/orth.h:12-13). */
int bcgspipi2(int m, int n, int s, double *Xsub, int ldXsub, double *R,
              int ldR, double *work, int lwork);

/* BCGSI+ (BCGS2) with intra-block QR in both passes.  Replaces
 * SRC/orth.c:464-548 (reference include/orth.h:14-15). */
int bcgsi2(int m, int n, int s, double *Xsub, int ldXsub, double *R, int ldR,
           double *work, int lwork);

/* BCGSI+P-2S, two synchronizations per block.  Replaces SRC/orth.c:550-697
 * (reference include/orth.h:16-17). */
int bcgsi2P2s(int m, int n, int s, double *Xsub, int ldXsub, double *R,
              int ldR, double *work, int lwork);

#ifdef __cplusplus
}
#endif

#endif
