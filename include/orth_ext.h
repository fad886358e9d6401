/*
 * Extended interface of liborth_hip.so: device-pointer and stream entry
 * points, status reporting, the accuracy harness, the mixed-precision
 * form and the batched form.  The five reference-compatible symbols are in
 * orth.h; nothing here has a counterpart in the reference's include/orth.h
 * unless its comment cites one.
 *
 * Conventions: matrices are column-major FP64 unless stated; sizes are
 * 64-bit; `stream` is a hipStream_t passed as void* (NULL = null stream).
 * Device entry points enqueue work and return without waiting; successive
 * calls on one device are ordered by the library even across streams,
 * because they share the library's device scratch.
 */
#ifndef ORTH_EXT_H
#define ORTH_EXT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Variant ids. */
enum {
    ORTH_P1S = 0,  /* bcgsi2P1s  */
    ORTH_1S = 1,   /* bcgsi21s   */
    ORTH_PIPI = 2, /* bcgspipi2  */
    ORTH_I2 = 3,   /* bcgsi2     */
    ORTH_P2S = 4   /* bcgsi2P2s  */
};

/* Precision of the first-pass products in the mixed form. */
enum {
    ORTH_PREC_FP64 = 0,   /* FP64 products; only the input storage is FP32    */
    ORTH_PREC_BF16 = 1,   /* bf16 MFMA operands, FP32 accumulation            */
    ORTH_PREC_FP16 = 2,   /* fp16 MFMA operands, FP32 accumulation            */
    ORTH_PREC_BF16X3 = 3, /* split bf16 operands (hi*lo + lo*hi + hi*hi)      */
    /* The same three for the first-pass projection and Gram only; the
     * first-pass update X - Q S stays FP64. */
    ORTH_PREC_BF16_PROJ = 4,
    ORTH_PREC_FP16_PROJ = 5,
    ORTH_PREC_BF16X3_PROJ = 6
};

typedef struct {
    int info_block;  /* 1-based block whose Cholesky broke down first, 0 = none */
    int info_pivot;  /* 1-based pivot of that breakdown                         */
    int hip_error;   /* last hipError_t seen by the library (0 = none)          */
    int passes;      /* streaming passes over the row slab launched by the call */
    int syncs;       /* global reductions (the reference's MPI_Allreduce count) */
} orth_status_t;

/* Status of the last completed call on the calling thread's device;
 * waits for that call's work to finish.  Returns 0. */
int orth_last_status(orth_status_t *st);

/* Device form of the five variants: dX (m x n, ld ldX) becomes Q in place,
 * dR (ld ldR) receives R exactly as in orth.h.  Returns 0, or 2 for an
 * invalid shape / HIP error. */
int orth_bcgs_device(int variant, int64_t m, int64_t n, int64_t s, double *dX,
                     int64_t ldX, double *dR, int64_t ldR, void *stream);

/* Allocation-free form of orth_bcgs_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc / hipFree and creates no event: its scratch is a
 * fixed 48 MiB region of device memory inside the library (one per device,
 * held until the process ends), whose address is looked up on the first call
 * on each device.  Limits: n <= 512 columns (larger n returns 2 with a
 * message on stderr); the other shape rules are orth_bcgs_device's, and the
 * passes use fewer row chunks when their partial sums would not fit.  R is
 * written as by orth_bcgs_device; the first breakdown (1-based block, pivot)
 * goes to dInfo (device memory, 2 ints, may be NULL) instead of
 * orth_last_status().  Every call on a device shares that device's region:
 * calls that overlap on one device (other streams, other threads) race.
 * Enqueues on `stream` and returns; 0, or 2 for an invalid or too large
 * shape or a HIP error. */
int orth_bcgs_device_noalloc(int variant, int64_t m, int64_t n, int64_t s, double *dX,
                             int64_t ldX, double *dR, int64_t ldR, int *dInfo,
                             void *stream);

/* Mixed-precision form of BCGSI+P-1S (no counterpart in the reference):
 * the input X32 (m x n FP32, ld ldX32, device, not modified) is
 * orthogonalized into Q (m x n FP64, ld ldQ, device) and R (FP64, as in
 * orth.h).  `precision` selects the arithmetic of the first-pass products
 * (the update X - Q S and the projection Q^T X with the Gram X^T X of a
 * block): ORTH_PREC_BF16 / _FP16 on MFMA with FP32 accumulation per 32-row
 * tile, ORTH_PREC_BF16X3 (FP32 values split into two bf16 terms, three
 * products), or ORTH_PREC_FP64; the _PROJ precisions reduce only the
 * projection and Gram and keep the update X - Q S in FP64 (so the residual
 * stays at the FP32 input's rounding).  The reorthogonalization (second-pass
 * update and projection), the Pythagorean subtraction, the Cholesky factors
 * and the triangular solves are FP64 in every case.  The reduced
 * precisions support variant ORTH_P1S with s = 16 and n <= 256 (the fused
 * mixed pass).  ORTH_PREC_FP64 is accepted for every variant and shape of
 * orth_bcgs_device: outside those limits the input is widened exactly into
 * Q and the FP64 form runs in place.  Returns 0, or 2 for an unsupported
 * shape / precision or a HIP error; breakdown is reported by
 * orth_last_status() as for the FP64 form. */
int orth_bcgs_mixed_device(int variant, int precision, int64_t m, int64_t n,
                           int64_t s, const float *dX32, int64_t ldX32,
                           double *dQ, int64_t ldQ, double *dR, int64_t ldR,
                           void *stream);

/* Batched form (no counterpart in the reference): `count` independent
 * orthogonalizations of uneven shapes, each as orth_bcgs_device (instance i:
 * m[i] x n[i] with block width s[i], device dX[i] (ld ldX[i]) becomes Q,
 * device dR[i] (ld ldR[i]) receives R).  Instances run side by side on
 * several internal streams, each with private scratch; the batch is
 * captured into hipGraphs and replayed while the same batch (shapes and
 * pointers) is passed again.  tune != 0 first times the stream count and
 * workgroup share on this batch and keeps the best for batches with the
 * same shapes; it also decides between replaying captured graphs and
 * launching directly on the streams, and picks the tile shape of the s = 16
 * passes.  When the environment variable ORTH_TUNE_CACHE names a file, the
 * tuned configurations are written there (one text line per device and
 * batch shape) and read back by later processes, so that a call with
 * tune == 0 uses a stored configuration.  Batches larger than a graph-size
 * limit are captured as several graphs launched back to back.  info (host,
 * 2 * count ints, may be NULL) receives each instance's breakdown (block,
 * pivot); config (host, 2 ints, may be NULL) the number of streams used
 * (negative when launched directly rather than from graphs) and the
 * workgroup share in thousandths.
 * Synchronous (waits for `stream` first).  Returns 0, or 2 for an invalid
 * instance or a HIP error. */
int orth_bcgs_batched_device(int variant, int count, const int64_t *m,
                             const int64_t *n, const int64_t *s,
                             double *const *dX, const int64_t *ldX,
                             double *const *dR, const int64_t *ldR, int *info,
                             int tune, int *config, void *stream);

/* The batched form of orth_bcgs_mixed_device: FP32 inputs dX32[i], FP64
 * outputs dQ[i] and dR[i].  Switched off in this version: every call
 * returns 2 (use orth_bcgs_mixed_device per instance). */
int orth_bcgs_batched_mixed_device(int variant, int precision, int count,
                                   const int64_t *m, const int64_t *n,
                                   const int64_t *s, const float *const *dX32,
                                   const int64_t *ldX32, double *const *dQ,
                                   const int64_t *ldQ, double *const *dR,
                                   const int64_t *ldR, int *info, int tune,
                                   int *config, void *stream);

/* Accuracy harness of the reference (SRC/util.c:36-75) on device data:
 * out[0] = ||Q^T Q - I||_F, out[1] = ||Q R - X||_F, out[2] = ||X||_F.
 * Q^T Q is accumulated with compensated (double-double) arithmetic, so its
 * own rounding stays far below the loss of orthogonality it measures.
 * Synchronous; out is host memory. */
int orth_accuracy_device(int64_t m, int64_t n, const double *dX, int64_t ldX,
                         const double *dQ, int64_t ldQ, const double *dR,
                         int64_t ldR, double *out, void *stream);

/* Block Gram primitive: dC (na x nb, ld ldc, device) = A^T B over m rows
 * with A = columns [a0, a0+na) and B = columns [b0, b0+nb) of dX; the
 * reduction order is fixed, so results are bitwise reproducible. */
int orth_gram_device(int64_t m, const double *dX, int64_t ldX, int64_t a0,
                     int64_t na, int64_t b0, int64_t nb, double *dC,
                     int64_t ldc, void *stream);

/* Block update primitive (one row-local pass over the slab):
 *   X(:, col0:col0+w) <- (X(:, col0:col0+w) - X(:, 0:c) W) T^{-1}
 * W is c x w (ld ldW, device), T is w x w upper (ld w, device) or NULL.
 * Requires c <= col0 and w <= 32.  Synchronous. */
int orth_update_device(int64_t m, double *dX, int64_t ldX, int64_t col0,
                       int64_t w, int64_t c, const double *dW, int64_t ldW,
                       const double *dT, void *stream);

/* Optional shift of the Pythagorean Cholesky steps (no counterpart in the
 * reference, which factors G - S^T S as is): later calls on every device
 * factor G - S^T S + c u tr(G) I instead, u = 2^-53 and G the Gram of the
 * block before the subtraction (P-1S, A-1S, PIPI+ and the P-2S second
 * pass; BCGSI+ has no Pythagorean step).  c = 0, the default, restores the
 * reference's arithmetic.  A shift keeps the factorization defined when
 * G - S^T S has lost positive definiteness to rounding (nearly dependent
 * blocks); it does not restore orthogonality there.  Returns 0, or 2 for a
 * negative or non-finite c. */
int orth_set_pythagorean_shift(double c);

/* Kernel classes for orth_kernel_times. */
enum {
    ORTH_K_PROJ = 0,    /* projection partials A^T B (MFMA FP64)       */
    ORTH_K_REDUCE = 1,  /* fixed-order sum of the partials             */
    ORTH_K_UPDATE = 2,  /* row-local block updates                     */
    ORTH_K_SMALL = 3,   /* one-workgroup s x s steps                   */
    ORTH_K_PASS = 4,    /* fused update + projection pass              */
    ORTH_K_COUNT = 8
};

/* Enable (1) or disable (0) per-launch event timing on this device and
 * clear what was recorded.  Events are recorded on the stream of each
 * launch, around the launch. */
int orth_kernel_timing(int enable);

/* Total elapsed milliseconds, launch count, algorithmic flops and
 * algorithmic HBM bytes recorded for one kernel class since the last
 * orth_kernel_timing call (waits for the device).  flops/bytes may be NULL. */
int orth_kernel_times(int kind, double *total_ms, int *count, double *flops,
                      double *bytes);

/* Streaming ceiling of the device (measurement aid, no counterpart in the
 * reference): mode 0 times a plain copy of `bytes` from src to dst, mode 1 a
 * plain read of `bytes` of src, both with 16-byte non-temporal accesses and
 * eight 256-thread workgroups per CU; *ms_out receives the HIP-event time of
 * the kernel.  Synchronous (waits for the device first).  Returns 0, or 2
 * for invalid arguments or a HIP error. */
int orth_stream_probe(int mode, double *dst, const double *src, int64_t bytes, double *ms_out);

/* Wait for all work the library enqueued on this device. */
int orth_sync(void);

/* Release the library's device scratch on this device, including the
 * batched form's captured graphs and per-stream scratch (the next call
 * allocates and captures again). */
int orth_release(void);

#ifdef __cplusplus
}
#endif

#endif
