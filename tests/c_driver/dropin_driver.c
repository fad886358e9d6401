/*
 * Link-level drop-in check: a plain C caller of include/orth.h, compiled
 * with gcc and linked against liborth_hip.so exactly as a C host code would
 * link the reference's archive.  It follows the reference driver's protocol
 * (TESTS/test_orth.c:12-60) at one rank: X from srand(0) / rand() with the
 * same formula, Q = copy of X, R from calloc, one call of the named variant
 * with a host work buffer of the reference driver's size, then the accuracy
 * harness of SRC/util.c:36-75 (||Q^T Q - I||_F and ||Q R - X||_F) in plain
 * loops.  A second call with lwork one short of the requirement must return
 * 1 and leave Q and R untouched.
 *
 * usage: dropin_driver m n s variant [R-output-file]
 * prints: rc=<rc> small_rc=<rc> loo=<..> resid=<..> xnorm=<..>
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orth.h"

typedef int (*orth_fn)(int, int, int, double *, int, double *, int, double *, int);

static orth_fn lookup(const char *name)
{
    if (!strcmp(name, "bcgsi2P1s")) return bcgsi2P1s;
    if (!strcmp(name, "bcgsi21s")) return bcgsi21s;
    if (!strcmp(name, "bcgspipi2")) return bcgspipi2;
    if (!strcmp(name, "bcgsi2")) return bcgsi2;
    if (!strcmp(name, "bcgsi2P2s")) return bcgsi2P2s;
    return NULL;
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s m n s variant [R-output-file]\n", argv[0]);
        return 2;
    }
    const int m = atoi(argv[1]), n = atoi(argv[2]), s = atoi(argv[3]);
    orth_fn f = lookup(argv[4]);
    if (!f || m < 1 || n < 1 || s < 1) return 2;

    const int is_i2 = !strcmp(argv[4], "bcgsi2");
    const int need = 5 * n * s + 9 * s * s + m * (is_i2 ? s : n);
    const int lwork = need + 2 * n * n;
    double *X = calloc((size_t)m * n, sizeof(double));
    double *Q = calloc((size_t)m * n, sizeof(double));
    double *R = calloc((size_t)n * n, sizeof(double));
    double *work = calloc((size_t)lwork, sizeof(double));
    if (!X || !Q || !R || !work) return 2;
    srand(0);
    for (long i = 0; i < (long)m * n; i++) X[i] = rand() % (10 * m * n) / (double)(10 * m * n + 1);

    /* Too small a workspace: return 1, nothing written. */
    memcpy(Q, X, sizeof(double) * (size_t)m * n);
    const int small_rc = f(m, n, s, Q, m, R, n, work, need - 1);
    for (long i = 0; i < (long)m * n; i++)
        if (Q[i] != X[i]) return 3;
    for (long i = 0; i < (long)n * n; i++)
        if (R[i] != 0.0) return 3;

    memcpy(Q, X, sizeof(double) * (size_t)m * n);
    const int rc = f(m, n, s, Q, m, R, n, work, lwork);

    double loo = 0.0, res = 0.0, xn = 0.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double g = 0.0;
            for (int r = 0; r < m; r++) g += Q[r + (long)i * m] * Q[r + (long)j * m];
            g -= (i == j) ? 1.0 : 0.0;
            loo += g * g;
        }
    for (int j = 0; j < n; j++)
        for (int r = 0; r < m; r++) {
            double v = -X[r + (long)j * m];
            for (int k = 0; k <= j; k++) v += Q[r + (long)k * m] * R[k + (long)j * n];
            res += v * v;
            xn += X[r + (long)j * m] * X[r + (long)j * m];
        }
    printf("rc=%d small_rc=%d loo=%.6e resid=%.6e xnorm=%.6e\n", rc, small_rc, sqrt(loo),
           sqrt(res), sqrt(xn));
    if (argc > 5) {
        FILE *o = fopen(argv[5], "wb");
        if (!o || fwrite(R, sizeof(double), (size_t)n * n, o) != (size_t)n * n) return 2;
        fclose(o);
    }
    free(X);
    free(Q);
    free(R);
    free(work);
    return 0;
}
