// Host-harness driver: replays the batched calls of the benchmark (uneven
// shapes, tuning then graph replays, both FP64 variants, and the mixed
// batched form) against the library's host schedules built with the HIP
// stand-in, so that AddressSanitizer / UndefinedBehaviorSanitizer check every
// host-side write (info arrays, configuration, status words, staging) and the
// stand-in checks the capture rules.  Exit status: 0 if every call behaved
// (returned 0, or 2 where a failure was injected) and the stand-in reported
// no misuse.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "orth_ext.h"

extern int g_stub_failures;

struct Batch {
    std::vector<int64_t> m, n, s, ldX, ldR, ldX32, ldQ;
    std::vector<double *> X, R, Q;
    std::vector<float *> X32;
};

static Batch make(int count, unsigned seed, bool mixed, int64_t mmax)
{
    Batch b;
    srand(seed);
    for (int i = 0; i < count; i++) {
        const double u = (double)rand() / RAND_MAX;
        const int64_t m = (int64_t)pow(2.0, 12.0 + u * (log2((double)mmax) - 12.0));
        const int64_t n = mixed ? (int64_t[]){32, 64, 128, 256}[rand() % 4] : (int64_t[]){32, 64, 128}[rand() % 3];
        const int64_t s = mixed ? 16 : (int64_t[]){4, 8, 16}[rand() % 3];
        b.m.push_back(m);
        b.n.push_back(n);
        b.s.push_back(s);
        b.ldX.push_back(m + (i % 3 == 0 ? 1 : 0));
        b.ldR.push_back(n);
        b.ldX32.push_back(m + (i % 5 == 0 ? 3 : 0));
        b.ldQ.push_back(m + (i % 4 == 0 ? 2 : 0));
        double *x = nullptr, *r = nullptr, *q = nullptr;
        float *x32 = nullptr;
        hipMalloc(&x, (size_t)b.ldX.back() * n * 8);
        hipMalloc(&r, (size_t)n * n * 8);
        hipMalloc(&q, (size_t)b.ldQ.back() * n * 8);
        hipMalloc(&x32, (size_t)b.ldX32.back() * n * 4);
        b.X.push_back(x);
        b.R.push_back(r);
        b.Q.push_back(q);
        b.X32.push_back(x32);
    }
    return b;
}

static void release(Batch &b)
{
    for (size_t i = 0; i < b.X.size(); i++) {
        hipFree(b.X[i]);
        hipFree(b.R[i]);
        hipFree(b.Q[i]);
        hipFree(b.X32[i]);
    }
}

static int bad = 0;

static void expect(const char *what, int rc, int want)
{
    printf("%-44s rc=%d\n", what, rc);
    if (rc != want) {
        printf("  unexpected return code (wanted %d)\n", want);
        bad++;
    }
}

int main(int argc, char **argv)
{
    const int count = argc > 1 ? atoi(argv[1]) : 176;
    const int64_t mmax = argc > 2 ? atoll(argv[2]) : (int64_t)1 << 16;
    setvbuf(stdout, nullptr, _IONBF, 0);
    const int want = getenv("STUB_FAIL_LAUNCH") ? -1 : 0;  // -1: 0 or 2 accepted
    hipStream_t st = nullptr;
    hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    auto check = [&](const char *what, int rc) {
        if (want < 0) {
            printf("%-44s rc=%d (failure injected)\n", what, rc);
            if (rc != 0 && rc != 2) bad++;
        } else {
            expect(what, rc, 0);
        }
    };
    // HARNESS_NO_TUNE: the first call of each batch does not tune (its
    // configuration comes from the library's table, or its tuning cache).
    const int tune = getenv("HARNESS_NO_TUNE") ? 0 : 1;
    for (int variant : {ORTH_P1S, ORTH_I2, ORTH_PIPI}) {
        Batch b = make(count, 7 + variant, false, mmax);
        std::vector<int> info(2 * count);
        int cfg[3] = {0, 0, 0};
        check("batched fp64: tune", orth_bcgs_batched_device(
                  variant, count, b.m.data(), b.n.data(), b.s.data(), b.X.data(), b.ldX.data(),
                  b.R.data(), b.ldR.data(), info.data(), tune, cfg, st));
        printf("batched fp64 variant %d config: %d %d\n", variant, cfg[0], cfg[1]);
        for (int rep = 0; rep < 3; rep++)
            check("batched fp64: replay", orth_bcgs_batched_device(
                      variant, count, b.m.data(), b.n.data(), b.s.data(), b.X.data(),
                      b.ldX.data(), b.R.data(), b.ldR.data(), info.data(), 0, cfg, st));
        // A sub-batch (other pointer set): a new graph beside the cached one.
        const int half = count / 2;
        check("batched fp64: half batch", orth_bcgs_batched_device(
                  variant, half, b.m.data(), b.n.data(), b.s.data(), b.X.data(), b.ldX.data(),
                  b.R.data(), b.ldR.data(), info.data(), 0, cfg, st));
        release(b);
    }
    for (int prec : {ORTH_PREC_FP64, ORTH_PREC_BF16, ORTH_PREC_BF16_PROJ}) {
        Batch b = make(count, 11 + prec, true, mmax);
        std::vector<int> info(2 * count);
        int cfg[3] = {0, 0, 0};
        const int rc = orth_bcgs_batched_mixed_device(
            ORTH_P1S, prec, count, b.m.data(), b.n.data(), b.s.data(),
            (const float *const *)b.X32.data(), b.ldX32.data(), b.Q.data(), b.ldQ.data(),
            b.R.data(), b.ldR.data(), info.data(), 1, cfg, st);
        printf("%-44s rc=%d\n", "batched mixed: tune", rc);
        if (rc == 0)
            for (int rep = 0; rep < 3; rep++)
                check("batched mixed: replay", orth_bcgs_batched_mixed_device(
                          ORTH_P1S, prec, count, b.m.data(), b.n.data(), b.s.data(),
                          (const float *const *)b.X32.data(), b.ldX32.data(), b.Q.data(),
                          b.ldQ.data(), b.R.data(), b.ldR.data(), info.data(), 0, cfg, st));
        release(b);
    }
    orth_release();
    hipStreamDestroy(st);
    printf("stand-in misuse reports: %d, unexpected returns: %d\n", g_stub_failures, bad);
    return g_stub_failures || bad ? 1 : 0;
}
