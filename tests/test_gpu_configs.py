"""The benchmark configurations as parity tests, through the C ABI.

1. Reference CPU shape: m = 1e5, n = 200, s = 10, kappa = 1e8, all five
   variants against the oracle on the same bits (LOO and relative residual
   within 10x, same breakdown block, R within kappa u of the oracle's).
2. Device shape n = 256, s = 16 at m = 2^22 for kappa in {1e2, 1e6, 1e10}:
   properties (LOO, residual, which variants break down) on the device, and
   the oracle's breakdown block at m = 2^16 on the same construction.
3. Mixed form at n = 256, s = 16, m = 2^21: FP64 / bf16 / fp16 / bf16x3
   first-pass products for kappa in {1e6, 1e10, 1e14}: FP64 products equal
   FP64 orthogonality with the FP32 rounding as residual; bf16 stays defined
   with a residual at bf16's unit roundoff; fp16 and bf16x3 report their
   Cholesky breakdown (no silent NaN) once u_prec kappa^2 >> 1.
4. Glued blocks at n = 512, s = 32 (delta in {1e-6, 1e-10, 1e-13}), all
   five variants: against the oracle at m = 2^16 (same breakdown block, LOO
   within 10x), properties at m = 2^22, and the FP32-storage form.
5. Batched form: 96 uneven instances (m log-uniform in [4096, 2^20], n in
   {32, 64, 128}, s in {4, 8, 16}, kappa log-uniform in [1e2, 1e10]), the
   tuning call's own output and a replay against the oracle per instance.
"""
import sys

import numpy as np
import pytest

import bcgs_hip as B
import device_inputs as D
import inputs as I
import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

VARIANTS = list(B.VARIANTS)
U = 2.0 ** -53


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def within10(a, b, floor):
    if a <= floor and b <= floor:
        return True
    return a <= 10 * max(b, floor) and b <= 10 * max(a, floor)


def device_run(variant, X0, m, n, s):
    """One factorization of the device input X0 (n x m row-major) on a copy;
    returns (status, loo, rel_residual, Q tensor, R tensor)."""
    X = X0.clone()
    R = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    B.bcgs_device(variant, m, n, s, X.data_ptr(), m, R.data_ptr(), n)
    B.sync()
    st = B.status()
    loo, res, xn = B.accuracy_device(m, n, X0.data_ptr(), m, X.data_ptr(), m, R.data_ptr(), n)
    return st, loo, res / xn, X, R


# ---------------------------------------------------------------- config 1

@pytest.fixture(scope="module")
def cpu_shape_input():
    return D.to_host(D.svd_input(torch, 100000, 200, 1e8, seed=2))


@pytest.mark.parametrize("variant", VARIANTS)
def test_config1_reference_cpu_shape(variant, cpu_shape_input):
    X = np.asfortranarray(cpu_shape_input)
    say("case config1", variant)
    Qo, Ro, info_o = O.bcgs(variant, X, 10)
    Qg, Rg, st = B.bcgs(variant, X, 10)
    assert st["hip_error"] == 0
    assert (st["info_block"] > 0) == (info_o[0] > 0)
    if info_o[0]:
        assert st["info_block"] == info_o[0]
        return
    lg, lo = I.loo(Qg), I.loo(Qo)
    rg, ro = I.rel_residual(X, Qg, Rg), I.rel_residual(X, Qo, Ro)
    say("  loo gpu %.2e oracle %.2e  resid gpu %.2e oracle %.2e" % (lg, lo, rg, ro))
    assert within10(lg, lo, 1e-14)
    assert within10(rg, ro, 1e-15)
    # The same algorithm in another rounding order: R agrees to kappa u.
    assert np.abs(Rg - Ro).max() <= 1e8 * 64 * U * np.abs(Ro).max()
    assert np.all(np.tril(Rg, -1) == 0)


# ---------------------------------------------------------------- config 2

@pytest.mark.parametrize("kappa", [1e2, 1e6, 1e10])
def test_config2_device_shape_properties(kappa):
    m, n, s = 1 << 22, 256, 16
    X0 = D.svd_input(torch, m, n, kappa, seed=1)
    for v in VARIANTS:
        say("case config2", v, "kappa %g" % kappa)
        st, loo, res, X, R = device_run(v, X0, m, n, s)
        del X, R
        say("  info %d loo %.2e resid %.2e" % (st["info_block"], loo, res))
        if kappa >= 1e10 and v in ("bcgsi2P1s", "bcgspipi2"):
            # The Pythagorean first pass needs u kappa^2 < 1.
            assert st["info_block"] > 0
            continue
        assert st["info_block"] == 0
        assert loo < (1e-12 if v == "bcgsi21s" else 1e-13)
        assert res < 1e-15
    del X0
    torch.cuda.empty_cache()


def test_config2_breakdown_block_matches_oracle():
    m, n, s = 1 << 16, 256, 16
    X0 = D.svd_input(torch, m, n, 1e10, seed=1)
    X = np.asfortranarray(D.to_host(X0))
    for v in VARIANTS:
        say("case config2 oracle", v)
        _, _, info_o = O.bcgs(v, X, s)
        st, loo, res, Q, R = device_run(v, X0, m, n, s)
        assert st["info_block"] == info_o[0]
        if info_o[0] == 0:
            Qo, Ro, _ = O.bcgs(v, X, s)
            assert within10(loo, I.loo(Qo), 1e-14)


# ---------------------------------------------------------------- config 3

@pytest.mark.parametrize("kappa", [1e6, 1e10, 1e14])
def test_config3_mixed_precisions(kappa):
    m, n, s = 1 << 21, 256, 16
    X0 = D.svd_input(torch, m, n, kappa, seed=1)
    X32 = X0.float()
    Q = torch.empty_like(X0)
    R = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    for prec in ("fp64", "bf16", "fp16", "bf16x3"):
        say("case config3", prec, "kappa %g" % kappa)
        Q.fill_(0.0)
        R.zero_()
        B.bcgs_mixed_device("bcgsi2P1s", prec, m, n, s, X32.data_ptr(), m, Q.data_ptr(), m,
                            R.data_ptr(), n)
        B.sync()
        st = B.status()
        loo, res, xn = B.accuracy_device(m, n, X0.data_ptr(), m, Q.data_ptr(), m,
                                         R.data_ptr(), n)
        res /= xn
        say("  info %d loo %.2e resid %.2e" % (st["info_block"], loo, res))
        if prec == "fp64":
            # FP64 products on the FP32-rounded input: FP64 orthogonality;
            # the residual against the FP64 input is the input rounding.
            assert st["info_block"] == 0
            assert loo < 1e-13
            assert 1e-9 < res < 1e-7
        else:
            # Reduced-precision first-pass projections perturb the
            # Pythagorean difference X^T X - S^T S at the format's unit
            # roundoff, so P-1S breaks down once that exceeds the block's
            # new directions (u_prec kappa^2 >> 1; data dependent: bf16 at
            # m = 2^24 stayed defined where it breaks down at m = 2^21).
            # A breakdown is always reported, never a silent NaN; a defined
            # factorization has the format's unit roundoff as residual.
            finite = bool(torch.isfinite(Q).all()) and bool(torch.isfinite(R).all())
            if not finite:
                assert st["info_block"] > 0
            if st["info_block"] == 0:
                assert finite
                assert res < 1e-2
                if prec == "bf16" and kappa <= 1e6:
                    assert loo < 1e-10
    del X0, X32, Q, R
    torch.cuda.empty_cache()


# ---------------------------------------------------------------- config 4

@pytest.mark.parametrize("delta", [1e-6, 1e-10, 1e-13])
def test_config4_glued_against_oracle(delta):
    m, n, s = 1 << 16, 512, 32
    X0 = D.glued_input(torch, m, n, s, delta, seed=3)
    X = np.asfortranarray(D.to_host(X0))
    for v in VARIANTS:
        say("case config4 oracle", v, "delta %g" % delta)
        Qo, Ro, info_o = O.bcgs(v, X, s)
        st, loo, res, Q, R = device_run(v, X0, m, n, s)
        say("  gpu info %d/%d oracle %s" % (st["info_block"], st["info_pivot"], info_o))
        assert st["info_block"] == info_o[0]
        if info_o[0] == 0:
            say("  loo gpu %.2e oracle %.2e" % (loo, I.loo(Qo)))
            assert within10(loo, I.loo(Qo), 1e-13)
            assert within10(res, I.rel_residual(X, Qo, Ro), 2e-15)


@pytest.mark.parametrize("delta", [1e-6, 1e-13])
def test_config4_glued_device_shape(delta):
    m, n, s = 1 << 22, 512, 32
    X0 = D.glued_input(torch, m, n, s, delta, seed=3)
    for v in VARIANTS:
        say("case config4", v, "delta %g" % delta)
        st, loo, res, X, R = device_run(v, X0, m, n, s)
        del X, R
        say("  info %d/%d loo %.2e resid %.2e" % (st["info_block"], st["info_pivot"], loo, res))
        if delta <= 1e-13 and v in ("bcgsi2P1s", "bcgspipi2"):
            assert st["info_block"] > 0
            continue
        assert st["info_block"] == 0
        # BCGSI+A-1S loses orthogonality to O(u kappa) on nearly dependent
        # blocks; the others keep O(u).
        assert loo < (1e-8 if v == "bcgsi21s" else 1e-12)
        assert res < 2e-15
    # FP32 storage (FP64 products): the rounding to FP32 perturbs every
    # block far above delta, so no variant breaks down.
    X32 = X0.float()
    Q = torch.empty_like(X0)
    R = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    for v in ("bcgsi2P1s", "bcgsi2"):
        say("case config4 fp32 storage", v, "delta %g" % delta)
        R.zero_()
        B.bcgs_mixed_device(v, "fp64", m, n, s, X32.data_ptr(), m, Q.data_ptr(), m,
                            R.data_ptr(), n)
        B.sync()
        st = B.status()
        loo, res, xn = B.accuracy_device(m, n, X0.data_ptr(), m, Q.data_ptr(), m,
                                         R.data_ptr(), n)
        say("  info %d loo %.2e resid %.2e" % (st["info_block"], loo, res / xn))
        assert st["info_block"] == 0
        assert loo < 1e-12
        assert res / xn < 1e-7
    del X0, X32, Q, R
    torch.cuda.empty_cache()


# ---------------------------------------------------------------- config 5

def _batch(count, seed):
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(count):
        m = int(2 ** rng.uniform(12, 20))
        shapes.append((m, int(rng.choice([32, 64, 128])), int(rng.choice([4, 8, 16])),
                       float(10 ** rng.uniform(2, 10))))
    # The largest instance of the configuration's range.
    shapes[0] = (1 << 20, 128, 16, 1e4)
    return shapes


@pytest.mark.parametrize("variant", ["bcgsi2P1s", "bcgsi2"])
def test_config5_batched_against_oracle(variant):
    shapes = _batch(96, seed=11)
    X0 = [D.svd_input(torch, m, n, k, seed=300 + i) for i, (m, n, s, k) in enumerate(shapes)]
    X = [t.clone() for t in X0]
    R = [torch.zeros(n * n, dtype=torch.float64, device="cuda") for (m, n, s, k) in shapes]
    ms = [sh[0] for sh in shapes]
    ns = [sh[1] for sh in shapes]
    ss = [sh[2] for sh in shapes]
    torch.cuda.synchronize()

    def call(tune):
        return B.bcgs_batched_device(variant, ms, ns, ss, [t.data_ptr() for t in X], ms,
                                     [r.data_ptr() for r in R], ns, tune=tune)

    oracle_out = {}

    def check(info, label):
        worst = 0.0
        for i, (m, n, s, k) in enumerate(shapes):
            Xh = np.asfortranarray(D.to_host(X0[i]))
            if i not in oracle_out:
                Qo, Ro, info_o = O.bcgs(variant, Xh, s)
                oracle_out[i] = (I.loo(Qo), I.rel_residual(Xh, Qo, Ro), info_o)
            loo_o, res_o, info_o = oracle_out[i]
            if info[i][0] != info_o[0]:
                # Only where u kappa^2 is near or above 1 may rounding decide
                # whether, and at which block, the Pythagorean Cholesky
                # factorization breaks down.
                assert variant in ("bcgsi2P1s", "bcgspipi2") and k > 1e7, (
                    label, i, shapes[i], info[i], info_o)
                continue
            if info_o[0]:
                continue
            loo, res, xn = B.accuracy_device(m, n, X0[i].data_ptr(), m, X[i].data_ptr(), m,
                                             R[i].data_ptr(), n)
            assert within10(loo, loo_o, 1e-14), (label, i, shapes[i], loo, loo_o)
            assert within10(res / xn, res_o, 1e-15), (label, i, shapes[i], res / xn, res_o)
            worst = max(worst, loo)
        say("  %s: worst LOO %.2e" % (label, worst))

    say("case config5", variant, "tuning call on", len(shapes), "instances")
    info, cfg = call(True)
    say("  tuned", cfg)
    # The tuning call's own output is the factorization of the caller's X.
    check(info, "tuning call")
    for t, t0, r in zip(X, X0, R):
        t.copy_(t0)
        r.zero_()
    torch.cuda.synchronize()
    say("case config5", variant, "replay")
    info, cfg2 = call(False)
    assert cfg2 == cfg
    check(info, "replay")
