"""GPU parity: liborth_hip.so (through the C ABI) against the CPU oracle on
the same seeded inputs, plus size-independent properties at larger sizes.

Tolerances (FP64 everywhere): R and Q agree with the oracle to
1e-11 * kappa-scaled max-norm for well-conditioned inputs; LOO and the
relative residual agree within 10x (or both sit below 1e-14 / 1e-15);
Cholesky breakdown is reported for the same block as the oracle.
"""
import os
import sys

import math

import numpy as np
import pytest

import bcgs_hip as B
import inputs as I
import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
VARIANTS = list(B.VARIANTS)

try:
    import torch
except Exception:  # pragma: no cover - torch is part of the image
    torch = None


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def within10(a, b, floor):
    if a <= floor and b <= floor:
        return True
    return a <= 10 * max(b, floor) and b <= 10 * max(a, floor)


def check_against_oracle(variant, X, s, rtol=1e-11):
    say("case", variant, X.shape, "s=%d" % s)
    Qo, Ro, info_o = O.bcgs(variant, X, s)
    Qg, Rg, st = B.bcgs(variant, X, s)
    assert st["hip_error"] == 0
    assert (st["info_block"], st["info_pivot"]) == info_o
    if info_o != (0, 0):
        return Qg, Rg, st
    scale = np.abs(Ro).max()
    assert np.abs(Rg - Ro).max() <= rtol * scale
    assert np.abs(Qg - Qo).max() <= rtol * 10
    assert np.all(np.tril(Rg, -1) == 0)
    assert within10(I.loo(Qg), I.loo(Qo), 1e-14)
    assert within10(I.rel_residual(X, Qg, Rg), I.rel_residual(X, Qo, Ro), 1e-15)
    return Qg, Rg, st


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("n", [35, 36, 37, 38, 39])
def test_reference_driver_input(variant, n):
    g = np.load(os.path.join(GOLDEN, "test_correct.npz"))
    X = np.asfortranarray(g["X%d" % n])
    Qg, Rg, _ = check_against_oracle(variant, X, 5)
    Rref = g["R%d" % n]
    assert np.abs(Rg - Rref).max() <= 1e-11 * np.abs(Rref).max()


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("shape", [(4096, 64, 16), (4096, 48, 16), (4096, 40, 16),
                                   (4096, 32, 16), (4096, 16, 16), (3000, 70, 8),
                                   (3000, 13, 4), (2000, 96, 32), (5000, 30, 5),
                                   (777, 21, 3)])
def test_shapes(variant, shape):
    m, n, s = shape
    X = I.svd_matrix(m, n, 1e3, seed=21)
    check_against_oracle(variant, X, s)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("kappa", [1e2, 1e6, 1e10])
def test_kappa_sweep(variant, kappa):
    """Same LOO/residual regime as the oracle; breakdown on the same block."""
    X = I.svd_matrix(20000, 128, kappa, seed=22)
    say("case kappa", variant, kappa)
    Qo, Ro, info_o = O.bcgs(variant, X, 16)
    Qg, Rg, st = B.bcgs(variant, X, 16)
    if info_o != (0, 0):
        assert st["info_block"] > 0
        return
    assert st["info_block"] == 0
    assert within10(I.loo(Qg), I.loo(Qo), 1e-14)
    assert within10(I.rel_residual(X, Qg, Rg), I.rel_residual(X, Qo, Ro), 1e-15)


@pytest.mark.parametrize("variant", ["bcgsi2P1s", "bcgsi2", "bcgsi2P2s"])
def test_glued_blocks(variant):
    X = I.glued_matrix(8192, 64, 8, 1e-13)
    say("case glued", variant)
    _, _, info_o = O.bcgs(variant, X, 8)
    Qg, Rg, st = B.bcgs(variant, X, 8)
    assert st["info_block"] == info_o[0]
    if info_o == (0, 0):
        assert I.loo(Qg) < 1e-13


def test_deterministic_bits():
    X = I.svd_matrix(30000, 64, 1e4, seed=23)
    say("case determinism")
    a = B.bcgs("bcgsi2P1s", X, 16)
    b = B.bcgs("bcgsi2P1s", X, 16)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("variant", VARIANTS)
def test_workspace_and_return_codes(variant):
    m, n, s = 600, 20, 5
    X = I.svd_matrix(m, n, 10.0, seed=24)
    need = 5 * n * s + 9 * s * s + m * (s if variant == "bcgsi2" else n)
    say("case workspace", variant)
    rc, Q, R = B.dropin(variant, X, s, lwork=need - 1, work=True)
    assert rc == 1
    rc, Q, R = B.dropin(variant, X, s, lwork=need, work=True)
    assert rc == 0 and I.loo(Q) < 1e-13


def test_sync_counts():
    """One reduction per block for the one-sync variants (plus the three of
    the block-1 intra QR), two for PIPI+."""
    X = I.svd_matrix(4096, 64, 1e2, seed=25)
    N = 4
    expect = {"bcgsi2P1s": 3 + N, "bcgsi21s": 3 + N, "bcgspipi2": 3 + 2 * (N - 1),
              "bcgsi2": 3 + (N - 1) * (2 + 6), "bcgsi2P2s": 3 + 1 + 3 + (N - 1) + 3 * (N - 2)}
    for v, k in expect.items():
        say("case syncs", v)
        _, _, st = B.bcgs(v, X, 16)
        assert st["syncs"] == k, (v, st)


@pytest.mark.skipif(torch is None, reason="torch missing")
def test_device_pointers_match_host_path():
    m, n, s = 10000, 48, 16
    X = I.svd_matrix(m, n, 1e5, seed=26)
    say("case device pointers")
    Qh, Rh, _ = B.bcgs("bcgsi2P1s", X, s)
    Xd = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    Rd = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    B.bcgs_device("bcgsi2P1s", m, n, s, Xd.data_ptr(), m, Rd.data_ptr(), n)
    B.sync()
    torch.cuda.synchronize()
    Qd = Xd.cpu().numpy().T
    R = Rd.cpu().numpy().reshape((n, n), order="F")
    assert np.array_equal(Qd, Qh) and np.array_equal(R, Rh)


@pytest.mark.skipif(torch is None, reason="torch missing")
def test_gram_and_accuracy_device():
    m, n = 50000, 40
    rng = np.random.default_rng(27)
    X = rng.standard_normal((m, n))
    say("case gram/accuracy")
    Xd = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    C = torch.zeros(n * 24, dtype=torch.float64, device="cuda")
    B.gram_device(m, Xd.data_ptr(), m, 3, 24, 10, 17, C.data_ptr(), 24)
    B.sync()
    Ch = C.cpu().numpy().reshape((24, n), order="F")[:, :17]
    ref = X[:, 3:27].T @ X[:, 10:27]
    assert np.abs(Ch - ref).max() <= 1e-12 * np.abs(ref).max()
    Q, R, _ = O.bcgs("bcgsi2", X, 8)
    Qn = np.asfortranarray(Q + 1e-7 * rng.standard_normal(Q.shape))
    Qd = torch.from_numpy(np.ascontiguousarray(Qn.T)).cuda()
    Rd = torch.from_numpy(np.ascontiguousarray(R.T)).cuda()
    got = B.accuracy_device(m, n, Xd.data_ptr(), m, Qd.data_ptr(), m, Rd.data_ptr(), n)
    ref = O.accuracy(X, Qn, R)
    for g_, r_ in zip(got, ref):
        assert abs(g_ - r_) <= 1e-8 * r_


def test_accuracy_gram_is_compensated():
    """The harness's Q^T Q is accumulated in double-double: on columns whose
    exact Gram is known in closed form and whose ||Q^T Q - I|| (~1e-14) is
    far below the rounding of an FP64 accumulation over m = 2^20 rows
    (~sqrt(m) u ~ 1e-13 per entry), the harness returns the exact value.
    Column 0: 2^-10 (1 + s_r 2^-30), column 1: 2^-10 e_r (1 + t_r 2^-30),
    e_r = +-1, with sum(s) = sum(t) = sum(e) = sum(e (s + t)) = 0, so that
    G00 - 1 = 2^-80 sum(s^2), G11 - 1 = 2^-80 sum(t^2) and
    G01 = 2^-80 sum(e s t) exactly (every entry is exact in FP64)."""
    m = 1 << 20
    h = m // 2
    rng = np.random.default_rng(29)
    s1 = rng.integers(-64, 65, h, dtype=np.int64)
    t1 = rng.integers(-64, 65, h, dtype=np.int64)
    e1 = rng.permutation(np.repeat(np.array([1, -1], dtype=np.int64), h // 2))
    # Second half: (s, t) of the first half negated with the same signs e
    # (balanced), so that every first-order sum vanishes while the products
    # e s t add up.
    s = np.concatenate([s1, -s1])
    t = np.concatenate([t1, -t1])
    e = np.concatenate([e1, e1])
    assert s.sum() == 0 and t.sum() == 0 and e.sum() == 0 and (e * (s + t)).sum() == 0
    q0 = 2.0 ** -10 * (1.0 + s.astype(np.float64) * 2.0 ** -30)
    q1 = 2.0 ** -10 * e.astype(np.float64) * (1.0 + t.astype(np.float64) * 2.0 ** -30)
    Q = np.stack([q0, q1], axis=1)
    ss, tt, est = int((s * s).sum()), int((t * t).sum()), int((e * s * t).sum())
    exact = math.sqrt(float(ss * ss + 2 * est * est + tt * tt)) * 2.0 ** -80
    say("case compensated gram", exact)
    Qd = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda()
    Rd = torch.eye(2, dtype=torch.float64, device="cuda")
    loo, res, xn = B.accuracy_device(m, 2, Qd.data_ptr(), m, Qd.data_ptr(), m, Rd.data_ptr(), 2)
    assert abs(loo - exact) <= 1e-12 * exact, (loo, exact)
    assert res == 0.0


@pytest.mark.skipif(torch is None, reason="torch missing")
@pytest.mark.parametrize("variant", VARIANTS)
def test_large_slab_properties(variant):
    """m = 2^22, n = 256, s = 16 (one eighth of the benchmark slab): LOO and
    residual at the oracle's level (same kappa, smaller m: both are
    size-independent for these variants)."""
    m, n, s = 1 << 22, 256, 16
    say("case large", variant)
    g = torch.Generator(device="cuda")
    g.manual_seed(28)
    G = torch.randn((n, m), generator=g, dtype=torch.float64, device="cuda")
    for _ in range(2):  # CholQR2 for an orthonormal U (rows of G^T)
        L = torch.linalg.cholesky(G @ G.T)
        G = torch.linalg.inv(L) @ G
    V, _ = torch.linalg.qr(torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda"))
    sig = torch.logspace(0, -6, n, dtype=torch.float64, device="cuda")
    X = ((V * sig) @ G).contiguous()  # rows of X^T: column-major m x n
    X0 = X.clone()
    Rd = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    B.bcgs_device(variant, m, n, s, X.data_ptr(), m, Rd.data_ptr(), n)
    B.sync()
    loo, res, xn = B.accuracy_device(m, n, X0.data_ptr(), m, X.data_ptr(), m, Rd.data_ptr(), n)
    st = B.status()
    assert st["info_block"] == 0
    assert loo < (1e-12 if variant == "bcgsi21s" else 1e-13)
    assert res / xn < 1e-15


@pytest.mark.parametrize("rowwise", [False, True])
@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("shape", [(4096, 64, 16), (3000, 37, 5), (2000, 96, 32), (3001, 80, 16),
                                   (2001, 96, 32)])
def test_unfused_path_matches_oracle(variant, shape, rowwise, monkeypatch):
    """The separate update / projection kernels (used when a shape does not
    fit the fused pass) stay at parity too, with the paired updates on MFMA
    (row counts that end inside a wave's 32 rows included) or row-wise."""
    monkeypatch.setenv("ORTH_UNFUSED", "1")
    if rowwise:
        monkeypatch.setenv("ORTH_UPDATE_ROWWISE", "1")
    m, n, s = shape
    X = I.svd_matrix(m, n, 1e4, seed=29)
    check_against_oracle(variant, X, s)


@pytest.mark.parametrize("variant", VARIANTS)
def test_fused_and_unfused_agree(variant, monkeypatch):
    X = I.svd_matrix(8192, 80, 1e5, seed=30)
    say("case fused/unfused", variant)
    Qf, Rf, stf = B.bcgs(variant, X, 16)
    monkeypatch.setenv("ORTH_UNFUSED", "1")
    Qu, Ru, stu = B.bcgs(variant, X, 16)
    assert stf["passes"] < stu["passes"]
    assert np.abs(Rf - Ru).max() <= 1e-11 * np.abs(Ru).max()
    assert np.abs(Qf - Qu).max() <= 1e-10


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("shape", [(20000, 512, 32), (12000, 300, 16)])
def test_wide_slabs(variant, shape):
    """n beyond the fused pass's 272-column tile: the separate projection and
    update kernels at the glued-stress width (n = 512, s = 32)."""
    m, n, s = shape
    X = I.svd_matrix(m, n, 1e3, seed=31)
    check_against_oracle(variant, X, s)


@pytest.mark.parametrize("env", ["ORTH_PASS256", "ORTH_GENERIC_PASS"])
@pytest.mark.parametrize("variant", VARIANTS)
def test_alternate_pass_kernels(variant, env, monkeypatch):
    """The 256-thread s = 16 pass and the generic pass agree with the oracle."""
    monkeypatch.setenv("ORTH_PASS256", "1")
    if env == "ORTH_GENERIC_PASS":
        monkeypatch.setenv("ORTH_GENERIC_PASS", "1")
    X = I.svd_matrix(6000, 80, 1e4, seed=32)
    check_against_oracle(variant, X, 16)


@pytest.mark.parametrize("shape", [(70000, 256, 16), (33334, 160, 16), (5000, 96, 16),
                                   (4096, 250, 16), (778, 48, 16), (40002, 112, 16),
                                   (40001, 112, 16)])
@pytest.mark.parametrize("onewg", [False, True])
@pytest.mark.parametrize("variant", VARIANTS)
def test_double_buffered_pass(variant, onewg, shape, monkeypatch):
    """The double-buffered s = 16 pass (direct-to-LDS loads; forced on every
    pass it applies to, with two 512-thread or one 1024-thread workgroup per
    CU) against the oracle and against the register-prefetch pass: partial
    last tiles (m not a multiple of the tile), a ragged last block (n = 250),
    tiny slabs, and an odd row count (which the double-buffered pass
    declines)."""
    m, n, s = shape
    X = I.svd_matrix(m, n, 1e4, seed=34)
    monkeypatch.setenv("ORTH_PASS16D_MINCOLS", "16")
    if onewg:
        monkeypatch.setenv("ORTH_PASS16D_ONEWG", "1")
    Qa, Ra, _ = check_against_oracle(variant, X, s)
    monkeypatch.setenv("ORTH_PASS_REGPF", "1")
    Qb, Rb, _ = check_against_oracle(variant, X, s)
    assert np.abs(Ra - Rb).max() <= 1e-12 * np.abs(Ra).max()


@pytest.mark.parametrize("shape", [(6000, 80, 16), (5000, 70, 8), (3000, 96, 32), (70000, 256, 16)])
@pytest.mark.parametrize("variant", VARIANTS)
def test_separate_reduction_matches_in_kernel(variant, shape, monkeypatch):
    """The synchronization inside the pass launch (ticket tree over the
    workgroups' partials, small step in the last workgroup) against the same
    sum and small step in one separate launch (ORTH_REDUCE_LAUNCH) and in two
    (ORTH_SEPARATE_REDUCE): all agree with the oracle, and bit for bit with
    each other (one summation order; each small-step entry is computed by one
    thread in a fixed order whatever the workgroup size)."""
    m, n, s = shape
    X = I.svd_matrix(m, n, 1e4, seed=33)
    # In-pass synchronization also below its default row bound.
    monkeypatch.setenv("ORTH_IN_PASS_SYNC", "1")
    Qa, Ra, _ = check_against_oracle(variant, X, s)
    monkeypatch.delenv("ORTH_IN_PASS_SYNC")
    monkeypatch.setenv("ORTH_REDUCE_LAUNCH", "1")
    Qb, Rb, _ = check_against_oracle(variant, X, s)
    monkeypatch.delenv("ORTH_REDUCE_LAUNCH")
    monkeypatch.setenv("ORTH_SEPARATE_REDUCE", "1")
    Qc, Rc, _ = check_against_oracle(variant, X, s)
    assert np.array_equal(Ra, Rb) and np.array_equal(Ra, Rc)
    assert np.array_equal(Qa, Qb) and np.array_equal(Qa, Qc)


@pytest.mark.parametrize("variant", VARIANTS)
def test_one_launch_per_synchronization(variant, monkeypatch):
    """Fused schedules synchronize inside the pass launch (forced below the
    default row bound): no reduction launch at all, one launch per pass;
    with ORTH_REDUCE_LAUNCH one reduction launch follows each
    synchronization."""
    X = I.svd_matrix(20000, 96, 1e4, seed=41)
    say("case launches per sync", variant)
    monkeypatch.setenv("ORTH_IN_PASS_SYNC", "1")
    B.kernel_timing(True)
    _, _, st = B.bcgs(variant, X, 16)
    kt = B.kernel_times()
    monkeypatch.delenv("ORTH_IN_PASS_SYNC")
    monkeypatch.setenv("ORTH_REDUCE_LAUNCH", "1")
    B.kernel_timing(True)
    _, _, st2 = B.bcgs(variant, X, 16)
    kt2 = B.kernel_times()
    B.kernel_timing(False)
    assert st["syncs"] > 0 and st2["syncs"] == st["syncs"]
    assert "reduce" not in kt and "proj" not in kt
    assert kt["pass"]["launches"] == st["passes"]
    assert kt2["reduce"]["launches"] == st2["syncs"]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("m", [65538, 9906, 96])
def test_tile_boundary_shapes(variant, m):
    """Several super-tiles per workgroup, a partial last super-tile, and a
    slab shorter than one super-tile, at n = 272 (the widest tile, 32 rows)
    and n = 48 (256-row super-tiles)."""
    for n in (272, 48):
        if m < n:
            continue
        say("case tile boundary", variant, m, n)
        X = I.svd_matrix(m, n, 1e3, seed=35)
        check_against_oracle(variant, X, 16)


@pytest.mark.parametrize("variant", ["bcgsi2P1s", "bcgsi2", "bcgsi2P2s", "bcgspipi2"])
def test_glued_wide(variant):
    """Glued blocks at n = 512, s = 32: breakdown reported on the same block
    as the oracle; otherwise LOO at the oracle's level."""
    X = I.glued_matrix(16384, 512, 32, 1e-10, seed=33)
    say("case glued wide", variant)
    Qo, Ro, info_o = O.bcgs(variant, X, 32)
    Qg, Rg, st = B.bcgs(variant, X, 32)
    assert st["info_block"] == info_o[0]
    if info_o == (0, 0):
        assert within10(I.loo(Qg), I.loo(Qo), 1e-13)


@pytest.mark.skipif(torch is None, reason="torch missing")
@pytest.mark.parametrize("s", [16, 8, 4])
@pytest.mark.parametrize("variant", ["bcgsi2P1s", "bcgsi2"])
@pytest.mark.parametrize("m,ld", [(4097, 4098), (4097, 4097), (6001, 6004), (4098, 4102)])
def test_padded_leading_dimension(variant, m, ld, s):
    """Odd row counts inside an even leading dimension (two-row loads with a
    last row that has no partner), odd leading dimensions (one-row loads) and
    even row counts with pad rows (two-row loads throughout), for the s = 16
    pass and the generic pass (s = 8, 4); the pad rows stay untouched."""
    n = 48
    X = I.svd_matrix(m, n, 1e3, seed=34)
    say("case padded ld", variant, m, ld, s)
    Xp = np.zeros((ld, n))
    Xp[:m] = X
    Xd = torch.from_numpy(np.ascontiguousarray(Xp.T)).cuda()
    Rd = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    B.bcgs_device(variant, m, n, s, Xd.data_ptr(), ld, Rd.data_ptr(), n)
    B.sync()
    torch.cuda.synchronize()
    Qp = Xd.cpu().numpy().T
    R = Rd.cpu().numpy().reshape((n, n), order="F")
    assert np.array_equal(Qp[m:], Xp[m:])  # rows past m untouched
    Qo, Ro, info = O.bcgs(variant, X, s)
    assert info == (0, 0)
    assert np.abs(R - Ro).max() <= 1e-11 * np.abs(Ro).max()
    assert np.abs(Qp[:m] - Qo).max() <= 1e-10


@pytest.mark.parametrize("variant", ["bcgsi2P1s", "bcgspipi2", "bcgsi21s", "bcgsi2P2s"])
def test_pythagorean_shift(variant):
    """The optional Pythagorean shift: GPU against the oracle with the same
    shift, and R's diagonal against its closed form on an input with known
    first-pass Grams."""
    m, n, s, c = 20000, 64, 16, 1.0e7
    X, d = I.layered_matrix(m, n, s)
    say("case pythagorean shift", variant)
    B.set_pythagorean_shift(c)
    O.set_pythagorean_shift(c)
    try:
        Qg, Rg, st = B.bcgs(variant, X, s)
        Qo, Ro, info_o = O.bcgs(variant, X, s)
    finally:
        B.set_pythagorean_shift(0.0)
        O.set_pythagorean_shift(0.0)
    assert st["hip_error"] == 0 and (st["info_block"], st["info_pivot"]) == info_o == (0, 0)
    assert np.abs(Rg - Ro).max() <= 1e-11 * np.abs(Ro).max()
    assert np.abs(Qg - Qo).max() <= 1e-10
    if variant != "bcgsi2P2s":
        want = I.shifted_diag(variant, X, d, s, c)
        assert np.abs(np.abs(np.diag(Rg)) - want).max() <= 1e-12 * want.max()


def test_pythagorean_shift_glued():
    """Glued blocks at delta = 1e-13: P-1S breaks down without the shift and
    completes with it, as in the oracle."""
    m, n, s = 16384, 128, 16
    X = I.glued_matrix(m, n, s, 1e-13, seed=36)
    say("case pythagorean shift glued")
    _, _, st0 = B.bcgs("bcgsi2P1s", X, s)
    assert st0["info_block"] > 0
    c = 11.0 * (m * s + s * (s + 1))
    B.set_pythagorean_shift(c)
    O.set_pythagorean_shift(c)
    try:
        Qg, Rg, st = B.bcgs("bcgsi2P1s", X, s)
        _, _, info_o = O.bcgs("bcgsi2P1s", X, s)
    finally:
        B.set_pythagorean_shift(0.0)
        O.set_pythagorean_shift(0.0)
    assert st["info_block"] == 0 and info_o == (0, 0)
    assert np.all(np.isfinite(Qg)) and np.all(np.isfinite(Rg))

