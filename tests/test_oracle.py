"""Pins the CPU oracle (oracle/bcgs_oracle.c) before it is used as the checker.

Pins used (the reference ships no expected outputs):
  * primitives against numpy/LAPACK on the same inputs;
  * the reference driver's own input (TESTS/test_orth.c:42-44) against
    LAPACK Householder R factors stored in tests/golden/test_correct.npz;
  * closed forms X = Q0 R0 (tests/golden/closed_form.npz) and X = Q0;
  * a 60-digit decimal Gram-Schmidt R on a tiny input (brute_force.npz);
  * QR uniqueness invariants (orthonormal Q, small residual, triangular R);
  * planted mistakes: each one must be rejected by these checks.
"""
import io
import os
import contextlib

import numpy as np
import pytest

import inputs as I
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
VARIANTS = list(O.VARIANTS)
EPS = np.finfo(np.float64).eps


def written_mask(n, s):
    """Entries of R that the reference writes (block column k: rows up to the
    end of block k)."""
    mask = np.zeros((n, n), dtype=bool)
    c = 0
    b = 0
    while c < n:
        w = s if b == 0 else min(s, n - c)
        mask[: c + w, c: c + w] = True
        c += w
        b += 1
    return mask


# ---------------------------------------------------------------- primitives

def test_gram_matches_numpy():
    rng = np.random.default_rng(0)
    A = np.asfortranarray(rng.standard_normal((3001, 7)))
    B = np.asfortranarray(rng.standard_normal((3001, 5)))
    C = O.gram(A, B)
    ref = A.T @ B
    assert np.abs(C - ref).max() <= 64 * 3001 * EPS * np.abs(A).max() * np.abs(B).max()


def test_potrf_matches_numpy():
    rng = np.random.default_rng(1)
    M = rng.standard_normal((12, 12))
    G = M.T @ M + 12 * np.eye(12)
    U, info = O.potrf(G)
    assert info == 0
    assert np.allclose(U, np.linalg.cholesky(G).T, rtol=1e-13, atol=1e-13)
    assert np.all(np.tril(U, -1) == 0)


def test_potrf_reports_first_bad_pivot():
    G = np.diag([4.0, 1.0, -2.0, 3.0])
    _, info = O.potrf(G)
    assert info == 3


def test_update_matches_formula():
    rng = np.random.default_rng(2)
    m, c, w = 1000, 9, 4
    Q = rng.standard_normal((m, c))
    X = rng.standard_normal((m, w))
    W = rng.standard_normal((c, w))
    T = np.triu(rng.standard_normal((w, w))) + 3 * np.eye(w)
    got = O.update(X, Q, W, T)
    ref = np.linalg.solve(T.T, (X - Q @ W).T).T
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
    got2 = O.update(X, Q, W, None)
    assert np.allclose(got2, X - Q @ W, rtol=1e-13, atol=1e-13)


def test_houseqr_matches_lapack():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((700, 9))
    Q, R = O.houseqr(A)
    Rl = np.linalg.qr(A, mode="r")
    Rl = Rl * np.sign(np.diag(Rl))[:, None]
    assert np.allclose(R, Rl, rtol=1e-12, atol=1e-12)
    assert np.all(np.diag(R) >= 0)
    assert I.loo(Q) < 1e-14
    assert I.rel_residual(A, Q, R) < 1e-15


# ---------------------------------------------------------------- golden inputs

@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("n", [35, 36, 37, 38, 39])
def test_reference_driver_input(variant, n):
    """The reference driver's input (m=550, s=5, ragged last block for n=36..39)
    against LAPACK Householder R."""
    g = np.load(os.path.join(GOLDEN, "test_correct.npz"))
    X = np.asfortranarray(g["X%d" % n])
    Rref = g["R%d" % n]
    Q, R, info = O.bcgs(variant, X, 5)
    assert info == (0, 0)
    assert np.abs(R - Rref).max() <= 1e-11 * np.abs(Rref).max()
    assert I.loo(Q) < 1e-13
    assert I.rel_residual(X, Q, R) < 1e-15


@pytest.mark.parametrize("n", [35, 36, 37, 38, 39])
def test_driver_input_regenerates_from_glibc(n):
    g = np.load(os.path.join(GOLDEN, "test_correct.npz"))
    X = I.test_correct_matrix(550, n, rank=0)
    assert np.array_equal(X, g["X%d" % n])


@pytest.mark.parametrize("variant", VARIANTS)
def test_closed_form_QR(variant):
    g = np.load(os.path.join(GOLDEN, "closed_form.npz"))
    Q0, R0 = g["Q0"], g["R0"]
    X = np.asfortranarray(Q0 @ R0)
    Q, R, info = O.bcgs(variant, X, 8)
    kappa = np.linalg.cond(R0)
    assert info == (0, 0)
    assert np.abs(R - R0).max() <= 1e-14 * kappa * np.abs(R0).max()
    assert np.abs(Q - Q0).max() <= 1e-13 * kappa


@pytest.mark.parametrize("variant", VARIANTS)
def test_closed_form_orthonormal_input(variant):
    g = np.load(os.path.join(GOLDEN, "closed_form.npz"))
    Q0 = np.asfortranarray(g["Q0"])
    Q, R, _ = O.bcgs(variant, Q0, 6)
    assert np.abs(R - np.eye(Q0.shape[1])).max() < 1e-14
    assert np.abs(Q - Q0).max() < 1e-14


@pytest.mark.parametrize("variant", VARIANTS)
def test_brute_force_tiny(variant):
    g = np.load(os.path.join(GOLDEN, "brute_force.npz"))
    X = np.asfortranarray(g["X"])
    Q, R, _ = O.bcgs(variant, X, 2)
    assert np.abs(R - g["R"]).max() <= 1e-13 * np.abs(g["R"]).max()


# ---------------------------------------------------------------- shapes, R layout

@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("shape", [(600, 16, 16), (600, 32, 16), (600, 40, 16),
                                   (600, 52, 16), (600, 7, 1), (600, 13, 4),
                                   (900, 96, 32)])
def test_edge_shapes(variant, shape):
    m, n, s = shape
    X = I.svd_matrix(m, n, 1e3, seed=5)
    Q, R, info = O.bcgs(variant, X, s)
    assert info == (0, 0)
    assert I.loo(Q) < 1e-12
    assert I.rel_residual(X, Q, R) < 1e-15
    assert np.all(np.tril(R, -1) == 0)
    assert np.all(np.diag(R) > 0)


@pytest.mark.parametrize("variant", VARIANTS)
def test_rows_below_blocks_untouched(variant):
    m, n, s = 500, 37, 5
    X = I.svd_matrix(m, n, 1e2, seed=6)
    Q = X.copy(order="F")
    R = np.full((n, n), 7.25, order="F")
    import ctypes
    info = (ctypes.c_int * 2)()
    rc = O.lib().oracle_bcgs(O.VARIANTS[variant], m, n, s, O.ptr(Q), m, O.ptr(R), n, info)
    assert rc == 0
    mask = written_mask(n, s)
    assert np.all(R[~mask] == 7.25)
    assert np.all(np.tril(R, -1)[mask] == 0)


def test_variants_agree_on_well_conditioned_input():
    X = I.svd_matrix(3000, 64, 1e3, seed=8)
    Rs = [O.bcgs(v, X, 8)[1] for v in VARIANTS]
    for R in Rs[1:]:
        assert np.abs(R - Rs[0]).max() <= 1e-11 * np.abs(Rs[0]).max()


def test_thread_count_does_not_change_bits():
    X = I.svd_matrix(20000, 48, 1e4, seed=9)
    out = []
    for nt in (1, 4):
        O.set_threads(nt)
        out.append(O.bcgs("bcgsi2P1s", X, 16))
    O.set_threads(1)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


# ---------------------------------------------------------------- drop-in wrappers

@pytest.mark.parametrize("variant", VARIANTS)
def test_workspace_check(variant):
    m, n, s = 300, 20, 5
    X = I.svd_matrix(m, n, 10.0, seed=10)
    need = 5 * n * s + 9 * s * s + m * (s if variant == "bcgsi2" else n)
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        rc, Q, R = O.dropin(variant, X, s, lwork=need - 1)
    assert rc == 1
    assert np.array_equal(Q, X)
    rc, Q, R = O.dropin(variant, X, s, lwork=need)
    assert rc == 0 and I.loo(Q) < 1e-13
    rc, Q, R = O.dropin(variant, X, s, work=False)
    assert rc == 0 and I.loo(Q) < 1e-13


def test_invalid_shapes_rejected():
    X = np.zeros((10, 8), order="F")
    for s in (0, 9):
        with pytest.raises(ValueError):
            O.bcgs("bcgsi2P1s", X, s)
    with pytest.raises(ValueError):
        O.bcgs("bcgsi2P1s", np.zeros((5, 8), order="F"), 2)


# ---------------------------------------------------------------- stability

def test_kappa_behaviour():
    """BCGSI+ and P2s keep O(u) orthogonality up to kappa = 1e14; the
    Pythagorean first pass of P1s and PIPI breaks down once u*kappa^2 >> 1."""
    X = I.svd_matrix(4000, 64, 1e6, seed=1)
    for v in VARIANTS:
        Q, R, info = O.bcgs(v, X, 8)
        assert info == (0, 0)
        assert I.loo(Q) < (1e-12 if v == "bcgsi21s" else 1e-14)
    X = I.svd_matrix(4000, 64, 1e14, seed=1)
    for v in ("bcgsi2", "bcgsi2P2s"):
        Q, R, info = O.bcgs(v, X, 8)
        assert info == (0, 0) and I.loo(Q) < 1e-14
    for v in ("bcgsi2P1s", "bcgspipi2"):
        _, _, info = O.bcgs(v, X, 8)
        assert info[0] > 0


def test_glued_blocks():
    X = I.glued_matrix(4000, 64, 8, 1e-13)
    Q, R, info = O.bcgs("bcgsi2", X, 8)
    assert info == (0, 0) and I.loo(Q) < 1e-13
    _, _, info = O.bcgs("bcgsi2P1s", X, 8)
    assert info[0] == 2


@pytest.mark.parametrize("variant", ["bcgsi2P1s", "bcgspipi2", "bcgsi21s"])
def test_pythagorean_shift_closed_form(variant):
    """The optional shift c u tr(G) of the Pythagorean Cholesky steps, on an
    input whose first-pass Gram X^T X - S^T S is diag(d)^2 exactly: R's
    diagonal follows the closed form (tr(G) includes the projected part, so
    a shift taken after the subtraction would fail), and c = 0 gives |d|."""
    m, n, s, c = 3000, 40, 8, 1.0e7
    X, d = I.layered_matrix(m, n, s)
    O.set_pythagorean_shift(c)
    try:
        Q, R, info = O.bcgs(variant, X, s)
    finally:
        O.set_pythagorean_shift(0.0)
    assert info == (0, 0)
    want = I.shifted_diag(variant, X, d, s, c)
    got = np.abs(np.diag(R))
    assert np.abs(got - want).max() <= 1e-12 * want.max()
    # The shift is visible (relative effect ~1e-9 >> rounding).
    assert np.abs(got[s:] - d[s:]).max() > 1e-10
    Q0, R0, _ = O.bcgs(variant, X, s)
    assert np.abs(np.abs(np.diag(R0)) - d).max() <= 1e-13 * d.max()


def test_pythagorean_shift_avoids_breakdown():
    """Glued blocks at delta = 1e-13 break P-1S's Pythagorean Cholesky; with
    the shift 11 (m s + s (s + 1)) u tr(G) every factorization is defined."""
    m, n, s = 4000, 64, 8
    X = I.glued_matrix(m, n, s, 1e-13)
    O.set_pythagorean_shift(11.0 * (m * s + s * (s + 1)))
    try:
        Q, R, info = O.bcgs("bcgsi2P1s", X, s)
    finally:
        O.set_pythagorean_shift(0.0)
    assert info == (0, 0)
    assert np.all(np.isfinite(Q)) and np.all(np.isfinite(R))


def test_accuracy_matches_numpy():
    """Harness values on a factorization perturbed well above rounding level,
    so both LOO and the residual are exactly representable targets."""
    rng = np.random.default_rng(12)
    X = I.svd_matrix(2000, 30, 1e4, seed=12)
    Q, R, _ = O.bcgs("bcgsi2", X, 5)
    Q = np.asfortranarray(Q + 1e-7 * rng.standard_normal(Q.shape))
    loo, res, xn = O.accuracy(X, Q, R)
    assert abs(loo - I.loo(Q)) <= 1e-8 * loo
    assert abs(res - np.linalg.norm(Q @ R - X)) <= 1e-8 * res
    assert abs(xn - np.linalg.norm(X)) <= 1e-13 * xn


# ---------------------------------------------------------------- planted mistakes

def _driver_case():
    g = np.load(os.path.join(GOLDEN, "test_correct.npz"))
    return np.asfortranarray(g["X37"]), g["R37"]


def _passes_pins(X, R, Q, Rref):
    return bool(np.isfinite(R).all() and I.loo(Q) < 1e-13
                and I.rel_residual(X, Q, R) < 1e-15
                and np.abs(R - Rref).max() <= 1e-11 * np.abs(Rref).max())


@pytest.mark.parametrize("variant", VARIANTS)
def test_second_pass_removes_first_pass_errors(variant):
    """A relative error of 1e-4 injected into every first-pass projection is
    removed by the FP64 second pass: the pins still hold.  This is what makes
    the reorthogonalization terms (Y Sigma, the Pythagorean Y^T Y, the lag
    correction) visible to the planted-mistake test below; in plain FP64
    they are rounding-level corrections."""
    X, Rref = _driver_case()
    O.set_first_pass_noise(1e-4)
    try:
        Q, R, info = O.bcgs(variant, X, 5)
    finally:
        O.set_first_pass_noise(0.0)
    assert info == (0, 0)
    assert _passes_pins(X, R, Q, Rref)


@pytest.mark.parametrize("mistake", [1, 2, 3, 4, 5])
def test_planted_mistakes_are_caught(mistake):
    """Each planted mistake fails the pins for at least one variant, with
    and without the injected first-pass error."""
    X, Rref = _driver_case()
    O.plant(mistake)
    try:
        caught = False
        for eta in (0.0, 1e-4):
            O.set_first_pass_noise(eta)
            for v in VARIANTS:
                Q, R, info = O.bcgs(v, X, 5)
                caught |= not _passes_pins(X, R, Q, Rref)
        assert caught
    finally:
        O.plant(0)
        O.set_first_pass_noise(0.0)


def test_reduced_precision_first_pass_limit():
    """Column-scaled error eta in the first-pass projections (a dot product
    in reduced precision): the second pass removes it while eta kappa^2 < 1,
    but the one-sync Pythagorean Cholesky chol(X^T X - S^T S) of BCGSI+P-1S
    breaks down once eta kappa^2 >> 1, whereas BCGSI+ (intra-block QR of the
    computed first-pass vectors) keeps O(u) orthogonality.  This bounds the
    conditioning at which a low-precision first pass can match FP64."""
    X = I.svd_matrix(8000, 64, 1e4, seed=44)
    try:
        O.set_first_pass_noise(1e-10, column_scaled=True)  # eta kappa^2 = 1e-2
        Q, R, info = O.bcgs("bcgsi2P1s", X, 16)
        assert info == (0, 0) and I.loo(Q) < 1e-13
        O.set_first_pass_noise(1e-5, column_scaled=True)   # eta kappa^2 = 1e3
        Q, R, info = O.bcgs("bcgsi2P1s", X, 16)
        assert info != (0, 0) or not I.loo(Q) < 1e-10
        Q, R, info = O.bcgs("bcgsi2", X, 16)                # eta kappa = 0.1
        assert info == (0, 0) and I.loo(Q) < 1e-13
    finally:
        O.set_first_pass_noise(0.0)
