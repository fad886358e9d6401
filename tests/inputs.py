"""Seeded test inputs (numpy, FP64, column-major)."""
import ctypes

import numpy as np


def orthonormal(rng, m, n):
    """Q-factor of an m x n standard Gaussian (LAPACK Householder)."""
    Q, R = np.linalg.qr(rng.standard_normal((m, n)))
    return Q * np.sign(np.diag(R))


def svd_matrix(m, n, kappa, seed=1):
    """X = U diag(sigma) V^T, sigma log-spaced from 1 to 1/kappa."""
    rng = np.random.default_rng(seed)
    U = orthonormal(rng, m, n)
    V = orthonormal(rng, n, n)
    sig = np.logspace(0.0, -np.log10(kappa), n)
    return np.asfortranarray((U * sig) @ V.T)


def glued_matrix(m, n, s, delta, seed=3):
    """Blocks X_k = Q0 C_k + delta E_k with one fixed orthonormal Q0 (m x s)."""
    rng = np.random.default_rng(seed)
    Q0 = orthonormal(rng, m, s)
    blocks = []
    for k in range((n + s - 1) // s):
        w = min(s, n - k * s)
        C = rng.standard_normal((s, w))
        E = rng.standard_normal((m, w))
        blocks.append(Q0 @ C + delta * E)
    return np.asfortranarray(np.hstack(blocks))


def layered_matrix(m, n, s, seed=5):
    """X with known QR: block b = Q0[:, :bs] A_b + Q0[:, bs:bs+s] diag(d_b),
    d in [1, 2], A_b standard Gaussian (so X^T X - S^T S = diag(d_b)^2 while
    the Gram of the block has trace ||A_b||_F^2 + sum d_b^2).  Returns
    (X, d)."""
    rng = np.random.default_rng(seed)
    Q0 = orthonormal(rng, m, n)
    d = 1.0 + rng.random(n)
    X = Q0 * d
    for b in range(1, (n + s - 1) // s):
        c0, c1 = b * s, min(n, (b + 1) * s)
        X[:, c0:c1] += Q0[:, :c0] @ rng.standard_normal((c0, c1 - c0))
    return np.asfortranarray(X), d


def shifted_diag(variant, X, d, s, c):
    """Diagonal of R that the shifted Pythagorean steps give for a
    layered_matrix input (blocks b >= 1), in closed form: with
    sigma = c u tr(X_b^T X_b) and sigma2 = c u tr(first-pass Gram of the
    normalized block), P-1S and PIPI+ give sqrt(d^2 + sigma2 (d^2 + sigma)),
    A-1S (no first-pass normalization) sqrt(d^2 + c u sum d^2)."""
    u = 2.0 ** -53
    n = X.shape[1]
    out = np.abs(d).copy()
    for b in range(1, (n + s - 1) // s):
        c0, c1 = b * s, min(n, (b + 1) * s)
        db = d[c0:c1]
        if variant == "bcgsi21s":
            out[c0:c1] = np.sqrt(db ** 2 + c * u * np.sum(db ** 2))
            continue
        sig = c * u * np.sum(X[:, c0:c1] ** 2)
        sig2 = c * u * np.sum(db ** 2 / (db ** 2 + sig))
        out[c0:c1] = np.sqrt(db ** 2 + sig2 * (db ** 2 + sig))
    return out


def test_correct_matrix(m, n, rank=0):
    """Input of the reference's test driver (TESTS/test_orth.c:42-44) at one
    rank: srand(rank); X[i] = rand() % (10 m n) / (10 m n + 1), filled in
    column-major order.  Needs glibc's rand()."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(rank)
    mod = 10 * m * n
    vals = np.fromiter((libc.rand() % mod for _ in range(m * n)), dtype=np.float64,
                       count=m * n)
    return np.asfortranarray((vals / float(mod + 1)).reshape((n, m)).T)


def loo(Q):
    n = Q.shape[1]
    return float(np.linalg.norm(Q.T @ Q - np.eye(n)))


def rel_residual(X, Q, R):
    return float(np.linalg.norm(Q @ R - X) / np.linalg.norm(X))
