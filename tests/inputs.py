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
