"""Writes the golden fixtures under tests/golden/ (numpy + decimal only).

The reference ships no expected outputs (its test driver only prints), so
the expected factors here come from independent computations:
  * test_correct.npz: the reference driver's own inputs X35..X39
    (TESTS/test_orth.c:42-44 with the calls of :284-288: glibc rand() after
    srand(0), X[i] = rand() % (10 m n) / (10 m n + 1) for m = 550 and each
    n = 35..39) and their LAPACK Householder R factors R35..R39
    (numpy.linalg.qr) with a non-negative diagonal;
  * closed_form.npz: X = Q0 R0 with known orthonormal Q0 and upper R0, so
    the exact factors are (Q0, R0);
  * brute_force.npz: a 10 x 6 matrix whose R factor is computed by classical
    Gram-Schmidt in 60-digit decimal arithmetic.
Run:  python tests/golden/make_fixtures.py
"""
import decimal
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from inputs import test_correct_matrix  # noqa: E402


def house_r(X):
    R = np.linalg.qr(X, mode="r")
    return R * np.sign(np.diag(R))[:, None]


def exact_r(X, digits=60):
    decimal.getcontext().prec = digits
    D = decimal.Decimal
    m, n = X.shape
    cols = [[D(float(X[i, j])) for i in range(m)] for j in range(n)]
    R = [[D(0)] * n for _ in range(n)]
    Q = []
    for j in range(n):
        v = list(cols[j])
        for i, q in enumerate(Q):
            r = sum(a * b for a, b in zip(q, cols[j]))
            R[i][j] = r
            v = [a - r * b for a, b in zip(v, q)]
        # one reorthogonalization pass keeps the exact-arithmetic result
        for i, q in enumerate(Q):
            r = sum(a * b for a, b in zip(q, v))
            R[i][j] += r
            v = [a - r * b for a, b in zip(v, q)]
        nrm = sum(a * a for a in v).sqrt()
        R[j][j] = nrm
        Q.append([a / nrm for a in v])
    return np.array([[float(R[i][j]) for j in range(n)] for i in range(n)])


def main():
    arrays = {}
    for n in range(35, 40):
        X = test_correct_matrix(550, n, rank=0)
        arrays["X%d" % n] = X
        arrays["R%d" % n] = house_r(X)
    np.savez_compressed(os.path.join(HERE, "test_correct.npz"), **arrays)

    rng = np.random.default_rng(7)
    m, n = 300, 24
    Q0, _ = np.linalg.qr(rng.standard_normal((m, n)))
    R0 = np.triu(rng.standard_normal((n, n)))
    R0[np.diag_indices(n)] = np.abs(np.diag(R0)) + 1.0
    np.savez_compressed(os.path.join(HERE, "closed_form.npz"), Q0=Q0, R0=R0)

    rng = np.random.default_rng(11)
    Xb = rng.standard_normal((10, 6))
    np.savez_compressed(os.path.join(HERE, "brute_force.npz"), X=Xb, R=exact_r(Xb))


if __name__ == "__main__":
    main()
