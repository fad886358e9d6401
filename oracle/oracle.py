"""ctypes loader for the CPU oracle (oracle/bcgs_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the benchmark's CPU
baseline leg use this module as the checker.  The product library never
imports it.

All matrices are numpy float64 arrays in Fortran (column-major) order, the
layout of the reference's SRC/orth.c.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

# Variant ids shared with the device library's extended API.
VARIANTS = {
    "bcgsi2P1s": 0,
    "bcgsi21s": 1,
    "bcgspipi2": 2,
    "bcgsi2": 3,
    "bcgsi2P2s": 4,
}

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_bcgs.restype = ctypes.c_int
        L.oracle_bcgs.argtypes = [ctypes.c_int, _i64, _i64, _i64, _dp, _i64,
                                  _dp, _i64, ctypes.POINTER(ctypes.c_int)]
        for name in VARIANTS:
            f = getattr(L, "oracle_" + name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                          ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int]
        L.oracle_houseqr.restype = None
        L.oracle_houseqr.argtypes = [_i64, _i64, _dp, _i64, _dp, _i64]
        L.oracle_accuracy.restype = None
        L.oracle_accuracy.argtypes = [_i64, _i64, _dp, _i64, _dp, _i64, _dp,
                                      _i64, _dp]
        L.oracle_gram.restype = None
        L.oracle_gram.argtypes = [_i64, _i64, _dp, _i64, _i64, _dp, _i64,
                                  _dp, _i64]
        L.oracle_potrf.restype = ctypes.c_int
        L.oracle_potrf.argtypes = [_i64, _dp, _i64]
        L.oracle_update.restype = None
        L.oracle_update.argtypes = [_i64, _i64, _dp, _i64, _i64, _dp, _i64,
                                    _dp, _i64, _dp, _i64]
        L.oracle_set_first_pass_noise.restype = None
        L.oracle_set_first_pass_noise.argtypes = [ctypes.c_double]
        L.oracle_set_noise_mode.restype = None
        L.oracle_set_noise_mode.argtypes = [ctypes.c_int]
        L.oracle_plant.restype = None
        L.oracle_plant.argtypes = [ctypes.c_int]
        L.oracle_set_pythagorean_shift.restype = None
        L.oracle_set_pythagorean_shift.argtypes = [ctypes.c_double]
        L.oracle_set_threads.restype = None
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def plant(mistake):
    """Plant a numbered mistake in the oracle (0 removes it)."""
    lib().oracle_plant(int(mistake))


def set_first_pass_noise(eta, column_scaled=False):
    """Perturb every first-pass projection by eta (0 removes it): relative to
    each entry, or (column_scaled) relative to the column norm, the error of
    a dot product computed in reduced precision."""
    lib().oracle_set_noise_mode(1 if column_scaled else 0)
    lib().oracle_set_first_pass_noise(float(eta))


def set_pythagorean_shift(c):
    """Shift the Pythagorean Cholesky steps by c u tr(G) (0 removes it)."""
    lib().oracle_set_pythagorean_shift(float(c))


def set_threads(n):
    lib().oracle_set_threads(int(n))


def _fortran(a):
    a = np.asarray(a, dtype=np.float64)
    return np.asfortranarray(a)


def bcgs(variant, X, s):
    """Orthogonalize a copy of X (m x n) with the named variant.

    Returns (Q, R, info) where info = (block, pivot) of the first Cholesky
    breakdown (0, 0 when none)."""
    Q = _fortran(X).copy(order="F")
    m, n = Q.shape
    R = np.zeros((n, n), order="F")
    info = (ctypes.c_int * 2)()
    rc = lib().oracle_bcgs(VARIANTS[variant], m, n, s, ptr(Q), m, ptr(R), n,
                           info)
    if rc != 0:
        raise ValueError("oracle_bcgs rejected the shape (rc=%d)" % rc)
    return Q, R, (info[0], info[1])


def dropin(variant, X, s, lwork=None, work=True):
    """Call the reference-signature wrapper; returns (rc, Q, R)."""
    Q = _fortran(X).copy(order="F")
    m, n = Q.shape
    R = np.zeros((n, n), order="F")
    need = 5 * n * s + 9 * s * s + m * (s if variant == "bcgsi2" else n)
    if lwork is None:
        lwork = need
    wbuf = np.zeros(max(int(lwork), 1)) if work else None
    rc = getattr(lib(), "oracle_" + variant)(m, n, s, ptr(Q), m, ptr(R), n,
                                             ptr(wbuf), int(lwork))
    return rc, Q, R


def houseqr(A):
    Q = _fortran(A).copy(order="F")
    m, w = Q.shape
    R = np.zeros((w, w), order="F")
    lib().oracle_houseqr(m, w, ptr(Q), m, ptr(R), w)
    return Q, R


def accuracy(X, Q, R):
    """(||Q^T Q - I||_F, ||Q R - X||_F, ||X||_F), SRC/util.c:36-75."""
    X = _fortran(X)
    Q = _fortran(Q)
    R = _fortran(R)
    m, n = X.shape
    out = np.zeros(3)
    lib().oracle_accuracy(m, n, ptr(X), m, ptr(Q), m, ptr(R), R.shape[0],
                          ptr(out))
    return float(out[0]), float(out[1]), float(out[2])


def gram(A, B):
    A = _fortran(A)
    B = _fortran(B)
    C = np.zeros((A.shape[1], B.shape[1]), order="F")
    lib().oracle_gram(A.shape[0], A.shape[1], ptr(A), A.shape[0], B.shape[1],
                      ptr(B), B.shape[0], ptr(C), C.shape[0])
    return C


def potrf(G):
    U = _fortran(G).copy(order="F")
    info = lib().oracle_potrf(U.shape[0], ptr(U), U.shape[0])
    return U, info


def update(X, Q, W, T=None):
    """X <- (X - Q W) T^{-1} (T upper, optional)."""
    Xo = _fortran(X).copy(order="F")
    Q = _fortran(Q)
    W = _fortran(W)
    Tp = None
    ldT = 1
    if T is not None:
        T = _fortran(T)
        Tp = ptr(T)
        ldT = T.shape[0]
    lib().oracle_update(Xo.shape[0], Xo.shape[1], ptr(Xo), Xo.shape[0],
                        Q.shape[1], ptr(Q), Q.shape[0], ptr(W), W.shape[0],
                        Tp, ldT)
    return Xo
