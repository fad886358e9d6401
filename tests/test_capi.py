"""C-ABI checks that need no GPU: liborth_hip.so loads, exports every
function declared in include/*.h, and rejects bad arguments before touching
the device (workspace rule of SRC/orth.c:82-86, invalid shapes)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import bcgs_hip as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("orth.h", "orth_ext.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|double)\s+\**(\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_symbols():
    names = declared_functions()
    for ref in ("bcgsi2P1s", "bcgsi21s", "bcgspipi2", "bcgsi2", "bcgsi2P2s"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(B.LIB_PATH), "build liborth_hip.so first"
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    L = B.lib()
    for n in declared_functions():
        assert getattr(L, n) is not None


def test_reference_signatures_are_c_linkage():
    L = ctypes.CDLL(B.LIB_PATH)
    for n in B.VARIANTS:
        assert hasattr(L, n)


@pytest.mark.parametrize("variant", list(B.VARIANTS))
def test_small_workspace_returns_1_without_device(variant):
    m, n, s = 200, 20, 5
    X = np.random.default_rng(0).standard_normal((m, n))
    need = 5 * n * s + 9 * s * s + m * (s if variant == "bcgsi2" else n)
    rc, Q, R = B.dropin(variant, X, s, lwork=need - 1, work=True)
    assert rc == 1
    assert np.array_equal(Q, X) and not R.any()


@pytest.mark.parametrize("shape", [(10, 8, 0), (10, 8, 9), (5, 8, 2), (100, 80, 40)])
def test_invalid_shape_returns_2_without_device(shape):
    m, n, s = shape
    X = np.zeros((m, n))
    rc, Q, R = B.dropin("bcgsi2P1s", X, s)
    assert rc == 2


def test_no_cpu_fallback_when_library_missing(tmp_path):
    code = ("import sys; sys.path.insert(0, %r); import bcgs_hip as B; "
            "B.LIB_PATH = %r\n"
            "try:\n    B.lib()\nexcept RuntimeError as e:\n    print('raised', e)\n"
            % (os.path.dirname(B.__file__), str(tmp_path / "nope.so")))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True,
                         text=True).stdout
    assert out.startswith("raised")


@pytest.mark.parametrize("args", [
    ("bcgsi2", "bf16", 4096, 32, 16),      # only BCGSI+P-1S has a mixed form
    ("bcgsi2P1s", "bf16", 4096, 32, 8),    # block width 16 only
    ("bcgsi2P1s", "bf16", 4096, 272, 16),  # n <= 256
])
def test_mixed_rejects_unsupported_without_device(args):
    v, prec, m, n, s = args
    L = B.lib()
    rc = L.orth_bcgs_mixed_device(B.VARIANTS[v], B.PRECISIONS[prec], m, n, s,
                                  ctypes.c_void_p(16), m, ctypes.c_void_p(16), m,
                                  ctypes.c_void_p(16), n, None)
    assert rc == 2
    rc = L.orth_bcgs_mixed_device(0, 7, m, 32, 16, ctypes.c_void_p(16), m,
                                  ctypes.c_void_p(16), m, ctypes.c_void_p(16), 32, None)
    assert rc == 2


def test_batched_rejects_invalid_without_device():
    i64 = ctypes.c_int64
    vp = ctypes.c_void_p
    L = B.lib()
    arr = lambda t, v: (t * len(v))(*v)
    # s > n for the second instance
    rc = L.orth_bcgs_batched_device(0, 2, arr(i64, [100, 100]), arr(i64, [16, 8]),
                                    arr(i64, [4, 16]), arr(vp, [16, 16]), arr(i64, [100, 100]),
                                    arr(vp, [16, 16]), arr(i64, [16, 16]), None, 0, None, None)
    assert rc == 2
    rc = L.orth_bcgs_batched_device(0, 0, arr(i64, [1]), arr(i64, [1]), arr(i64, [1]),
                                    arr(vp, [16]), arr(i64, [1]), arr(vp, [16]), arr(i64, [1]),
                                    None, 0, None, None)
    assert rc == 2


def test_pythagorean_shift_argument_checked_without_device():
    L = B.lib()
    for bad in (-1.0, float("nan"), float("inf")):
        assert L.orth_set_pythagorean_shift(bad) == 2
    assert L.orth_set_pythagorean_shift(0.0) == 0


def test_batched_mixed_rejected_without_device():
    """The batched mixed form is switched off: every precision returns 2
    before touching the device."""
    i64 = ctypes.c_int64
    vp = ctypes.c_void_p
    L = B.lib()
    arr = lambda t, v: (t * len(v))(*v)
    for prec in ("fp64", "bf16", "fp16", "bf16x3"):
        rc = L.orth_bcgs_batched_mixed_device(
            0, B.PRECISIONS[prec], 1, arr(i64, [4096]), arr(i64, [32]), arr(i64, [16]),
            arr(vp, [16]), arr(i64, [4096]), arr(vp, [16]), arr(i64, [4096]), arr(vp, [16]),
            arr(i64, [32]), None, 0, None, None)
        assert rc == 2

