"""Link-level drop-in: a plain C caller (tests/c_driver/dropin_driver.c)
compiled with gcc against include/orth.h and linked with -lorth_hip, the
way C host code links the reference's archive (include/orth.h:8-17).  It
follows the reference driver's protocol (TESTS/test_orth.c:12-60): X from
srand(0) / rand(), Q = X, R from calloc, each variant with a host work
buffer, then ||Q^T Q - I||_F and ||Q R - X||_F; a workspace one short of
SRC/orth.c:82-86's requirement must return 1 and write nothing.  On the
device the R factors must match the golden Householder R factors of the
same inputs (tests/golden/test_correct.npz)."""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "c_driver"))
import make_driver as driver_build  # noqa: E402  tests/c_driver/make_driver.py

VARIANTS = ["bcgsi2P1s", "bcgsi21s", "bcgspipi2", "bcgsi2", "bcgsi2P2s"]


def driver():
    if not os.path.exists(driver_build.OUT):
        driver_build.build()
    return driver_build.OUT


def run(args):
    r = subprocess.run([driver()] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=120)
    fields = {}
    for tok in r.stdout.split():
        if "=" in tok:
            k, v = tok.split("=", 1)
            fields[k] = float(v)
    return r.returncode, fields


def test_driver_links_and_checks_workspace_without_device():
    """Needs no GPU: the workspace check returns 1 before the device is
    touched, through the linked C symbols."""
    rc, f = run([550, 35, 5, "bcgsi2P1s"])
    assert rc == 0, rc
    assert f["small_rc"] == 1


@pytest.mark.gpu
@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("n", [35, 36, 37, 38, 39])
def test_c_caller_matches_golden(variant, n, tmp_path):
    out = tmp_path / "R.bin"
    print("case c driver", variant, n, flush=True)
    rc, f = run([550, n, 5, variant, out])
    assert rc == 0
    assert f["rc"] == 0 and f["small_rc"] == 1
    assert f["loo"] < 1e-13
    assert f["resid"] / f["xnorm"] < 1e-15
    R = np.fromfile(out, dtype=np.float64).reshape((n, n), order="F")
    g = np.load(os.path.join(HERE, "golden", "test_correct.npz"))
    Rref = g["R%d" % n]
    assert np.abs(R - Rref).max() <= 1e-11 * np.abs(Rref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("variant", VARIANTS)
def test_c_caller_larger_slab(variant):
    print("case c driver 20000x64 s=16", variant, flush=True)
    rc, f = run([20000, 64, 16, variant])
    assert rc == 0 and f["rc"] == 0 and f["small_rc"] == 1
    assert f["loo"] < 1e-13
    assert f["resid"] / f["xnorm"] < 1e-15
