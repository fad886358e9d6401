"""The allocation-free entry point orth_bcgs_device_noalloc against
orth_bcgs_device on the same inputs, through the C ABI.

Calls in a row per variant, each with a smaller slab than the one before
(the fixed scratch keeps the previous call's contents, so later calls check
that every buffer that must start at zero is cleared).  Tolerance: for
kappa <= 1e6 the oracle parity tests' (R to 1e-11 of its max norm, Q to
1e-10 absolute); beyond, R to kappa * 1e-15 of its max norm; LOO and the
relative residual within 10x in every case.  The two entry points run the
same kernels, and at these sizes with the same row chunks, so they normally
agree bit for bit (printed).  Also: the breakdown
(block, pivot) written to the caller's device buffer equals the status of
the regular call, and shapes beyond the fixed scratch are rejected.
"""
import sys

import numpy as np
import pytest

import bcgs_hip as B
import inputs as I

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _device(X):
    return torch.from_numpy(np.ascontiguousarray(X.T)).cuda()


def run_regular(variant, X, s):
    m, n = X.shape
    Xd = _device(X)
    Rd = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    B.bcgs_device(variant, m, n, s, Xd.data_ptr(), m, Rd.data_ptr(), n)
    B.sync()
    torch.cuda.synchronize()
    st = B.status()
    return (Xd.cpu().numpy().T, Rd.cpu().numpy().reshape((n, n), order="F"),
            (st["info_block"], st["info_pivot"]))


def run_noalloc(variant, X, s):
    m, n = X.shape
    Xd = _device(X)
    Rd = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    info = torch.full((2,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = B.bcgs_device_noalloc(variant, m, n, s, Xd.data_ptr(), m, Rd.data_ptr(), n,
                               info.data_ptr())
    torch.cuda.synchronize()
    inf = info.cpu().numpy()
    return rc, Xd.cpu().numpy().T, Rd.cpu().numpy().reshape((n, n), order="F"), (int(inf[0]),
                                                                                 int(inf[1]))


@pytest.mark.parametrize("in_pass", [False, True])
@pytest.mark.parametrize("variant", list(B.VARIANTS))
def test_noalloc_matches_device_entry(variant, in_pass, monkeypatch):
    # in_pass: the passes synchronize inside their launch also at these
    # lengths (on the fixed scratch the group sums take room from the
    # partials, so the passes use fewer row chunks).
    if in_pass:
        monkeypatch.setenv("ORTH_IN_PASS_SYNC", "1")
    # The second shape is smaller in every dimension but s.
    for m, n, s, kappa in ((120000, 256, 16, 1e6), (30001, 96, 8, 1e3), (4096, 40, 16, 1e10)):
        X = I.svd_matrix(m, n, kappa, seed=41)
        say("case noalloc", variant, (m, n, s), kappa)
        Qa, Ra, ia = run_regular(variant, X, s)
        rc, Qb, Rb, ib = run_noalloc(variant, X, s)
        assert rc == 0
        assert ib == ia
        if ia != (0, 0):
            continue
        say("  bit-identical:", bool(np.array_equal(Qa, Qb) and np.array_equal(Ra, Rb)))
        assert np.all(np.tril(Rb, -1) == 0)
        if kappa <= 1e6:
            assert np.abs(Rb - Ra).max() <= 1e-11 * np.abs(Ra).max()
            assert np.abs(Qb - Qa).max() <= 1e-10
        else:
            # Rounding differences (fewer row chunks when partials would not
            # fit) are amplified by kappa: compare to kappa u and the metrics.
            assert np.abs(Rb - Ra).max() <= kappa * 1e-15 * np.abs(Ra).max()
        la, lb = I.loo(Qa), I.loo(Qb)
        ra_, rb_ = I.rel_residual(X, Qa, Ra), I.rel_residual(X, Qb, Rb)
        assert lb <= 10 * max(la, 1e-14) and la <= 10 * max(lb, 1e-14)
        assert rb_ <= 10 * max(ra_, 1e-15) and ra_ <= 10 * max(rb_, 1e-15)


def test_noalloc_rejects_shapes_beyond_the_fixed_scratch():
    m, n, s = 2048, 520, 8
    X = torch.zeros((n, m), dtype=torch.float64, device="cuda")
    R = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    say("case noalloc reject")
    assert B.bcgs_device_noalloc("bcgsi2P1s", m, n, s, X.data_ptr(), m, R.data_ptr(), n) == 2
    # Invalid shapes are rejected as by orth_bcgs_device.
    assert B.bcgs_device_noalloc("bcgsi2P1s", 100, 200, 8, X.data_ptr(), 100, R.data_ptr(),
                                 200) == 2
