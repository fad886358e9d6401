"""GPU checks of the mixed-precision form (orth_bcgs_mixed_device).

* FP64 products on FP32-stored input: same schedule and arithmetic as the
  FP64 form, so Q and R match the CPU oracle run on the FP32-rounded input to
  the FP64 parity tolerance (1e-11 relative in R).
* Reduced precision for the first-pass projection and Gram only (the _proj
  precisions): the residual stays at FP64 level and the loss of
  orthogonality within 10x of the oracle's.
* bf16 / fp16 / bf16x3 first-pass products: the FP64 second pass restores
  orthogonality, so at kappa = 1e2 the loss of orthogonality stays within 10x
  of the oracle's FP64 value; the first-pass update error stays in the
  residual, which must sit at the level of each format's unit roundoff
  (within [0.1x, 10x] of 2^-9 for bf16, 2^-12 for fp16, 2^-18 for bf16x3,
  scaled by the measured constant 1.0), which also shows that the reduced
  precision is really used.
"""
import numpy as np
import pytest

import bcgs_hip as B
import inputs as I
import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)


def mixed(X, s, precision, variant="bcgsi2P1s", ld=None, pads=None):
    """Run the mixed form on X rounded to FP32; returns (Q, R, status).
    With ld > m, X32 and Q get leading dimension ld and NaN pad rows; the
    pad rows of both after the call are appended to pads when given."""
    m, n = X.shape
    ld = m if ld is None else ld
    X32 = np.asfortranarray(X.astype(np.float32))
    x32 = torch.full((n, ld), float("nan"), dtype=torch.float32)
    x32[:, :m] = torch.from_numpy(np.ascontiguousarray(X32.T))
    x32 = x32.cuda()
    q = torch.full((n, ld), float("nan"), dtype=torch.float64, device="cuda")
    r = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    B.bcgs_mixed_device(variant, precision, m, n, s, x32.data_ptr(), ld, q.data_ptr(), ld,
                        r.data_ptr(), n)
    B.sync()
    torch.cuda.synchronize()
    st = B.status()
    Q = q[:, :m].cpu().numpy().T
    R = r.cpu().numpy().reshape((n, n), order="F")
    if pads is not None:
        pads.append((x32[:, m:].cpu().numpy(), q[:, m:].cpu().numpy()))
    return Q, R, st, X32.astype(np.float64)


KERNELS = ["pass512", "pass32", "dbuf"]


def pick(kernel, monkeypatch):
    """pass512: the 512-thread pass with super-tiles (default below 193
    columns); pass32: the 32-row mixed pass (also the fallback for odd leading
    dimensions); dbuf: the double-buffered pass forced on every slab it
    applies to (FP64 products and reduced-precision projections; passes with
    a reduced-precision update, odd row counts and odd leading dimensions
    fall back to the 512-thread pass)."""
    if kernel == "pass32":
        monkeypatch.setenv("ORTH_MIXED_PASS32", "1")
    if kernel == "dbuf":
        monkeypatch.setenv("ORTH_PASS16D_MINCOLS", "16")


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("shape", [(20000, 128), (8192, 64), (5000, 80), (5001, 48), (300, 256)])
def test_fp64_products_match_oracle(shape, kernel, monkeypatch):
    pick(kernel, monkeypatch)
    m, n = shape
    X = I.svd_matrix(m, n, 1e4, seed=41)
    say("case mixed fp64", shape, kernel)
    Q, R, st, Xr = mixed(X, 16, "fp64")
    Qo, Ro, info = O.bcgs("bcgsi2P1s", Xr, 16)
    assert st["hip_error"] == 0 and st["info_block"] == 0 and info == (0, 0)
    assert np.abs(R - Ro).max() <= 1e-11 * np.abs(Ro).max()
    assert np.abs(Q - Qo).max() <= 1e-10
    assert np.all(np.tril(R, -1) == 0)


@pytest.mark.parametrize("case", [("bcgsi2", 5000, 80, 16), ("bcgspipi2", 6000, 64, 8),
                                  ("bcgsi2P1s", 20000, 512, 32), ("bcgsi2P2s", 4000, 40, 10)])
def test_fp64_products_other_variants_and_shapes(case):
    """FP64 products on FP32 storage outside the fused mixed pass (other
    variants, s != 16, n > 256): the input is widened into Q and the FP64
    form runs; Q and R match the oracle on the rounded input."""
    variant, m, n, s = case
    X = I.svd_matrix(m, n, 1e4, seed=43)
    say("case mixed fp64 widened", case)
    Q, R, st, Xr = mixed(X, s, "fp64", variant)
    Qo, Ro, info = O.bcgs(variant, Xr, s)
    assert st["hip_error"] == 0 and (st["info_block"], st["info_pivot"]) == info == (0, 0)
    assert np.abs(R - Ro).max() <= 1e-11 * np.abs(Ro).max()
    assert np.abs(Q - Qo).max() <= 1e-10


LEVEL = {"bf16": 2.0 ** -9, "fp16": 2.0 ** -12, "bf16x3": 2.0 ** -18}


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("precision", ["bf16", "fp16", "bf16x3"])
def test_reduced_precision_first_pass(precision, kernel, monkeypatch):
    pick(kernel, monkeypatch)
    m, n = 65536, 128
    X = I.svd_matrix(m, n, 1e2, seed=42)
    say("case mixed", precision, kernel)
    Q, R, st, Xr = mixed(X, 16, precision)
    assert st["hip_error"] == 0 and st["info_block"] == 0
    Qo, Ro, _ = O.bcgs("bcgsi2P1s", Xr, 16)
    loo, loo_o = I.loo(Q), I.loo(Qo)
    assert loo <= 10 * max(loo_o, 1e-14), (loo, loo_o)
    res = I.rel_residual(Xr, Q, R)
    lvl = LEVEL[precision]
    assert 0.1 * lvl <= res <= 10 * lvl, (res, lvl)
    # R's upper triangle stays close to the FP64 factorization at that level.
    assert np.abs(R - Ro).max() <= 100 * lvl * np.abs(Ro).max()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("precision", ["bf16_proj", "fp16_proj", "bf16x3_proj"])
@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_reduced_precision_projection_only(kappa, precision, kernel, monkeypatch):
    """Reduced precision for the first-pass projection and Gram only: the
    first-pass update stays FP64, so X = QR holds to FP64 rounding of the
    FP32-rounded input (R carries the first pass's projection as computed
    and the reorthogonalization's correction), and the FP64 second pass
    keeps the loss of orthogonality within 10x of the oracle's FP64 value.
    The Pythagorean first-pass factor chol(X^T X - S^T S) cancels to
    sigma_min^2 / sigma_max^2 of the block, which a reduced-precision S
    (bf16, fp16, even bf16x3) cannot resolve at kappa = 1e6: there a
    breakdown may be reported (and Q is not trusted)."""
    pick(kernel, monkeypatch)
    m, n = 65536, 128
    X = I.svd_matrix(m, n, kappa, seed=45)
    say("case mixed projection-only", precision, kernel, kappa)
    Q, R, st, Xr = mixed(X, 16, precision)
    assert st["hip_error"] == 0
    if kappa > 1e4 and st["info_block"] != 0:
        return
    assert st["info_block"] == 0
    Qo, Ro, _ = O.bcgs("bcgsi2P1s", Xr, 16)
    loo, loo_o = I.loo(Q), I.loo(Qo)
    assert loo <= 10 * max(loo_o, 1e-14), (loo, loo_o)
    res = I.rel_residual(Xr, Q, R)
    assert res <= 1e-14, res
    assert np.all(np.tril(R, -1) == 0)


@pytest.mark.parametrize("precision", ["fp64", "bf16_proj"])
@pytest.mark.parametrize("shape", [(70000, 256), (33334, 208), (8192, 96)])
def test_double_buffered_mixed_matches_512_thread_pass(shape, precision, monkeypatch):
    """The double-buffered mixed pass (forced on every slab, including the
    intra-block QR's Gram of raw columns with no pending group) against the
    512-thread mixed pass on the same input: the reduced-precision tile
    products are the same 32-row MFMAs in both, so R agrees to the FP64
    rounding of their different summation orders."""
    m, n = shape
    X = I.svd_matrix(m, n, 1e4 if precision == "fp64" else 1e2, seed=46)
    say("case mixed dbuf vs 512", shape, precision)
    monkeypatch.setenv("ORTH_PASS16D_MINCOLS", "16")
    Qa, Ra, sta, Xr = mixed(X, 16, precision)
    monkeypatch.setenv("ORTH_PASS_REGPF", "1")
    Qb, Rb, stb, _ = mixed(X, 16, precision)
    assert sta["hip_error"] == 0 and stb["hip_error"] == 0
    assert sta["info_block"] == 0 and stb["info_block"] == 0
    assert np.abs(Ra - Rb).max() <= 1e-11 * np.abs(Rb).max()
    assert np.abs(Qa - Qb).max() <= 1e-10


@pytest.mark.parametrize("precision", ["fp64", "bf16"])
@pytest.mark.parametrize("shape", [(5001, 48, 5004), (4099, 128, 4100), (301, 256, 304),
                                   (20003, 64, 20004)])
def test_padded_leading_dimension(shape, precision):
    """Row counts that are not a multiple of 4 inside a leading dimension
    that is: the 512-thread pass moves raw FP32 columns four rows per lane,
    so the last group straddles the end of the rows.  Results match the
    oracle (FP64 products) or keep the reduced-precision bounds (bf16), and
    the NaN pad rows of X32 and Q are neither read into the result nor
    written."""
    m, n, ld = shape
    X = I.svd_matrix(m, n, 1e2, seed=44)
    say("case mixed padded", shape, precision)
    pads = []
    Q, R, st, Xr = mixed(X, 16, precision, ld=ld, pads=pads)
    assert st["hip_error"] == 0 and st["info_block"] == 0
    assert np.all(np.isnan(pads[0][0])) and np.all(np.isnan(pads[0][1]))
    assert np.all(np.isfinite(Q)) and np.all(np.isfinite(R))
    Qo, Ro, info = O.bcgs("bcgsi2P1s", Xr, 16)
    if precision == "fp64":
        assert np.abs(R - Ro).max() <= 1e-11 * np.abs(Ro).max()
        assert np.abs(Q - Qo).max() <= 1e-10
    else:
        assert I.loo(Q) <= 10 * max(I.loo(Qo), 1e-14)
        assert I.rel_residual(Xr, Q, R) <= 10 * LEVEL[precision]


def test_mixed_rejects_unsupported():
    x = torch.zeros((32, 4096), dtype=torch.float32, device="cuda")
    q = torch.zeros((32, 4096), dtype=torch.float64, device="cuda")
    r = torch.zeros(32 * 32, dtype=torch.float64, device="cuda")
    for v, s in [("bcgsi2", 16), ("bcgsi2P1s", 8)]:
        with pytest.raises(RuntimeError):
            B.bcgs_mixed_device(v, "bf16", 4096, 32, s, x.data_ptr(), 4096, q.data_ptr(), 4096,
                                r.data_ptr(), 32)
