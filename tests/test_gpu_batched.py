"""GPU checks of the batched form (orth_bcgs_batched_device): every instance
of an uneven batch agrees with the CPU oracle as a single call does; a
replayed batch (same pointers, captured graph) reproduces the first run bit
for bit; the tuner's choice keeps parity; breakdown is reported per
instance."""
import numpy as np
import pytest

import bcgs_hip as B
import inputs as I
import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)


def make_batch(count, seed, kappa_hi=1e6, s_choices=(4, 8, 16), n_choices=(32, 64, 128)):
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(count):
        m = int(2 ** rng.uniform(np.log2(3000), np.log2(40000)))
        n = int(rng.choice(n_choices))
        s = int(rng.choice(s_choices))
        kappa = float(10 ** rng.uniform(2, np.log10(kappa_hi)))
        shapes.append((m, n, s, kappa))
    return shapes


class Batch:
    def __init__(self, shapes, seed):
        self.shapes = shapes
        self.X = [I.svd_matrix(m, n, kappa, seed=seed + i) for i, (m, n, s, kappa) in
                  enumerate(shapes)]
        self.x = [torch.from_numpy(np.ascontiguousarray(X.T)).cuda() for X in self.X]
        self.x0 = [t.clone() for t in self.x]
        self.r = [torch.zeros(n * n, dtype=torch.float64, device="cuda")
                  for (m, n, s, k) in shapes]

    def restore(self):
        for t, t0 in zip(self.x, self.x0):
            t.copy_(t0)
        for r in self.r:
            r.zero_()
        torch.cuda.synchronize()

    def run(self, variant, tune=False):
        ms = [m for (m, n, s, k) in self.shapes]
        ns = [n for (m, n, s, k) in self.shapes]
        ss = [s for (m, n, s, k) in self.shapes]
        return B.bcgs_batched_device(variant, ms, ns, ss, [t.data_ptr() for t in self.x], ms,
                                     [r.data_ptr() for r in self.r], ns, tune=tune)

    def result(self, i):
        m, n, s, k = self.shapes[i]
        return (self.x[i].cpu().numpy().T,
                self.r[i].cpu().numpy().reshape((n, n), order="F"))


@pytest.mark.parametrize("variant", ["bcgsi2P1s", "bcgsi2"])
def test_batch_matches_oracle(variant):
    batch = Batch(make_batch(24, 51), 51)
    say("case batch", variant)
    info, cfg = batch.run(variant)
    for i, (m, n, s, kappa) in enumerate(batch.shapes):
        Qo, Ro, info_o = O.bcgs(variant, batch.X[i], s)
        assert info[i] == info_o, (i, info[i], info_o)
        if info_o != (0, 0):
            continue
        Q, R = batch.result(i)
        assert np.abs(R - Ro).max() <= 1e-11 * np.abs(Ro).max() * max(1.0, kappa / 1e3), i
        assert I.loo(Q) <= 10 * max(I.loo(Qo), 1e-14), i
        assert I.rel_residual(batch.X[i], Q, R) <= 10 * max(I.rel_residual(batch.X[i], Qo, Ro),
                                                            1e-15), i


@pytest.mark.parametrize("variant", ["bcgsi21s", "bcgspipi2", "bcgsi2P2s"])
def test_batch_other_variants(variant):
    """The remaining variants through the batched form: 12 uneven instances
    (s = 4, 8, 16; n up to 128), each against the oracle as a single call."""
    batch = Batch(make_batch(12, 56, kappa_hi=1e4), 56)
    say("case batch", variant)
    info, cfg = batch.run(variant)
    for i, (m, n, s, kappa) in enumerate(batch.shapes):
        Qo, Ro, info_o = O.bcgs(variant, batch.X[i], s)
        assert info[i] == info_o, (i, info[i], info_o)
        if info_o != (0, 0):
            continue
        Q, R = batch.result(i)
        assert np.abs(R - Ro).max() <= 1e-11 * np.abs(Ro).max() * max(1.0, kappa / 1e3), i
        assert I.loo(Q) <= 10 * max(I.loo(Qo), 1e-14), i


def test_replay_is_bitwise_identical():
    batch = Batch(make_batch(12, 52), 52)
    say("case batch replay")
    batch.run("bcgsi2P1s")
    first = [batch.result(i) for i in range(len(batch.shapes))]
    for _ in range(2):
        batch.restore()
        batch.run("bcgsi2P1s")
        for i in range(len(batch.shapes)):
            Q, R = batch.result(i)
            assert np.array_equal(Q, first[i][0]) and np.array_equal(R, first[i][1])


def test_tuned_batch_keeps_parity():
    batch = Batch(make_batch(10, 53), 53)
    say("case batch tune")
    info, cfg = batch.run("bcgsi2P1s", tune=True)
    assert 1 <= cfg[0] <= 8 and 0 < cfg[1] <= 1 and cfg[2] in (True, False)
    batch.restore()
    info, cfg2 = batch.run("bcgsi2P1s")
    assert cfg2 == cfg
    for i, (m, n, s, kappa) in enumerate(batch.shapes):
        Qo, Ro, info_o = O.bcgs("bcgsi2P1s", batch.X[i], s)
        if info_o != (0, 0):
            continue
        Q, R = batch.result(i)
        assert np.abs(R - Ro).max() <= 1e-11 * np.abs(Ro).max() * max(1.0, kappa / 1e3), i


def test_breakdown_reported_per_instance():
    shapes = [(4000, 64, 16, 1e2), (4000, 64, 16, 1e12), (5000, 32, 8, 1e3)]
    batch = Batch(shapes, 54)
    say("case batch breakdown")
    info, _ = batch.run("bcgsi2P1s")
    for i, (m, n, s, kappa) in enumerate(shapes):
        _, _, info_o = O.bcgs("bcgsi2P1s", batch.X[i], s)
        assert (info[i][0] > 0) == (info_o[0] > 0)
    assert info[0] == (0, 0) and info[1][0] > 0


@pytest.mark.skip(reason="the batched mixed form is switched off: three runs of its tuning in the "
                         "batched benchmark ended the process with a host heap corruption whose cause "
                         "is not identified; the entry point returns 2")
def test_mixed_batch():
    shapes = [(6000, 64, 16, 1e2), (9000, 128, 16, 1e2), (3000, 32, 16, 1e2)]
    batch = Batch(shapes, 55)
    say("case batch mixed")
    x32 = [torch.from_numpy(np.ascontiguousarray(X.astype(np.float32).T)).cuda() for X in batch.X]
    ms = [m for (m, n, s, k) in shapes]
    ns = [n for (m, n, s, k) in shapes]
    info, _ = B.bcgs_batched_device("bcgsi2P1s", ms, ns, [16] * 3,
                                    [t.data_ptr() for t in batch.x], ms,
                                    [r.data_ptr() for r in batch.r], ns, precision="fp64",
                                    x32_ptrs=[t.data_ptr() for t in x32], ldx32s=ms)
    for i, (m, n, s, kappa) in enumerate(shapes):
        Xr = batch.X[i].astype(np.float32).astype(np.float64)
        Qo, Ro, _ = O.bcgs("bcgsi2P1s", Xr, 16)
        Q, R = batch.result(i)
        assert np.abs(R - Ro).max() <= 1e-11 * np.abs(Ro).max()
