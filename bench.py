#!/usr/bin/env python3
"""Benchmark of the BCGS row-slab orthogonalization on one MI355X.

Workload (one step = one full orthogonalization, inputs resident in HBM):
BCGSI+P-1S in FP64 on a row slab X of m = 2^24 rows, n = 256 columns,
block width s = 16 (16 blocks), X = U diag(sigma) V^T with sigma log-spaced
over kappa = 1e6 (seeded, generated on the device).

Prints one JSON line.  `value` is the achieved HBM bandwidth on the
algorithmic byte count B_alg = 8 m (2n + s N(N-1)/2) of one factorization
(read X once, write Q once, re-read the Q prefix once per block) divided by
the HIP-event time of the call.  `roofline` describes the dominant kernel
class, timed with HIP events around each of its launches on the stream it
runs on; `cpu_baseline` times the CPU oracle (a C restatement of the
reference) on the same input on the host cores (all m rows by default).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("one-synchronization-bcgs_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "BCGSI+P-1S ms & HBM GB/s, m=2^24 n=256 s=16: mixed vs FP64 at equal LOO"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def finite(v):
    """JSON-safe copy: NaN / inf (a reported breakdown) become null."""
    if isinstance(v, float):
        return v if v == v and abs(v) != float("inf") else None
    if isinstance(v, dict):
        return {k: finite(x) for k, x in v.items()}
    if isinstance(v, list):
        return [finite(x) for x in v]
    return v


def b_alg(m, n, s, elem=8):
    N = 1 + (n - s + s - 1) // s
    return elem * m * (2 * n + s * N * (N - 1) // 2)


def b_alg_mixed(m, n, s):
    """Algorithmic bytes of the mixed form: read X once in FP32, write Q once
    and re-read the Q prefix once per block in FP64."""
    N = 1 + (n - s + s - 1) // s
    return 4 * m * n + 8 * m * (n + s * N * (N - 1) // 2)


def f_alg(m, n):
    return 4.0 * m * n * n


def make_input(torch, m, n, kappa, seed):
    """X^T as an n x m row-major tensor (= column-major m x n, ld m)."""
    import device_inputs

    return device_inputs.svd_input(torch, m, n, kappa, seed)


def host_cpu():
    """(model name, logical CPUs visible) of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    return model, ncpu


def cpu_baseline(X0, variant, n, s, rows, threads, m_full):
    """Time the CPU oracle on the first `rows` rows of the device input X0
    (the same bits the GPU factorizes); scaled linearly in m when rows < m."""
    import inputs
    import oracle

    oracle.set_threads(threads)
    X = X0[:, :rows].cpu().numpy().T  # rows x n, column-major
    t0 = time.perf_counter()
    Q, R, info = oracle.bcgs(variant, X, s)
    t = time.perf_counter() - t0
    loo = inputs.loo(Q) if rows <= (1 << 22) else None
    del Q, R, X
    t_full = t * (m_full / rows)
    model, ncpu = host_cpu()
    scaled = "" if rows == m_full else ", scaled linearly in m to m=%d" % m_full
    return {
        "value": round(b_alg(m_full, n, s) / t_full / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": ("CPU oracle (C restatement of SRC/orth.c, OpenMP, %d threads) %s FP64 on "
                   "rows 0..%d of the same device input (n=%d, s=%d), %.2f s%s: %.1f ms per "
                   "factorization%s" % (
                       threads, variant, rows, n, s, t, scaled, t_full * 1e3,
                       "" if loo is None else "; sample LOO %.1e" % loo)),
        "ms_per_factorization": round(t_full * 1e3, 1),
        "host": {"cpu_model": model, "logical_cpus_visible": ncpu, "threads_used": threads},
    }


def pass_traffic(variant, m, n, s):
    """HBM bytes per streaming-pass launch from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE + WRITE_SIZE, scaled by the
    same counters read on a plain copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "pass_traffic_%s_m%d_n%d_s%d.md" % (variant, m, n, s))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def make_glued(torch, m, n, s, delta, seed):
    """Blocks X_k = Q0 C_k + delta E_k (one fixed orthonormal m x s basis Q0),
    as an n x m row-major tensor (= column-major m x n)."""
    import device_inputs

    return device_inputs.glued_input(torch, m, n, s, delta, seed)


def run_one(torch, B, variant, X0, m, n, s, reps=3):
    """Time `reps` factorizations of X0 (restored before each, outside the
    events); returns (ms, status, loo, rel_residual) of the last one."""
    X = torch.empty_like(X0)
    R = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream()
    times = []
    for k in range(reps + 1):
        X.copy_(X0)
        R.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        B.bcgs_device(variant, m, n, s, X.data_ptr(), m, R.data_ptr(), n, st.cuda_stream)
        e1.record(st)
        torch.cuda.synchronize()
        if k:
            times.append(e0.elapsed_time(e1))
    status = B.status()
    loo, res, xn = B.accuracy_device(m, n, X0.data_ptr(), m, X.data_ptr(), m, R.data_ptr(), n,
                                     st.cuda_stream)
    del X, R
    return sorted(times)[len(times) // 2], status, loo, res / xn


def run_mixed(torch, B, prec, X0, X32, m, n, s, reps=2, variant="bcgsi2P1s"):
    """Time `reps` mixed-form factorizations of the FP32 input X32; returns
    (ms, status, loo, rel_residual) of the last one, measured against the
    FP64 input X0."""
    Q = torch.empty_like(X0)
    R = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream()
    times = []
    for k in range(reps + 1):
        R.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        B.bcgs_mixed_device(variant, prec, m, n, s, X32.data_ptr(), m, Q.data_ptr(), m,
                            R.data_ptr(), n, st.cuda_stream)
        e1.record(st)
        torch.cuda.synchronize()
        if k:
            times.append(e0.elapsed_time(e1))
    status = B.status()
    loo, res, xn = B.accuracy_device(m, n, X0.data_ptr(), m, Q.data_ptr(), m, R.data_ptr(), n,
                                     st.cuda_stream)
    del Q, R
    return sorted(times)[len(times) // 2], status, loo, res / xn


def suite(args):
    """The measurement configurations beside the headline line: every variant
    at m=2^24, n=256, s=16 for kappa in {1e2, 1e6, 1e10}; the mixed form at
    the same shape for kappa in {1e2, 1e6, 1e10, 1e14} and each first-pass
    precision, beside FP64 BCGSI+P-1S on the same input; the reference's CPU
    shape (m=1e5, n=200, s=10, kappa=1e8) on the GPU beside the CPU oracle on
    the same input; and glued near-rank-deficient blocks at m=2^22, n=512,
    s=32 for delta in {1e-6, 1e-10, 1e-13}.  One JSON object per case on
    stdout (this mode prints no headline line)."""
    import numpy as np
    import torch

    import bcgs_hip as B
    import inputs
    import oracle

    B.lib()
    variants = list(B.VARIANTS)
    out = []

    def emit(d):
        out.append(d)
        print(json.dumps(finite(d)), flush=True)

    # Reference CPU shape: GPU and CPU oracle on identical bits.
    m, n, s = 100000, 200, 10
    X0 = make_input(torch, m, n, 1e8, seed=2)
    Xh = X0.cpu().numpy().T
    oracle.set_threads(args.cpu_threads)
    for v in variants:
        ms, st, loo, res = run_one(torch, B, v, X0, m, n, s)
        t0 = time.perf_counter()
        Qo, Ro, info = oracle.bcgs(v, Xh, s)
        tc = time.perf_counter() - t0
        emit({"config": "cpu-shape", "variant": v, "m": m, "n": n, "s": s, "kappa": 1e8,
              "gpu_ms": round(ms, 3), "gpu_loo": loo, "gpu_rel_residual": res,
              "gpu_info_block": st["info_block"], "cpu_ms": round(tc * 1e3, 1),
              "cpu_threads": args.cpu_threads, "cpu_loo": inputs.loo(Qo),
              "cpu_rel_residual": inputs.rel_residual(Xh, Qo, Ro), "cpu_info_block": info[0]})
    del X0

    # North-star shape, every variant, three conditionings; the mixed form
    # (BCGSI+P-1S on the FP32-rounded input, four first-pass precisions)
    # beside FP64 BCGSI+P-1S on the same input, at four conditionings.
    m, n, s = 1 << 24, 256, 16
    for kappa in (1e2, 1e6, 1e10, 1e14):
        X0 = make_input(torch, m, n, kappa, seed=1)
        fp64_p1s = None
        for v in (variants if kappa <= 1e10 else ["bcgsi2P1s"]):
            ms, st, loo, res = run_one(torch, B, v, X0, m, n, s)
            if v == "bcgsi2P1s":
                fp64_p1s = (ms, loo)
            if kappa <= 1e10:
                emit({"config": "fp64", "variant": v, "m": m, "n": n, "s": s, "kappa": kappa,
                      "ms": round(ms, 3), "GBs_alg": round(b_alg(m, n, s) / (ms / 1e3) / 1e9, 1),
                      "loo": loo, "rel_residual": res, "info_block": st["info_block"],
                      "passes": st["passes"], "syncs": st["syncs"]})
        X32 = X0.float()
        for prec in ("fp64", "bf16", "fp16", "bf16x3", "bf16_proj", "fp16_proj", "bf16x3_proj"):
            ms, st, loo, res = run_mixed(torch, B, prec, X0, X32, m, n, s)
            emit({"config": "mixed", "precision": prec, "m": m, "n": n, "s": s,
                  "kappa": kappa, "ms": round(ms, 3),
                  "GBs_alg": round(b_alg_mixed(m, n, s) / (ms / 1e3) / 1e9, 1),
                  "loo": loo, "rel_residual": res, "info_block": st["info_block"],
                  "fp64_ms": round(fp64_p1s[0], 3), "fp64_loo": fp64_p1s[1],
                  "speedup_vs_fp64": round(fp64_p1s[0] / ms, 3)})
        del X0, X32
        torch.cuda.empty_cache()

    # Glued stress.
    m, n, s = 1 << 22, 512, 32
    for delta in (1e-6, 1e-10, 1e-13):
        X0 = make_glued(torch, m, n, s, delta, seed=3)
        for v in variants:
            ms, st, loo, res = run_one(torch, B, v, X0, m, n, s, reps=1)
            emit({"config": "glued", "variant": v, "m": m, "n": n, "s": s, "delta": delta,
                  "ms": round(ms, 3), "GBs_alg": round(b_alg(m, n, s) / (ms / 1e3) / 1e9, 1),
                  "loo": loo, "rel_residual": res, "info_block": st["info_block"],
                  "info_pivot": st["info_pivot"]})
        # The Pythagorean variants again with the optional Cholesky shift
        # c u tr(G), c = 11 (m s + s (s + 1)).
        pc = 11.0 * (m * s + s * (s + 1))
        B.set_pythagorean_shift(pc)
        for v in ("bcgsi2P1s", "bcgspipi2", "bcgsi21s", "bcgsi2P2s"):
            ms, st, loo, res = run_one(torch, B, v, X0, m, n, s, reps=1)
            emit({"config": "glued-shift", "variant": v, "m": m, "n": n, "s": s,
                  "delta": delta, "shift_c": pc, "ms": round(ms, 3), "loo": loo,
                  "rel_residual": res, "info_block": st["info_block"],
                  "info_pivot": st["info_pivot"]})
        B.set_pythagorean_shift(0.0)
        # Mixed form with FP64 products on the FP32-rounded input (widened
        # into Q, then the FP64 form: n = 512 is outside the fused mixed pass).
        X32 = X0.float()
        for v in ("bcgsi2P1s", "bcgsi2"):
            ms, st, loo, res = run_mixed(torch, B, "fp64", X0, X32, m, n, s, reps=1, variant=v)
            emit({"config": "glued-mixed", "variant": v, "precision": "fp64", "m": m, "n": n,
                  "s": s, "delta": delta, "ms": round(ms, 3), "loo": loo, "rel_residual": res,
                  "info_block": st["info_block"], "info_pivot": st["info_pivot"]})
        del X32
        del X0
        torch.cuda.empty_cache()
    return out


def batched(args):
    """Batched form: `--batch-count` independent orthogonalizations of uneven
    shapes (m log-uniform in [4096, 2^20], n in {32, 64, 128}, s in
    {4, 8, 16}, kappa log-uniform in [1e2, 1e10]) in one graph-captured
    launch after one tuning call.  One JSON object per variant / precision:
    aggregate algorithmic GB/s, ms, tuned configuration, breakdown count and
    per-instance LOO statistics (instances without breakdown), plus the CPU
    oracle's LOO on the smallest instances for comparison."""
    import numpy as np
    import torch

    import bcgs_hip as B
    import inputs
    import oracle

    B.lib()
    rng = np.random.default_rng(7)
    shapes = []
    for _ in range(args.batch_count):
        m = int(2 ** rng.uniform(12, 20))
        shapes.append((m, int(rng.choice([32, 64, 128])), int(rng.choice([4, 8, 16])),
                       float(10 ** rng.uniform(2, 10))))
    log("generating %d instances" % len(shapes))
    X0 = [make_input(torch, m, n, kappa, seed=100 + i) for i, (m, n, s, kappa) in
          enumerate(shapes)]
    X = [t.clone() for t in X0]
    R = [torch.zeros(n * n, dtype=torch.float64, device="cuda") for (m, n, s, k) in shapes]
    ms_ = [m for (m, n, s, k) in shapes]
    ns_ = [n for (m, n, s, k) in shapes]
    ss_ = [s for (m, n, s, k) in shapes]
    small = sorted(range(len(shapes)), key=lambda i: shapes[i][0] * shapes[i][1])[:8]
    oracle.set_threads(args.cpu_threads)

    def restore(idx):
        for i in idx:
            X[i].copy_(X0[i])
            R[i].zero_()
        torch.cuda.synchronize()

    def measure(name, idx, call):
        restore(idx)
        log("batched %s: tuning on %d instances" % (name, len(idx)))
        info, cfg = call(True)          # tune (captures and times candidates)
        log("batched %s: tuned %s" % (name, cfg))
        times = []
        for _ in range(args.batch_reps):
            restore(idx)
            t0 = time.perf_counter()
            info, cfg = call(False)     # graph replay
            times.append(time.perf_counter() - t0)
        t = sorted(times)[len(times) // 2]
        byt = sum(b_alg(shapes[i][0], shapes[i][1], shapes[i][2]) for i in idx)
        loos, brk = [], 0
        for k, i in enumerate(idx):
            m, n, s, kappa = shapes[i]
            if info[k][0]:
                brk += 1
                continue
            lo, re_, xn = B.accuracy_device(m, n, X0[i].data_ptr(), m, X[i].data_ptr(), m,
                                            R[i].data_ptr(), n)
            loos.append(lo)
        d = {"config": "batched", "case": name, "instances": len(idx), "ms": round(t * 1e3, 2),
             "GBs_alg": round(byt / t / 1e9, 1), "streams": cfg[0], "wg_share": cfg[1],
             "graphs": cfg[2],
             "breakdowns": brk,
             "loo_median": float(np.median(loos)) if loos else None,
             "loo_max": float(max(loos)) if loos else None}
        return d, info

    for variant in ("bcgsi2P1s", "bcgsi2"):
        idx = list(range(len(shapes)))
        d, info = measure(variant, idx, lambda tune: B.bcgs_batched_device(
            variant, ms_, ns_, ss_, [t.data_ptr() for t in X], ms_, [r.data_ptr() for r in R],
            ns_, tune=tune))
        ol = []
        for i in small:
            m, n, s, kappa = shapes[i]
            Qo, Ro, info_o = oracle.bcgs(variant, X0[i].cpu().numpy().T, s)
            if info_o == (0, 0) and info[i] == (0, 0):
                lo, _, _ = B.accuracy_device(m, n, X0[i].data_ptr(), m, X[i].data_ptr(), m,
                                             R[i].data_ptr(), n)
                ol.append((lo, inputs.loo(Qo)))
        d["oracle_sample"] = [{"gpu_loo": a, "cpu_loo": b} for a, b in ol]
        print(json.dumps(finite(d)), flush=True)

    # The batched mixed form is switched off (the entry point returns 2).


def mixed_only(args):
    """Only the mixed form at one first-pass precision (for profiling its
    pass kernel): `--steps` timed factorizations of the FP32-rounded input,
    one JSON object with the mean ms, LOO and residual against the FP64
    input."""
    import torch

    import bcgs_hip as B

    B.lib()
    m, n, s = args.m, args.n, args.s
    X0 = make_input(torch, m, n, args.kappa, seed=1)
    X32 = X0.float()
    Q = torch.empty_like(X0)
    R = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    ts = []
    for k in range(args.warmup + args.steps):
        R.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        B.bcgs_mixed_device("bcgsi2P1s", args.mixed_only, m, n, s, X32.data_ptr(), m,
                            Q.data_ptr(), m, R.data_ptr(), n, stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        if k >= args.warmup:
            ts.append(e0.elapsed_time(e1))
    st = B.status()
    loo, res, xn = B.accuracy_device(m, n, X0.data_ptr(), m, Q.data_ptr(), m, R.data_ptr(), n,
                                     stream.cuda_stream)
    print(json.dumps(finite({"config": "mixed-only", "precision": args.mixed_only, "m": m,
                             "n": n, "s": s, "kappa": args.kappa,
                             "ms": round(sum(ts) / len(ts), 3),
                             "info_block": st["info_block"], "loo": loo,
                             "rel_residual": res / xn})), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=1 << 24)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--s", type=int, default=16)
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--variant", default="bcgsi2P1s")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="rows of the CPU baseline sample (default: all m rows)")
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS", "16")),
                    help="OpenMP threads of the CPU baseline (default: OMP_NUM_THREADS, the "
                         "host share of this GPU, else 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision leg (FP32-stored input)")
    ap.add_argument("--batched", action="store_true",
                    help="run the batched form on an uneven batch instead of the headline "
                         "line (one JSON object per case)")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--suite", action="store_true",
                    help="run the per-variant / per-kappa / glued measurement cases "
                         "instead of the headline line (one JSON object per case)")
    ap.add_argument("--mixed-only", choices=["fp64", "bf16", "fp16", "bf16x3", "bf16_proj",
                                             "fp16_proj", "bf16x3_proj"],
                    help="run only the mixed form at this first-pass precision (profiling)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after timing, run one extra plain elementwise copy of X (read 8mn, "
                         "write 8mn bytes) to calibrate HBM byte counters")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the row slab of one rank at its real "
              "shape; multi-GPU runs are not part of this build (--gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    if args.suite:
        suite(args)
        return
    if args.batched:
        batched(args)
        return
    if args.mixed_only:
        mixed_only(args)
        return

    import torch

    import bcgs_hip as B

    B.lib()
    m, n, s = args.m, args.n, args.s
    log("generating X: m=%d n=%d kappa=%g" % (m, n, args.kappa))
    X0 = make_input(torch, m, n, args.kappa, seed=1)
    X = torch.empty_like(X0)
    R = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def step():
        B.bcgs_device(args.variant, m, n, s, X.data_ptr(), m, R.data_ptr(), n, sp)

    for _ in range(args.warmup):
        X.copy_(X0)
        step()
    torch.cuda.synchronize()
    log("warmup done")

    B.kernel_timing(True)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        X.copy_(X0)  # restore the input; outside the event-timed region
        evs[k][0].record(stream)
        step()
        evs[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = B.kernel_times()
    B.kernel_timing(False)
    step_ms = [a.elapsed_time(b) for a, b in evs]
    ms = sum(step_ms) / len(step_ms)
    st = B.status()
    if args.copy_probe:
        torch.mul(X0, 1.0, out=X)
        torch.cuda.synchronize()
        log("copy probe: %d bytes read, %d bytes written" % (8 * m * n, 8 * m * n))
    log("steps ms:", ["%.2f" % v for v in step_ms], "wall %.3f s" % wall)
    log("kernel classes:", json.dumps(kt))

    acc = {}
    if not args.no_accuracy:
        loo, res, xn = B.accuracy_device(m, n, X0.data_ptr(), m, X.data_ptr(), m,
                                         R.data_ptr(), n, sp)
        acc = {"loo": loo, "rel_residual": res / xn}
        log("LOO %.3e  residual %.3e" % (loo, res / xn))

    # Streaming ceiling of this device (after the accuracy check: the copy
    # overwrites the factorization's Q): the library's plain 16-byte copy
    # X0 -> X (8mn bytes read, 8mn written) and plain read of X0, best of
    # three each, HIP-event timed; torch's copy beside them.
    cp_ms = min(B.stream_probe(0, X.data_ptr(), X0.data_ptr(), 8 * m * n) for _ in range(3))
    rd_ms = min(B.stream_probe(1, 0, X0.data_ptr(), 8 * m * n) for _ in range(3))
    copy_gbs = 16.0 * m * n / (cp_ms / 1e3) / 1e9
    read_gbs = 8.0 * m * n / (rd_ms / 1e3) / 1e9
    tc = []
    for _ in range(3):
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record(stream)
        X.copy_(X0)
        c1.record(stream)
        torch.cuda.synchronize()
        tc.append(c0.elapsed_time(c1))
    torch_copy_gbs = 16.0 * m * n / (min(tc) / 1e3) / 1e9
    log("streaming ceiling: copy %.1f GB/s, read %.1f GB/s (torch copy %.1f GB/s)" % (
        copy_gbs, read_gbs, torch_copy_gbs))

    # Mixed-precision leg: the same input rounded to FP32 storage; first-pass
    # products in each reduced precision, FP64 reorthogonalization.
    mixed = {}
    if not args.no_mixed and args.variant == "bcgsi2P1s" and s == 16 and n <= 256:
        X32 = X0.float()
        base_loo = acc.get("loo")
        for prec in ("fp64", "bf16", "fp16", "bf16x3", "bf16_proj", "fp16_proj", "bf16x3_proj"):
            ts = []
            for k in range(1 + min(args.steps, 3)):
                R.zero_()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                B.bcgs_mixed_device(args.variant, prec, m, n, s, X32.data_ptr(), m, X.data_ptr(),
                                    m, R.data_ptr(), n, sp)
                e1.record(stream)
                torch.cuda.synchronize()
                if k:
                    ts.append(e0.elapsed_time(e1))
            mms = sum(ts) / len(ts)
            stm = B.status()
            d = {"ms": round(mms, 3),
                 "GBs_alg": round(b_alg_mixed(m, n, s) / (mms / 1e3) / 1e9, 1),
                 "speedup_vs_fp64": round(ms / mms, 3), "info_block": stm["info_block"]}
            if not args.no_accuracy:
                lo, re_, xn = B.accuracy_device(m, n, X0.data_ptr(), m, X.data_ptr(), m,
                                                R.data_ptr(), n, sp)
                d.update({"loo": lo, "rel_residual": re_ / xn,
                          "equal_loo": bool(base_loo is not None and lo <= 10 * max(base_loo, 1e-14))})
            mixed[prec] = d
            log("mixed %s: %s" % (prec, json.dumps(finite(d))))
        del X32

    # Dominant kernel class and its roofline.  The fused pass carries every
    # O(m) product of the factorization, so its algorithmic work per
    # factorization is F_alg = 4 m n^2 flop and B_alg bytes; per launch that
    # is F_alg (B_alg) over the launches of one factorization, and
    # achieved = F_alg / (time of the class per factorization), i.e.
    # frac = max(B_alg / BW, F_alg / P) / t.  The schedule's own flops (with
    # the triangular solves and the partly used product tiles) and the
    # counter-measured traffic are reported beside it.
    dom = max(kt.items(), key=lambda kv: kv[1]["ms"])
    name, d = dom
    sec = d["ms"] / 1e3
    sched_gbs = d["bytes"] / sec / 1e9
    sched_tfs = d["flops"] / sec / 1e12
    per_fact = sec / args.steps
    launches_per_fact = d["launches"] / args.steps
    if name == "pass":
        fa, ba = f_alg(m, n), float(b_alg(m, n, s))
    else:
        fa, ba = d["flops"] / args.steps, d["bytes"] / args.steps
    t_f, t_b = fa / (FP64_PEAK_TFLOPS * 1e12), ba / (HBM_PEAK_GBS * 1e9)
    if t_f >= t_b:
        ach = fa / per_fact / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4)}
    else:
        ach = ba / per_fact / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4)}
    traffic = pass_traffic(args.variant, m, n, s) if name == "pass" else None
    roof.update({"traffic": traffic, "kernel": name, "launches": d["launches"],
                 "avg_launch_ms": round(d["ms"] / d["launches"], 4),
                 "alg_flops_per_launch": fa / launches_per_fact,
                 "alg_bytes_per_launch": ba / launches_per_fact,
                 "share_of_step": round(d["ms"] / (ms * args.steps), 3),
                 "alg_gbs": round(ba / per_fact / 1e9, 1),
                 "schedule_tflops": round(sched_tfs, 3), "schedule_gbs": round(sched_gbs, 1),
                 "copy_gbs": round(copy_gbs, 1), "read_gbs": round(read_gbs, 1),
                 "torch_copy_gbs": round(torch_copy_gbs, 1)})

    value = b_alg(m, n, s) / (ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded U diag(sigma) V^T generated on the device)",
        "config": {"workload": "%s FP64 m=%d n=%d s=%d kappa=%g" % (args.variant, m, n, s,
                                                                     args.kappa),
                   "variant": args.variant, "m": m, "n": n, "s": s, "kappa": args.kappa},
        "fp64_tflops_alg": round(f_alg(m, n) / (ms / 1e3) / 1e12, 3),
        "frac_hbm_peak": round(value / HBM_PEAK_GBS, 4),
        "frac_fp64_peak": round(f_alg(m, n) / (ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS, 4),
        "passes": st["passes"],
        "syncs": st["syncs"],
        "info_block": st["info_block"],
        "roofline": roof,
    }
    out.update(acc)
    if mixed:
        out["mixed"] = mixed
    if not args.no_cpu:
        log("CPU baseline ...")
        rows = m if args.cpu_rows <= 0 else min(args.cpu_rows, m)
        out["cpu_baseline"] = cpu_baseline(X0, args.variant, n, s, rows, args.cpu_threads, m)
    print(json.dumps(finite(out)), flush=True)


if __name__ == "__main__":
    main()
