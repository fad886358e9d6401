"""Host-side sanitizer run of the library's schedules (no GPU).

one-synchronization-bcgs_amd/csrc/orth_host.cpp is compiled with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer against a host-only stand-in
of the HIP runtime (tests/host_harness/hip/hip_runtime.h, hip_stub.cpp:
"device" memory is host memory, kernels do nothing, streams / events /
graphs are bookkeeping that enforces the capture rules the library relies
on).  tests/host_harness/driver.cpp then replays the batched benchmark's
calls: uneven batches, tuning and graph replays, sub-batches, three FP64
variants and the mixed batched form -- whose entry point is switched off in
the library, so the harness build switches it back on in its private copy of
the source.  Runs with an injected launch failure at several points exercise
the error paths (a failing capture, a failing tuning candidate): every call
must return 0 or 2, the sanitizers must stay silent, and the stand-in must
report no misuse (allocation or synchronization while capturing, a capture
ended with unjoined streams, unknown handles).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "one-synchronization-bcgs_amd", "csrc")
HARNESS = os.path.join(HERE, "host_harness")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("host_harness")
    src = open(os.path.join(CSRC, "orth_host.cpp")).read()
    src = src.replace('"../../include/orth.h"', '"orth.h"')
    src = src.replace('"../../include/orth_ext.h"', '"orth_ext.h"')
    off = "    if (a.mx) return false;\n"
    assert src.count(off) == 1, "the batched mixed switch moved; update the harness"
    src = src.replace(off, "    // (switched on for the host harness)\n")
    copy = d / "orth_host_harness.cpp"
    copy.write_text(src)
    exe = d / "harness"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined",
           "-I" + HARNESS, "-I" + CSRC, "-I" + os.path.join(ROOT, "include"), str(copy),
           os.path.join(HARNESS, "hip_stub.cpp"), os.path.join(HARNESS, "driver.cpp"),
           "-o", str(exe), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    return str(exe)


def run(exe, count, mmax, fail_at=None, extra_env=None):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    env.pop("ORTH_TUNE_CACHE", None)
    if fail_at is not None:
        env["STUB_FAIL_LAUNCH"] = str(fail_at)
    env.update(extra_env or {})
    r = subprocess.run([exe, str(count), str(mmax)], capture_output=True, text=True,
                       timeout=600, env=env)
    return r


def test_batched_paths_clean_under_sanitizers(harness):
    r = run(harness, 96, 1 << 14)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "stand-in misuse reports: 0, unexpected returns: 0" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out


@pytest.mark.parametrize("fail_at", [1, 5, 50, 400, 2500, 9000])
def test_batched_error_paths_clean_under_sanitizers(harness, fail_at):
    r = run(harness, 32, 1 << 13, fail_at)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "stand-in misuse reports: 0, unexpected returns: 0" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out


def configs(out):
    """{variant: (streams, share)} printed by the driver."""
    res = {}
    for line in out.splitlines():
        if line.startswith("batched fp64 variant "):
            w = line.split()
            res[int(w[3])] = (int(w[5]), int(w[6]))
    return res


def test_tuning_cache_round_trip(harness, tmp_path):
    """ORTH_TUNE_CACHE: a tuning run writes one line per batch signature for
    this device; a later process that does not tune takes its configuration
    from the file (here a hand-edited line that no tuning candidate
    produces); lines of another device are kept when the file is rewritten,
    lines that do not parse are dropped."""
    cache = tmp_path / "tune.txt"
    env = {"ORTH_TUNE_CACHE": str(cache)}
    r = run(harness, 24, 1 << 13, extra_env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    lines = cache.read_text().splitlines()
    assert all(l.startswith("orth-tune-1 gfx950/256 ") for l in lines)
    # Fields: tag, device, variant, mixed precision (0: FP64), count, m n s...
    fp64 = [l for l in lines if l.split()[3] == "0"]
    assert sorted(int(l.split()[2]) for l in fp64) == [0, 2, 3]
    assert all(l.split()[4] == "24" and len(l.split(" : ")[0].split()) == 5 + 3 * 24 for l in fp64)
    # Edit the FP64 P-1S line to a configuration no candidate gives: 3
    # streams, share 0.375, direct launches, double-buffered pass from 64
    # columns.
    edited = []
    for l in lines:
        if l.split()[2:4] == ["0", "0"]:
            l = l.split(" : ")[0] + " : 3 0.375 0 64"
        edited.append(l)
    foreign = "orth-tune-1 gfx942/304 0 0 1 4096 64 16 : 2 0.5 1 0"
    cache.write_text("\n".join(edited + [foreign, "not a tuning line"]) + "\n")
    r = run(harness, 24, 1 << 13, extra_env=dict(env, HARNESS_NO_TUNE="1"))
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "stand-in misuse reports: 0, unexpected returns: 0" in out
    cfg = configs(out)
    assert cfg[0] == (-3, 375)
    # Without the cache the untuned default applies (4 streams, graphs, half
    # the device per pass).
    r = run(harness, 24, 1 << 13, extra_env={"HARNESS_NO_TUNE": "1"})
    assert configs(r.stdout)[0] == (4, 500)
    # The harness's mixed batches tuned in that run and rewrote the file: the
    # other device's line and the edited line are kept, the line that does
    # not parse is dropped.
    text = cache.read_text()
    assert foreign in text and "not a tuning line" not in text
    assert sum(1 for l in text.splitlines() if l.endswith(" : 3 0.375 0 64")) == 1
