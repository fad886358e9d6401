"""ctypes binding of liborth_hip.so (the MI355X BCGS library).

Mirrors the reference's C interface (include/orth.h: bcgsi2P1s, bcgsi21s,
bcgspipi2, bcgsi2, bcgsi2P2s with the same argument meaning and return codes)
for host numpy arrays, and exposes the device entry points of
include/orth_ext.h for device pointers (e.g. torch tensors' data_ptr()).

There is no CPU fallback: loading fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liborth_hip.so")

VARIANTS = {
    "bcgsi2P1s": 0,
    "bcgsi21s": 1,
    "bcgspipi2": 2,
    "bcgsi2": 3,
    "bcgsi2P2s": 4,
}
PRECISIONS = {"fp64": 0, "bf16": 1, "fp16": 2, "bf16x3": 3,
              "bf16_proj": 4, "fp16_proj": 5, "bf16x3_proj": 6}

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


class Status(ctypes.Structure):
    _fields_ = [("info_block", ctypes.c_int), ("info_pivot", ctypes.c_int),
                ("hip_error", ctypes.c_int), ("passes", ctypes.c_int),
                ("syncs", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load liborth_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "liborth_hip.so is missing (%s); build it with "
            "one-synchronization-bcgs_amd/build.py" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    for name in VARIANTS:
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                      ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int]
    L.orth_last_status.restype = ctypes.c_int
    L.orth_last_status.argtypes = [ctypes.POINTER(Status)]
    L.orth_bcgs_device.restype = ctypes.c_int
    L.orth_bcgs_device.argtypes = [ctypes.c_int, _i64, _i64, _i64, _vp, _i64,
                                   _vp, _i64, _vp]
    L.orth_accuracy_device.restype = ctypes.c_int
    L.orth_accuracy_device.argtypes = [_i64, _i64, _vp, _i64, _vp, _i64, _vp,
                                       _i64, _dp, _vp]
    L.orth_gram_device.restype = ctypes.c_int
    L.orth_gram_device.argtypes = [_i64, _vp, _i64, _i64, _i64, _i64, _i64,
                                   _vp, _i64, _vp]
    L.orth_sync.restype = ctypes.c_int
    L.orth_sync.argtypes = []
    L.orth_release.restype = ctypes.c_int
    L.orth_release.argtypes = []
    L.orth_update_device.restype = ctypes.c_int
    L.orth_update_device.argtypes = [_i64, _vp, _i64, _i64, _i64, _i64, _vp,
                                     _i64, _vp, _vp]
    L.orth_set_pythagorean_shift.restype = ctypes.c_int
    L.orth_set_pythagorean_shift.argtypes = [ctypes.c_double]
    L.orth_kernel_timing.restype = ctypes.c_int
    L.orth_kernel_timing.argtypes = [ctypes.c_int]
    L.orth_kernel_times.restype = ctypes.c_int
    L.orth_kernel_times.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_int), _dp, _dp]
    L.orth_bcgs_mixed_device.restype = ctypes.c_int
    L.orth_bcgs_mixed_device.argtypes = [ctypes.c_int, ctypes.c_int, _i64, _i64, _i64, _vp,
                                         _i64, _vp, _i64, _vp, _i64, _vp]
    _ip = ctypes.POINTER(ctypes.c_int)
    _i64p = ctypes.POINTER(ctypes.c_int64)
    _vpp = ctypes.POINTER(ctypes.c_void_p)
    L.orth_bcgs_batched_device.restype = ctypes.c_int
    L.orth_bcgs_batched_device.argtypes = [ctypes.c_int, ctypes.c_int, _i64p, _i64p, _i64p,
                                           _vpp, _i64p, _vpp, _i64p, _ip, ctypes.c_int, _ip,
                                           _vp]
    L.orth_bcgs_batched_mixed_device.restype = ctypes.c_int
    L.orth_bcgs_batched_mixed_device.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 _i64p, _i64p, _i64p, _vpp, _i64p, _vpp,
                                                 _i64p, _vpp, _i64p, _ip, ctypes.c_int, _ip,
                                                 _vp]
    L.orth_stream_probe.restype = ctypes.c_int
    L.orth_stream_probe.argtypes = [ctypes.c_int, _vp, _vp, _i64, _dp]
    if hasattr(L, "orth_bcgs_device_noalloc"):
        L.orth_bcgs_device_noalloc.restype = ctypes.c_int
        L.orth_bcgs_device_noalloc.argtypes = [ctypes.c_int, _i64, _i64, _i64, _vp, _i64, _vp,
                                               _i64, _vp, _vp]
    _lib = L
    return L


def _addr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def dropin(variant, X, s, lwork=None, work=False):
    """Call a reference-signature symbol on host memory.

    X (m x n) is copied to a Fortran-order array which the call overwrites
    with Q.  Returns (rc, Q, R).  work=True passes a host work buffer of
    lwork doubles (default: exactly the reference's requirement)."""
    Q = np.array(X, dtype=np.float64, order="F", copy=True)
    m, n = Q.shape
    R = np.zeros((n, n), order="F")
    need = 5 * n * s + 9 * s * s + m * (s if variant == "bcgsi2" else n)
    wbuf = None
    if work:
        if lwork is None:
            lwork = need
        wbuf = np.zeros(max(int(lwork), 1))
    else:
        lwork = 0
    rc = getattr(lib(), variant)(m, n, s, _addr(Q), m, _addr(R), n,
                                 _addr(wbuf), int(lwork))
    return rc, Q, R


def bcgs(variant, X, s):
    """Orthogonalize a host matrix; returns (Q, R, status dict)."""
    rc, Q, R = dropin(variant, X, s)
    if rc != 0:
        raise RuntimeError("%s returned %d" % (variant, rc))
    return Q, R, status()


def status():
    st = Status()
    lib().orth_last_status(ctypes.byref(st))
    return st.as_dict()


def bcgs_device(variant, m, n, s, x_ptr, ldx, r_ptr, ldr, stream=0):
    """Enqueue a variant on device pointers (no host wait)."""
    rc = lib().orth_bcgs_device(VARIANTS[variant], m, n, s, x_ptr, ldx, r_ptr,
                                ldr, stream or None)
    if rc != 0:
        raise RuntimeError("orth_bcgs_device returned %d" % rc)


def bcgs_device_noalloc(variant, m, n, s, x_ptr, ldx, r_ptr, ldr, info_ptr=None, stream=0):
    """Enqueue a variant on device pointers using only the library's fixed
    device scratch (no allocation); info_ptr (device, 2 int32) receives the
    first breakdown (block, pivot).  Returns the entry point's code."""
    return lib().orth_bcgs_device_noalloc(VARIANTS[variant], m, n, s, x_ptr, ldx, r_ptr, ldr,
                                          info_ptr, stream or None)


def bcgs_mixed_device(variant, precision, m, n, s, x32_ptr, ldx, q_ptr, ldq,
                      r_ptr, ldr, stream=0):
    rc = lib().orth_bcgs_mixed_device(VARIANTS[variant], PRECISIONS[precision],
                                      m, n, s, x32_ptr, ldx, q_ptr, ldq,
                                      r_ptr, ldr, stream or None)
    if rc != 0:
        raise RuntimeError("orth_bcgs_mixed_device returned %d" % rc)


def _arr(ctype, vals):
    return (ctype * len(vals))(*vals)


def bcgs_batched_device(variant, ms, ns, ss, x_ptrs, ldxs, r_ptrs, ldrs, tune=False,
                        stream=0, precision=None, x32_ptrs=None, ldx32s=None):
    """Run a batch of independent orthogonalizations on device pointers
    (synchronous).  With precision set, x32_ptrs are the FP32 inputs and
    x_ptrs the FP64 outputs Q.  Returns (info, config): info[i] = (block,
    pivot) of instance i's first breakdown, config = (streams, workgroup
    share, graphs used) of the run."""
    cnt = len(ms)
    info = (ctypes.c_int * (2 * cnt))()
    cfg = (ctypes.c_int * 2)()
    i64 = ctypes.c_int64
    vp = ctypes.c_void_p
    common = (_arr(i64, ms), _arr(i64, ns), _arr(i64, ss))
    if precision is None:
        rc = lib().orth_bcgs_batched_device(VARIANTS[variant], cnt, *common, _arr(vp, x_ptrs),
                                            _arr(i64, ldxs), _arr(vp, r_ptrs), _arr(i64, ldrs),
                                            info, 1 if tune else 0, cfg, stream or None)
    else:
        rc = lib().orth_bcgs_batched_mixed_device(
            VARIANTS[variant], PRECISIONS[precision], cnt, *common, _arr(vp, x32_ptrs),
            _arr(i64, ldx32s), _arr(vp, x_ptrs), _arr(i64, ldxs), _arr(vp, r_ptrs),
            _arr(i64, ldrs), info, 1 if tune else 0, cfg, stream or None)
    if rc != 0:
        raise RuntimeError("batched call returned %d" % rc)
    return ([(info[2 * i], info[2 * i + 1]) for i in range(cnt)],
            (abs(cfg[0]), cfg[1] / 1000.0, cfg[0] > 0))


def accuracy_device(m, n, x_ptr, ldx, q_ptr, ldq, r_ptr, ldr, stream=0):
    out = (ctypes.c_double * 3)()
    rc = lib().orth_accuracy_device(m, n, x_ptr, ldx, q_ptr, ldq, r_ptr, ldr,
                                    out, stream or None)
    if rc != 0:
        raise RuntimeError("orth_accuracy_device returned %d" % rc)
    return out[0], out[1], out[2]


def gram_device(m, x_ptr, ldx, a0, na, b0, nb, c_ptr, ldc, stream=0):
    rc = lib().orth_gram_device(m, x_ptr, ldx, a0, na, b0, nb, c_ptr, ldc,
                                stream or None)
    if rc != 0:
        raise RuntimeError("orth_gram_device returned %d" % rc)


def update_device(m, x_ptr, ldx, col0, w, c, w_ptr, ldw, t_ptr, stream=0):
    rc = lib().orth_update_device(m, x_ptr, ldx, col0, w, c, w_ptr, ldw,
                                  t_ptr, stream or None)
    if rc != 0:
        raise RuntimeError("orth_update_device returned %d" % rc)


def stream_probe(mode, dst_ptr, src_ptr, nbytes):
    """Milliseconds of one plain streaming copy (mode 0) or read (mode 1) of
    nbytes on the device."""
    ms = ctypes.c_double()
    rc = lib().orth_stream_probe(mode, dst_ptr, src_ptr, nbytes, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError("orth_stream_probe returned %d" % rc)
    return ms.value


def sync():
    lib().orth_sync()


KERNELS = {"proj": 0, "reduce": 1, "update": 2, "small": 3, "pass": 4}


def set_pythagorean_shift(c):
    """Shift the Pythagorean Cholesky steps of later calls by c u tr(G)
    (0 restores the reference's arithmetic)."""
    if lib().orth_set_pythagorean_shift(float(c)) != 0:
        raise ValueError("shift constant must be finite and >= 0")


def kernel_timing(enable):
    """Enable/disable per-launch HIP-event timing (clears the record)."""
    lib().orth_kernel_timing(1 if enable else 0)


def kernel_times():
    """{kernel class: dict(ms, launches, flops, bytes)} recorded since
    kernel_timing(); flops and bytes are the algorithmic work of those
    launches."""
    out = {}
    for name, k in KERNELS.items():
        t = ctypes.c_double()
        c = ctypes.c_int()
        f = ctypes.c_double()
        b = ctypes.c_double()
        lib().orth_kernel_times(k, ctypes.byref(t), ctypes.byref(c),
                                ctypes.byref(f), ctypes.byref(b))
        if c.value:
            out[name] = {"ms": t.value, "launches": c.value, "flops": f.value,
                         "bytes": b.value}
    return out
