"""ctypes binding of libatrip_hip.so (include/atrip_hip.h).

The product path: every call goes to the HIP library; there is no CPU
fallback.  A missing library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libatrip_hip.so")

FP64 = 0
SPLIT_FP32 = 1
FP32 = 2
BF16X3 = 3
PRECISIONS = {"fp64": FP64, "split_fp32": SPLIT_FP32,
              "fp32": FP32, "bf16x3": BF16X3}

# Exported C-ABI symbols (kept in sync with include/atrip_hip.h).
SYMBOLS = (
    "atrip_hip_create", "atrip_hip_destroy", "atrip_hip_last_error",
    "atrip_hip_upload", "atrip_hip_upload_synthetic", "atrip_hip_release",
    "atrip_hip_upload_synthetic_scaled", "atrip_hip_upload_stream",
    "atrip_hip_upload_complex", "atrip_hip_upload_synthetic_complex",
    "atrip_hip_run", "atrip_hip_tuple_energies", "atrip_hip_set_option",
    "atrip_hip_run_windows",
    "atrip_hip_get_option", "atrip_hip_kernel_stats", "atrip_hip_reset_stats",
    "atrip_hip_sys_info", "atrip_hip_device_count",
    "atrip_hip_batch_run", "atrip_hip_batch_tune", "atrip_hip_batch_plan",
    "atrip_hip_batch_model", "atrip_hip_serve_create", "atrip_hip_serve_destroy",
    "atrip_hip_serve_set_policy", "atrip_hip_serve_clock_ms",
    "atrip_hip_serve_submit_synthetic", "atrip_hip_serve_submit",
    "atrip_hip_serve_step", "atrip_hip_serve_result", "atrip_hip_serve_stats",
    "atrip_hip_run_noalloc", "atrip_hip_plan_model",
)

TENSORS = ("ei", "ea", "Tai", "Tabij", "Vabij", "Vijka", "Vabci", "Jijka", "Jabci")


STREAM_WITH_J = 1
STREAM_COMPLEX = 2

# int (*)(void *user, int tensor, int64_t first, int64_t count, double *out)
READ_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int64, C.c_int64,
                      C.POINTER(C.c_double))


class HipError(RuntimeError):
    pass


class SysDesc(C.Structure):
    _fields_ = [("No", C.c_int64), ("Nv", C.c_int64)] + [
        (name, C.POINTER(C.c_double)) for name in TENSORS
    ]


_lib = None


def load():
    """Load the HIP library; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipError(
            f"{LIB_PATH} is missing: build it with `make -C atrip_amd` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.atrip_hip_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.atrip_hip_destroy.argtypes = [vp]
    L.atrip_hip_last_error.argtypes = [vp]
    L.atrip_hip_last_error.restype = C.c_char_p
    L.atrip_hip_upload.argtypes = [vp, C.POINTER(SysDesc), C.POINTER(C.c_int)]
    L.atrip_hip_upload_synthetic.argtypes = [
        vp, C.c_int64, C.c_int64, C.c_uint64, C.c_int, C.POINTER(C.c_int)]
    L.atrip_hip_upload_synthetic_scaled.argtypes = [
        vp, C.c_int64, C.c_int64, C.c_uint64, C.c_int, C.c_double,
        C.POINTER(C.c_int)]
    L.atrip_hip_upload_stream.argtypes = [
        vp, C.c_int64, C.c_int64, C.c_int, READ_FN, C.c_void_p,
        C.POINTER(C.c_int)]
    L.atrip_hip_release.argtypes = [vp, C.c_int]
    L.atrip_hip_upload_complex.argtypes = [vp, C.POINTER(SysDesc), C.POINTER(C.c_int)]
    L.atrip_hip_upload_synthetic_complex.argtypes = [
        vp, C.c_int64, C.c_int64, C.c_uint64, C.c_int, C.POINTER(C.c_int)]
    L.atrip_hip_run.argtypes = [
        vp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
        C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.atrip_hip_run_noalloc.argtypes = [
        vp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
        C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.atrip_hip_run_windows.argtypes = [
        vp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
        C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    L.atrip_hip_tuple_energies.argtypes = [
        vp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
        C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.atrip_hip_set_option.argtypes = [vp, C.c_char_p, C.c_int64]
    L.atrip_hip_get_option.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
    L.atrip_hip_kernel_stats.argtypes = [
        vp, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
        C.POINTER(C.c_double)]
    L.atrip_hip_reset_stats.argtypes = [vp]
    L.atrip_hip_sys_info.argtypes = [
        vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.atrip_hip_device_count.argtypes = []
    L.atrip_hip_device_count.restype = C.c_int
    ip = C.POINTER(C.c_int)
    L.atrip_hip_batch_run.argtypes = [vp, ip, C.c_int, C.c_int,
                                      C.POINTER(C.c_double)]
    L.atrip_hip_batch_tune.argtypes = [vp, ip, C.c_int, C.c_int, C.c_int]
    L.atrip_hip_batch_plan.argtypes = [vp, C.c_int, C.c_int, ip, ip, ip, ip,
                                       C.POINTER(C.c_double)]
    L.atrip_hip_batch_model.argtypes = [C.c_int64, C.c_int64, C.c_int,
                                        C.c_int64, C.c_int, ip, ip, ip]
    L.atrip_hip_plan_model.argtypes = [C.c_int64, C.c_int64, C.c_int, C.c_int,
                                       ip, ip, ip]
    dp = C.POINTER(C.c_double)
    u64p = C.POINTER(C.c_uint64)
    L.atrip_hip_serve_create.argtypes = [vp, C.c_uint64, C.c_int, C.c_double,
                                         C.c_int, C.POINTER(vp)]
    L.atrip_hip_serve_destroy.argtypes = [vp]
    L.atrip_hip_serve_set_policy.argtypes = [vp, C.c_int]
    L.atrip_hip_serve_clock_ms.argtypes = [vp]
    L.atrip_hip_serve_clock_ms.restype = C.c_double
    L.atrip_hip_serve_submit_synthetic.argtypes = [
        vp, C.c_int64, C.c_int64, C.c_uint64, C.c_double, C.POINTER(C.c_int64)]
    L.atrip_hip_serve_submit.argtypes = [vp, C.POINTER(SysDesc), C.c_double,
                                         C.POINTER(C.c_int64)]
    L.atrip_hip_serve_step.argtypes = [vp, ip]
    L.atrip_hip_serve_result.argtypes = [vp, C.c_int64, ip, dp, dp, dp, dp]
    L.atrip_hip_serve_stats.argtypes = [vp, ip, ip, u64p, u64p, u64p, u64p, dp]
    _lib = L
    return L


def device_count() -> int:
    return int(load().atrip_hip_device_count())


def batch_model(no: int, nv: int, scratch_mb: int = 8192, grid: int = 256,
                precision="fp64"):
    """Cost-model plan (window edge, fm, fn) of the batched form; host only."""
    nb, fm, fn = C.c_int(), C.c_int(), C.c_int()
    rc = load().atrip_hip_batch_model(no, nv, PRECISIONS[precision], scratch_mb,
                                      grid, C.byref(nb), C.byref(fm), C.byref(fn))
    if rc != 0:
        raise HipError(f"atrip_hip_batch_model({no}, {nv}) failed ({rc})")
    return nb.value, fm.value, fn.value


def plan_model(no: int, nv: int, precision="fp64", window: int = 0):
    """Plan of the plain path (window edge, fm, fn) that atrip_hip_run takes
    for this shape with the automatic tile; host only."""
    nb, fm, fn = C.c_int(), C.c_int(), C.c_int()
    rc = load().atrip_hip_plan_model(no, nv, PRECISIONS[precision], window,
                                     C.byref(nb), C.byref(fm), C.byref(fn))
    if rc != 0:
        raise HipError(f"atrip_hip_plan_model({no}, {nv}, {precision}, {window}) "
                       f"failed ({rc})")
    return nb.value, fm.value, fn.value


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class KernelStats:
    launches: int
    total_ms: float
    flop: float

    @property
    def avg_ms(self) -> float:
        return self.total_ms / self.launches if self.launches else 0.0


class Context:
    """One HIP device context (not thread-safe; one host thread each)."""

    def __init__(self, device: int = 0):
        self._L = load()
        self._h = C.c_void_p()
        rc = self._L.atrip_hip_create(device, C.byref(self._h))
        if rc != 0:
            raise HipError(f"atrip_hip_create(device={device}) failed ({rc})")

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.atrip_hip_last_error(self._h).decode()
            raise HipError(f"{what}: {msg}")

    def close(self):
        if self._h:
            self._L.atrip_hip_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- systems ---------------------------------------------------------
    def upload(self, no: int, nv: int, t: dict) -> int:
        """Upload CTF-layout float64 arrays (flat, first index fastest)."""
        d = SysDesc()
        d.No, d.Nv = no, nv
        keep = []
        for name in TENSORS:
            arr = t.get(name)
            if arr is None:
                setattr(d, name, C.POINTER(C.c_double)())
                continue
            arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
            keep.append(arr)
            setattr(d, name, _p(arr))
        sys_ = C.c_int()
        self._check(self._L.atrip_hip_upload(self._h, C.byref(d), C.byref(sys_)),
                    "upload")
        return sys_.value

    def upload_synthetic(self, no: int, nv: int, seed: int, with_j=False) -> int:
        sys_ = C.c_int()
        self._check(self._L.atrip_hip_upload_synthetic(
            self._h, no, nv, seed, int(with_j), C.byref(sys_)), "upload_synthetic")
        return sys_.value

    def upload_synthetic_scaled(self, no: int, nv: int, seed: int, scale: float,
                                with_j=False) -> int:
        """Synthetic system with amplitudes and integrals times `scale`."""
        sys_ = C.c_int()
        self._check(self._L.atrip_hip_upload_synthetic_scaled(
            self._h, no, nv, seed, int(with_j), float(scale), C.byref(sys_)),
            "upload_synthetic_scaled")
        return sys_.value

    def upload_stream(self, no: int, nv: int, read, with_j=False, complex_=False) -> int:
        """Upload through a reader: read(tensor_name, first, count) returns the
        values [first, first + count) of that tensor in CTF order (float64, or
        complex128 for complex systems); the library asks for slabs."""
        comp = 2 if complex_ else 1

        def cb(_user, tensor, first, count, out):
            try:
                vals = np.asarray(read(TENSORS[tensor], first, count))
                if complex_:
                    vals = np.ascontiguousarray(vals, dtype=np.complex128).view(np.float64)
                else:
                    vals = np.ascontiguousarray(vals, dtype=np.float64)
                if vals.size != count * comp:
                    return 1
                np.ctypeslib.as_array(out, shape=(count * comp,))[:] = vals
                return 0
            except Exception:  # reported through the library's status
                return 1

        fn = READ_FN(cb)
        flags = (STREAM_WITH_J if with_j else 0) | (STREAM_COMPLEX if complex_ else 0)
        sys_ = C.c_int()
        self._check(self._L.atrip_hip_upload_stream(
            self._h, no, nv, flags, fn, None, C.byref(sys_)), "upload_stream")
        return sys_.value

    def upload_complex(self, no: int, nv: int, t: dict) -> int:
        """Upload complex CTF-layout arrays (complex128, first index fastest)."""
        d = SysDesc()
        d.No, d.Nv = no, nv
        keep = []
        for name in TENSORS:
            arr = t.get(name)
            if arr is None:
                setattr(d, name, C.POINTER(C.c_double)())
                continue
            arr = np.ascontiguousarray(arr, dtype=np.complex128).reshape(-1)
            keep.append(arr)
            setattr(d, name, arr.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)))
        sys_ = C.c_int()
        self._check(self._L.atrip_hip_upload_complex(self._h, C.byref(d), C.byref(sys_)),
                    "upload_complex")
        return sys_.value

    def upload_synthetic_complex(self, no: int, nv: int, seed: int, with_j=False) -> int:
        sys_ = C.c_int()
        self._check(self._L.atrip_hip_upload_synthetic_complex(
            self._h, no, nv, seed, int(with_j), C.byref(sys_)), "upload_synthetic_complex")
        return sys_.value

    def release(self, sys_: int):
        self._check(self._L.atrip_hip_release(self._h, sys_), "release")

    def sys_info(self, sys_: int):
        no, nv = C.c_int64(), C.c_int64()
        nt, nbytes = C.c_uint64(), C.c_uint64()
        self._check(self._L.atrip_hip_sys_info(
            self._h, sys_, C.byref(no), C.byref(nv), C.byref(nt), C.byref(nbytes)),
            "sys_info")
        return dict(No=no.value, Nv=nv.value, n_tuples=nt.value, bytes=nbytes.value)

    # -- runs ------------------------------------------------------------
    def run(self, sys_: int, precision="fp64", ijkabc=False, first=0, count=None):
        e, ect = C.c_double(), C.c_double()
        cnt = (1 << 64) - 1 if count is None else count
        self._check(self._L.atrip_hip_run(
            self._h, sys_, PRECISIONS[precision], int(ijkabc), first, cnt,
            C.byref(e), C.byref(ect)), "run")
        return e.value, ect.value

    def run_noalloc(self, sys_: int, precision="fp64", ijkabc=False, first=0,
                    count=None):
        """run() with every piece of device scratch in the library's fixed
        per-GPU region (no device allocation)."""
        e, ect = C.c_double(), C.c_double()
        cnt = (1 << 64) - 1 if count is None else count
        self._check(self._L.atrip_hip_run_noalloc(
            self._h, sys_, PRECISIONS[precision], int(ijkabc), first, cnt,
            C.byref(e), C.byref(ect)), "run_noalloc")
        return e.value, ect.value

    def run_windows(self, sys_: int, stride: int, offset: int = 0, precision="fp64",
                    ijkabc=False):
        """E(T), E(cT) and tuple count of every stride-th window of the plan."""
        e, ect, nt = C.c_double(), C.c_double(), C.c_uint64()
        self._check(self._L.atrip_hip_run_windows(
            self._h, sys_, PRECISIONS[precision], int(ijkabc), stride, offset,
            C.byref(e), C.byref(ect), C.byref(nt)), "run_windows")
        return e.value, ect.value, nt.value

    def tuple_energies(self, sys_: int, first: int, count: int, precision="fp64",
                       ijkabc=False, with_ct=False):
        e = np.zeros(max(count, 1))
        ect = np.zeros(max(count, 1)) if with_ct else None
        self._check(self._L.atrip_hip_tuple_energies(
            self._h, sys_, PRECISIONS[precision], int(ijkabc), first, count,
            _p(e), _p(ect) if with_ct else C.POINTER(C.c_double)()),
            "tuple_energies")
        return (e[:count], ect[:count]) if with_ct else e[:count]

    # -- batched form ----------------------------------------------------
    def batch_run(self, systems, precision="fp64") -> np.ndarray:
        """E(T) of every listed resident system, all in one persistent launch."""
        arr = (C.c_int * max(len(systems), 1))(*systems)
        e = np.zeros(max(len(systems), 1))
        self._check(self._L.atrip_hip_batch_run(
            self._h, arr, len(systems), PRECISIONS[precision], _p(e)), "batch_run")
        return e[:len(systems)]

    def batch_tune(self, systems, measure=True, precision="fp64"):
        arr = (C.c_int * max(len(systems), 1))(*systems)
        self._check(self._L.atrip_hip_batch_tune(
            self._h, arr, len(systems), PRECISIONS[precision],
            1 if measure else 0), "batch_tune")

    def batch_plan(self, sys_: int, precision="fp64") -> dict:
        nb, fm, fn, meas = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        sec = C.c_double()
        self._check(self._L.atrip_hip_batch_plan(
            self._h, sys_, PRECISIONS[precision], C.byref(nb), C.byref(fm),
            C.byref(fn), C.byref(meas), C.byref(sec)), "batch_plan")
        return dict(nb=nb.value, fm=fm.value, fn=fn.value,
                    measured=bool(meas.value), sec_window=sec.value)

    # -- knobs -----------------------------------------------------------
    def set_option(self, key: str, value: int):
        self._check(self._L.atrip_hip_set_option(self._h, key.encode(), int(value)),
                    f"set_option({key})")

    def get_option(self, key: str) -> int:
        v = C.c_int64()
        self._check(self._L.atrip_hip_get_option(self._h, key.encode(), C.byref(v)),
                    f"get_option({key})")
        return v.value

    def kernel_stats(self, name: str) -> KernelStats:
        n, ms, fl = C.c_uint64(), C.c_double(), C.c_double()
        self._check(self._L.atrip_hip_kernel_stats(
            self._h, name.encode(), C.byref(n), C.byref(ms), C.byref(fl)),
            "kernel_stats")
        return KernelStats(n.value, ms.value, fl.value)

    def reset_stats(self):
        self._check(self._L.atrip_hip_reset_stats(self._h), "reset_stats")
