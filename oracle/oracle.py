"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The product path never uses it.

Tensors are flat float64 numpy arrays in CTF column-major layout (first index
fastest), lens as in the reference's bench/main.cxx:198-200.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

TENSORS = ("ei", "ea", "Tai", "Tabij", "Vabij", "Vijka", "Vabci", "Jijka", "Jabci")
TENSOR_ID = {name: i for i, name in enumerate(TENSORS)}

MISTAKES = {
    "none": 0,
    "p3_perm": 1,
    "h0_sign": 2,
    "drop_p4": 3,
    "same_as_distinct": 4,
    "factor": 5,
    "h1_transpose": 6,
    "singles_index": 7,
}


class OracleInput(C.Structure):
    _fields_ = [("No", C.c_size_t), ("Nv", C.c_size_t)] + [
        (name, C.POINTER(C.c_double)) for name in TENSORS
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        L.oracle_n_tuples.restype = C.c_uint64
        L.oracle_n_tuples.argtypes = [C.c_uint64]
        L.oracle_all_tuples.argtypes = [C.c_uint64, C.POINTER(C.c_uint64)]
        L.oracle_nth_tuple.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
        L.oracle_run.restype = C.c_int
        L.oracle_run.argtypes = [
            C.POINTER(OracleInput), C.c_int, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
            C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
            C.POINTER(C.c_double), C.POINTER(C.c_double),
        ]
        L.oracle_fill_synthetic.argtypes = [
            C.c_uint64, C.c_int, C.c_size_t, C.c_size_t, C.POINTER(C.c_double), C.c_size_t,
            C.c_int,
        ]
        L.oracle_blas_init.restype = C.c_char_p
        L.oracle_blas_init.argtypes = [C.c_char_p]
        L.oracle_blas_name.restype = C.c_char_p
        dptr = C.POINTER(C.c_double)
        L.oracle_doubles_loop.argtypes = [C.c_size_t, C.c_size_t] + [dptr] * 16
        L.oracle_doubles_gemm.argtypes = [C.c_size_t, C.c_size_t] + [dptr] * 18
        L.oracle_singles.argtypes = [C.c_size_t] * 5 + [dptr] * 5
        L.oracle_energy_distinct.restype = C.c_double
        L.oracle_energy_distinct.argtypes = [C.c_double, C.c_size_t, dptr, dptr, dptr]
        L.oracle_energy_same.restype = C.c_double
        L.oracle_energy_same.argtypes = [C.c_double, C.c_size_t, dptr, dptr, dptr]
        L.oracle_run_synthetic.restype = C.c_int
        L.oracle_run_synthetic.argtypes = [
            C.c_uint64, C.c_size_t, C.c_size_t, C.c_int, C.c_double, C.c_int,
            C.c_uint64, C.c_uint64, C.c_int, C.c_int, dptr, dptr]
        L.oracle_run_z.restype = C.c_int
        L.oracle_run_z.argtypes = [
            C.POINTER(OracleInput), C.c_int, C.c_uint64, C.c_uint64, C.c_int,
            dptr, dptr, dptr, dptr]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def n_tuples(nv: int) -> int:
    return int(lib().oracle_n_tuples(nv))


def all_tuples(nv: int) -> np.ndarray:
    n = n_tuples(nv)
    out = np.zeros(3 * max(n, 1), dtype=np.uint64)
    lib().oracle_all_tuples(nv, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out[: 3 * n].reshape(n, 3)


def nth_tuple(it: int, nv: int):
    out = (C.c_uint64 * 3)()
    lib().oracle_nth_tuple(it, nv, out)
    return (int(out[0]), int(out[1]), int(out[2]))


def tensor_size(name: str, no: int, nv: int) -> int:
    return {
        "ei": no, "ea": nv, "Tai": nv * no,
        "Tabij": nv * nv * no * no, "Vabij": nv * nv * no * no,
        "Vijka": no ** 3 * nv, "Jijka": no ** 3 * nv,
        "Vabci": nv ** 3 * no, "Jabci": nv ** 3 * no,
    }[name]


def synthetic(no: int, nv: int, seed: int, with_j: bool = False, nthreads: int = 1) -> dict:
    """Seeded synthetic inputs (include/atrip_synth.h), CTF layout, flat."""
    out = {}
    names = TENSORS if with_j else TENSORS[:7]
    for name in names:
        n = tensor_size(name, no, nv)
        arr = np.empty(n, dtype=np.float64)
        lib().oracle_fill_synthetic(seed, TENSOR_ID[name], no, nv, _p(arr), n, nthreads)
        out[name] = arr
    return out


IMAG = 16  # imaginary-part flag of the generator's tensor ids


def synthetic_z(no: int, nv: int, seed: int, with_j: bool = False,
                nthreads: int = 1) -> dict:
    """Seeded complex inputs: real parts as synthetic(), imaginary parts from
    the tensor ids | IMAG (orbital energies real); complex128, CTF layout."""
    out = {}
    names = TENSORS if with_j else TENSORS[:7]
    for name in names:
        n = tensor_size(name, no, nv)
        re = np.empty(n, dtype=np.float64)
        im = np.empty(n, dtype=np.float64)
        lib().oracle_fill_synthetic(seed, TENSOR_ID[name], no, nv, _p(re), n, nthreads)
        lib().oracle_fill_synthetic(seed, TENSOR_ID[name] | IMAG, no, nv, _p(im), n,
                                    nthreads)
        out[name] = re + 1j * im
    return out


def run_z(no: int, nv: int, t: dict, *, ijkabc=False, first=0, count=None,
          nthreads=1, per_tuple=False):
    """Complex (T) loop (oracle_run_z); complex128 inputs in CTF layout."""
    n = n_tuples(nv)
    if count is None:
        count = n - first
    keep = {k: np.ascontiguousarray(v, dtype=np.complex128) for k, v in t.items()}
    inp = OracleInput()
    inp.No, inp.Nv = no, nv
    for name in TENSORS:
        arr = keep.get(name)
        setattr(inp, name, arr.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
                if arr is not None else C.POINTER(C.c_double)())
    e, ect = C.c_double(), C.c_double()
    pt = np.zeros(max(count, 1))
    ptc = np.zeros(max(count, 1))
    rc = lib().oracle_run_z(C.byref(inp), int(ijkabc), first, count, nthreads,
                            C.byref(e), C.byref(ect), _p(pt), _p(ptc))
    if rc != 0:
        raise RuntimeError(f"oracle_run_z failed ({rc})")
    if per_tuple:
        return e.value, ect.value, pt[:count], ptc[:count]
    return e.value, ect.value


def make_input(no: int, nv: int, t: dict) -> OracleInput:
    inp = OracleInput()
    inp.No, inp.Nv = no, nv
    for name in TENSORS:
        arr = t.get(name)
        setattr(inp, name, _p(arr) if arr is not None else C.POINTER(C.c_double)())
    return inp


def run(no: int, nv: int, t: dict, *, ijkabc=False, first=0, count=None,
        use_gemm=False, nthreads=1, mistake="none", per_tuple=False):
    """Full single-rank (T) loop; returns (energy, ct_energy[, e_t, e_ct_t])."""
    n = n_tuples(nv)
    if count is None:
        count = n - first
    inp = make_input(no, nv, t)
    e, ect = C.c_double(), C.c_double()
    pt = np.zeros(max(count, 1)) if per_tuple else None
    ptc = np.zeros(max(count, 1)) if per_tuple else None
    null = C.POINTER(C.c_double)()
    rc = lib().oracle_run(
        C.byref(inp), int(ijkabc), first, count, int(use_gemm), nthreads,
        MISTAKES[mistake], C.byref(e), C.byref(ect),
        _p(pt) if per_tuple else null, _p(ptc) if per_tuple else null)
    if rc != 0:
        raise RuntimeError(f"oracle_run failed ({rc})")
    if per_tuple:
        return e.value, ect.value, pt[:count], ptc[:count]
    return e.value, ect.value


def run_synthetic(no: int, nv: int, seed: int, first: int, count: int, *,
                  with_j=False, scale=1.0, ijkabc=False, use_gemm=True, nthreads=1):
    """Raw per-tuple energies (and (cT) energies) of tuples [first,
    first+count) of the seeded synthetic system, slices drawn from the
    generator per tuple (no whole tensors in host memory)."""
    n = n_tuples(nv)
    first = min(first, n)
    count = max(0, min(count, n - first))
    pt = np.zeros(max(count, 1))
    ptc = np.zeros(max(count, 1))
    rc = lib().oracle_run_synthetic(seed, no, nv, int(with_j), float(scale),
                                    int(ijkabc), first, count, int(use_gemm),
                                    nthreads, _p(pt), _p(ptc))
    if rc != 0:
        raise RuntimeError(f"oracle_run_synthetic failed ({rc})")
    return pt[:count], ptc[:count]


def blas_init(path=None):
    r = lib().oracle_blas_init(path.encode() if path else None)
    return r.decode() if r else None


# ---- per-tuple kernels on explicit slices (for closed-form pins) ----------

def doubles(no, nv, slices, *, gemm=False):
    """slices: dict with VAB,VAC,VBC,VBA,VCA,VCB,VhhhA,VhhhB,VhhhC,TA,TB,TC,
    TAB,TAC,TBC (flat arrays in the reference slice layouts)."""
    order = ["VAB", "VAC", "VBC", "VBA", "VCA", "VCB", "VhhhA", "VhhhB", "VhhhC",
             "TA", "TB", "TC", "TAB", "TAC", "TBC"]
    w = np.zeros(no ** 3)
    args = [_p(np.ascontiguousarray(slices[k], dtype=np.float64)) for k in order]
    if gemm:
        tb, vb = np.zeros(no ** 3), np.zeros(no ** 3)
        lib().oracle_doubles_gemm(no, nv, *args, _p(w), _p(tb), _p(vb))
    else:
        lib().oracle_doubles_loop(no, nv, *args, _p(w))
    return w


def energy(epsabc, no, epsi, w, z, same=False):
    f = lib().oracle_energy_same if same else lib().oracle_energy_distinct
    return f(epsabc, no, _p(np.ascontiguousarray(epsi)), _p(w), _p(z))
