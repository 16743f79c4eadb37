"""ctypes front end of the test oracle.  TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (the C restatement in fm_oracle.c) and, when built,
oracle/_ref/libfmref.so (the reference's own headers compiled by
oracle/Makefile).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import sys
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# coefficient tables only (not the fastmm package, which loads the GPU library)
sys.path.insert(0, os.path.join(ROOT, "fast-matmul_amd", "fastmm"))
import tables as _tables  # noqa: E402

MKL_PATH = "/opt/conda/lib/libmkl_rt.so.1"

_lib = None
_ref = None
_mkl = None
_reproducible = None  # MKL conditional numerical reproducibility mode set?
MKL_CBWR_COMPATIBLE = 3  # MKL_CBWR_COMPATIBLE in mkl_service.h


def _dp():
    return ctypes.POINTER(ctypes.c_double)


class FmoTable(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("k", ctypes.c_int), ("n", ctypes.c_int),
                ("R", ctypes.c_int), ("nsx", ctypes.c_int), ("ntx", ctypes.c_int),
                ("nmx", ctypes.c_int),
                ("U", ctypes.c_void_p), ("V", ctypes.c_void_p), ("W", ctypes.c_void_p),
                ("SX", ctypes.c_void_p), ("TX", ctypes.c_void_p), ("MX", ctypes.c_void_p)]


_lib_name = "liboracle.so"


def use_parallel_adds():
    """Load the build whose add loops run over OpenMP threads
    (liboracle_par.so) instead of the sequential one; for the CPU baseline,
    in its own process, before anything else loads the oracle."""
    global _lib_name
    if _lib is not None and _lib_name != "liboracle_par.so":
        raise RuntimeError("the sequential oracle is already loaded in this process")
    _lib_name = "liboracle_par.so"


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, _lib_name)
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = ctypes.CDLL(path)
        for nm in ("fmo_max_rel_diff", "fmo_max_abs_diff", "fmo_frobenius_diff"):
            getattr(_lib, nm).restype = ctypes.c_double
    return _lib


def use_mkl(threads: Optional[int] = 1, reproducible: bool = True) -> bool:
    """Install the image's MKL dgemm_/sgemm_ as the oracle's leaf BLAS (the
    library the reference links).  Returns False when MKL is absent, in which
    case the oracle keeps its own triple-loop leaf.

    reproducible: put MKL in its COMPATIBLE conditional-numerical-
    reproducibility mode (one code path on every x86 CPU) so results are bit
    identical on any host; must be chosen before the first MKL call of the
    process."""
    global _mkl, _reproducible
    if _mkl is None:
        if not os.path.exists(MKL_PATH):
            return False
        _mkl = ctypes.CDLL(MKL_PATH, mode=ctypes.RTLD_GLOBAL)
        if reproducible:
            ok = _mkl.MKL_CBWR_Set(ctypes.c_int(MKL_CBWR_COMPATIBLE)) == 0
            # MKL_CBWR_Get(MKL_CBWR_BRANCH = 1) reports the active branch
            _reproducible = ok and _mkl.MKL_CBWR_Get(ctypes.c_int(1)) == MKL_CBWR_COMPATIBLE
        else:
            _reproducible = False
    if threads is not None:
        _mkl.MKL_Set_Num_Threads(ctypes.c_int(threads))
    d = ctypes.cast(_mkl.dgemm_, ctypes.c_void_p)
    s = ctypes.cast(_mkl.sgemm_, ctypes.c_void_p)
    lib().fmo_set_blas(d, s)
    return True


def reproducible() -> bool:
    """True when the oracle's MKL runs in the cross-CPU reproducible mode."""
    return bool(_reproducible)


def use_naive():
    lib().fmo_set_blas(None, None)


def ref():
    """The reference compiled from its own sources, or None if not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libfmref.so")
        if not os.path.exists(path):
            return None
        _ref = ctypes.CDLL(path)
        for nm in ("fmref_fastmatmul_d", "fmref_fastmatmul_s", "fmref_matmul_d",
                   "fmref_max_rel_diff"):
            getattr(_ref, nm).restype = ctypes.c_double
    return _ref


# ---------------------------------------------------------------------------
class PackedTable:
    """A Table flattened to the row-major double arrays of fmo_table."""

    def __init__(self, name: str, x: float = 1e-8):
        t = _tables.get_table(name)
        self.table = t

        def arr(grp, width):
            if not grp:
                return np.zeros((0, width))
            return np.array([[c.eval(x) for c in row] for row in grp], dtype=np.float64)

        R = t.R
        self.U = np.ascontiguousarray(arr(t.U, R))
        self.V = np.ascontiguousarray(arr(t.V, R))
        self.W = np.ascontiguousarray(arr(t.W, R + len(t.MX)))
        self.SX = np.ascontiguousarray(arr(t.SX, t.m * t.k))
        self.TX = np.ascontiguousarray(arr(t.TX, t.k * t.n))
        self.MX = np.ascontiguousarray(arr(t.MX, R))
        self.c = FmoTable(t.m, t.k, t.n, R, len(t.SX), len(t.TX), len(t.MX),
                          self.U.ctypes.data, self.V.ctypes.data, self.W.ctypes.data,
                          self.SX.ctypes.data, self.TX.ctypes.data, self.MX.ctypes.data)


_packed = {}


def packed(name: str, x: float = 1e-8) -> PackedTable:
    key = (name, x)
    if key not in _packed:
        _packed[key] = PackedTable(name, x)
    return _packed[key]


def _fortran(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


def fastmatmul(name: str, A: np.ndarray, B: np.ndarray, C: Optional[np.ndarray] = None,
               steps: int = 1, alpha: float = 1.0, beta: float = 0.0, x: float = 1e-8):
    """Oracle C = alpha A B + beta C via the restated recursion.  Arrays are
    column-major (Fortran order); dtype float64 or float32 selects the path."""
    dtype = A.dtype
    A = _fortran(A, dtype)
    B = _fortran(B, dtype)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), dtype=dtype, order="F") if C is None else _fortran(C.copy(), dtype)
    pt = packed(name, x)
    L = lib()
    if dtype == np.float64:
        fn, cs = L.fmo_fastmatmul_d, ctypes.c_double
    elif dtype == np.float32:
        fn, cs = L.fmo_fastmatmul_s, ctypes.c_float
    else:
        raise TypeError(dtype)
    fn(ctypes.byref(pt.c), M, K, N, ctypes.c_void_p(A.ctypes.data), M,
       ctypes.c_void_p(B.ctypes.data), K, ctypes.c_void_p(C.ctypes.data), M,
       steps, cs(alpha), cs(beta))
    return C


def gemm(A, B, C=None, alpha=1.0, beta=0.0):
    dtype = A.dtype
    A = _fortran(A, dtype)
    B = _fortran(B, dtype)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), dtype=dtype, order="F") if C is None else _fortran(C.copy(), dtype)
    L = lib()
    fn, cs = (L.fmo_gemm_d, ctypes.c_double) if dtype == np.float64 else (L.fmo_gemm_s, ctypes.c_float)
    fn(M, K, N, ctypes.c_void_p(A.ctypes.data), M, ctypes.c_void_p(B.ctypes.data), K,
       ctypes.c_void_p(C.ctypes.data), M, cs(alpha), cs(beta))
    return C


def ref_fastmatmul(name, A, B, C=None, steps=1, alpha=1.0, beta=0.0, x=1e-8):
    r = ref()
    if r is None:
        return None
    dtype = A.dtype
    A = _fortran(A.copy(), dtype)
    B = _fortran(B.copy(), dtype)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), dtype=dtype, order="F") if C is None else _fortran(C.copy(), dtype)
    if dtype == np.float64:
        fn, cs = r.fmref_fastmatmul_d, ctypes.c_double
    else:
        fn, cs = r.fmref_fastmatmul_s, ctypes.c_float
    ms = fn(name.encode(), M, K, N, ctypes.c_void_p(A.ctypes.data), M,
            ctypes.c_void_p(B.ctypes.data), K, ctypes.c_void_p(C.ctypes.data), M,
            steps, ctypes.c_double(x), cs(alpha), cs(beta))
    if ms < 0:
        raise KeyError(name)
    return C


def quad_matmul(A, B):
    A = _fortran(A, np.float64)
    B = _fortran(B, np.float64)
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), order="F")
    lib().fmo_quad_matmul(M, K, N, ctypes.c_void_p(A.ctypes.data), M,
                          ctypes.c_void_p(B.ctypes.data), K, ctypes.c_void_p(C.ctypes.data), M)
    return C


def quad_entries(A, B, rows, cols):
    """Quad-precision values of (A B)[rows[e], cols[e]] and sum |a||b|."""
    A = _fortran(A, np.float64)
    B = _fortran(B, np.float64)
    M, K = A.shape
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    out = np.zeros(len(rows))
    absab = np.zeros(len(rows))
    lib().fmo_quad_entries(K, ctypes.c_void_p(A.ctypes.data), M, ctypes.c_void_p(B.ctypes.data),
                           K, ctypes.c_void_p(rows.ctypes.data), ctypes.c_void_p(cols.ctypes.data),
                           len(rows), ctypes.c_void_p(out.ctypes.data),
                           ctypes.c_void_p(absab.ctypes.data))
    return out, absab


def max_rel_diff(A, B):
    """MaxRelativeDiff(A, B): max |a-b|/|a|, dividing by the FIRST argument."""
    A = _fortran(A, np.float64)
    B = _fortran(B, np.float64)
    m, n = A.shape
    return lib().fmo_max_rel_diff(m, n, ctypes.c_void_p(A.ctypes.data), m,
                                  ctypes.c_void_p(B.ctypes.data), m)


def max_abs_diff(A, B):
    A = _fortran(A, np.float64)
    B = _fortran(B, np.float64)
    m, n = A.shape
    return lib().fmo_max_abs_diff(m, n, ctypes.c_void_p(A.ctypes.data), m,
                                  ctypes.c_void_p(B.ctypes.data), m)


def frobenius_diff(A, B):
    A = _fortran(A, np.float64)
    B = _fortran(B, np.float64)
    m, n = A.shape
    return lib().fmo_frobenius_diff(m, n, ctypes.c_void_p(A.ctypes.data), m,
                                    ctypes.c_void_p(B.ctypes.data), m)


def normwise_err(Cf, D, A, B):
    """max|C - D| / (max|A| * max|B| * K): the normwise error measure."""
    K = A.shape[1]
    den = float(np.abs(A).max()) * float(np.abs(B).max()) * K
    return float(np.abs(np.asarray(Cf, dtype=np.float64) - D).max()) / den
