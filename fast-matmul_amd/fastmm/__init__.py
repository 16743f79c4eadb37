"""fastmm -- fast (Strassen-like) matrix multiplication on one MI355X.

Host-side mirror of the reference's API over the C ABI of libfastmm.so:

* ``Matrix``: the column-major view type of linalg/linalg.hpp:41-192
  (``m``, ``n``, ``stride``, ``multiplier``, ``Subblock``, ``Submatrix``),
  backed by a GPU tensor;
* one namespace object per algorithm with ``FastMatmul(A, B, C, num_steps,
  x=1e-8, alpha=1, beta=0) -> ms`` (algorithms/strassen.hpp:662-722), reached
  as ``fastmm.strassen``, ``fastmm.grey424_26_206`` ... or ``fastmm.alg(name)``;
* ``RunAlgorithm(ENUM_NAME, A, B, C, num_steps)`` and ``MatMul(A, B, C, beta)``
  (algorithms/all_algorithms.hpp:93-243, linalg/linalg.hpp:261-269);
* tensor-level ``fastmatmul`` / ``gemm`` for callers that hold torch tensors.

All arithmetic runs in the HIP kernels of libfastmm.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib
from . import tables
from ._lib import LIB, FastMMError, check  # noqa: F401

__all__ = ["Matrix", "fastmatmul", "fastmatmul_scaled", "gemm", "MatMul", "RunAlgorithm", "alg", "algorithm_names",
           "workspace_bytes", "release_caches", "options", "Plan", "scaled_workspace_bytes"]

_DT = {torch.float64: _lib.FM_F64, torch.float32: _lib.FM_F32}


def algorithm_names():
    return [LIB.fm_algorithm_name(i).decode() for i in range(LIB.fm_num_algorithms())]


def _is_col_major(t: torch.Tensor) -> bool:
    m, n = t.shape
    if m <= 1:  # one row: any column stride >= 1 is a valid leading dimension
        return n <= 1 or t.stride(1) >= 1
    if n <= 1:  # one column: rows must be contiguous
        return t.stride(0) == 1
    return t.stride(0) == 1 and t.stride(1) >= m


def _col_major(t: torch.Tensor, name: str) -> torch.Tensor:
    if t.dim() != 2:
        raise ValueError("%s must be 2-D" % name)
    if _is_col_major(t):
        return t
    raise ValueError("%s must be a column-major view (stride(0) == 1); use "
                     "fastmm.col_major(%s)" % (name, name))


def col_major(t: torch.Tensor) -> torch.Tensor:
    """Column-major copy of a 2-D tensor (no copy if it already is one)."""
    if _is_col_major(t):
        return t
    return t.t().contiguous().t()


def empty(m: int, n: int, dtype=torch.float64, device="cuda") -> torch.Tensor:
    """Uninitialised column-major m x n tensor."""
    return torch.empty((n, m), dtype=dtype, device=device).t()


def _ld(t: torch.Tensor) -> int:
    m, n = t.shape
    if n > 1:
        return max(1, m, t.stride(1)) if m > 1 else max(1, t.stride(1))
    return max(1, m)


def _stream_handle(t: torch.Tensor):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _operands(A: torch.Tensor, B: torch.Tensor, C: Optional[torch.Tensor], beta: float):
    """Check A (m x k), B (k x n) and C (m x n, allocated when None): 2-D
    column-major views of one float64/float32 dtype on one GPU.  Every entry
    point runs this before anything reaches the library, so a mismatched
    call raises here instead of letting a kernel write past C."""
    A = _col_major(A, "A")
    B = _col_major(B, "B")
    m, k = A.shape
    k2, n = B.shape
    if k2 != k:
        raise ValueError("inner dimensions differ: A is %d x %d, B is %d x %d" % (m, k, k2, n))
    if A.dtype not in _DT:
        raise TypeError("A must be float64 or float32, not %s" % A.dtype)
    if B.dtype != A.dtype:
        raise TypeError("B must have A's dtype %s, not %s" % (A.dtype, B.dtype))
    if not A.is_cuda or not B.is_cuda:
        raise ValueError("A and B must be GPU tensors")
    if B.device != A.device:
        raise ValueError("A and B must be on one device (%s vs %s)" % (A.device, B.device))
    if C is None:
        if beta != 0.0:
            raise ValueError("beta != 0 needs an input C")
        C = empty(m, n, A.dtype, A.device)
    C = _col_major(C, "C")
    if tuple(C.shape) != (m, n):
        raise ValueError("C must be %d x %d, not %d x %d" % (m, n, C.shape[0], C.shape[1]))
    if C.dtype != A.dtype:
        raise TypeError("C must have A's dtype %s, not %s" % (A.dtype, C.dtype))
    if not C.is_cuda or C.device != A.device:
        raise ValueError("C must be a GPU tensor on A's device %s" % A.device)
    return A, B, C, m, k, n


def options(**kw) -> "_lib.FmOptions":
    """An fm_options block: the process defaults with the given fields
    replaced (gemm_tile, dfs_levels, workspace_limit, split_remainder,
    stage_fusion, ozaki_slices)."""
    o = _lib.FmOptions()
    LIB.fm_options_default(ctypes.byref(o))
    names = {f[0] for f in _lib.FmOptions._fields_}
    for key, val in kw.items():
        if key not in names:
            raise TypeError("unknown option %r" % key)
        setattr(o, key, int(val))
    return o


def _options_ref(opts):
    if opts is None:
        return None
    if isinstance(opts, dict):
        opts = options(**opts)
    return ctypes.byref(opts)


def _workspace(workspace, A):
    if workspace is None:
        return None, 0
    if not workspace.is_cuda or workspace.device != A.device:
        raise ValueError("workspace must be a GPU tensor on A's device")
    return ctypes.c_void_p(workspace.data_ptr()), workspace.numel() * workspace.element_size()


def fastmatmul(alg_name: str, A: torch.Tensor, B: torch.Tensor, C: Optional[torch.Tensor] = None,
               num_steps: int = 1, x: float = 1e-8, alpha: float = 1.0, beta: float = 0.0,
               mode: str = "plain", workspace: Optional[torch.Tensor] = None,
               timed: bool = False, opts=None):
    """C := alpha * A @ B + beta * C with ``num_steps`` levels of ``alg_name``.

    A, B, C are column-major GPU tensors of one dtype (float64 or float32).
    ``opts`` (a dict or ``options(...)``) overrides the process defaults for
    this call only.  Returns C, or (C, ms) when ``timed``."""
    A, B, C, m, k, n = _operands(A, B, C, beta)
    ws_ptr, ws_bytes = _workspace(workspace, A)
    ms = ctypes.c_float(0.0)
    st = LIB.fm_fastmatmul_ex(alg_name.encode(), _DT[A.dtype], _lib.MODES[mode], m, k, n,
                              ctypes.c_void_p(A.data_ptr()), _ld(A), ctypes.c_void_p(B.data_ptr()),
                              _ld(B), ctypes.c_void_p(C.data_ptr()), _ld(C), int(num_steps),
                              float(x), float(alpha), float(beta), _options_ref(opts),
                              _stream_handle(A), ws_ptr, ws_bytes,
                              ctypes.byref(ms) if timed else None)
    check(st)
    return (C, ms.value) if timed else C


class Plan:
    """One multiply's launch sequence, built once (fm_plan_create) and run
    for any A, B, C of the planned shape and leading dimensions
    (fm_plan_run).  ``alg`` None plans the classical product."""

    def __init__(self, alg: Optional[str], m: int, k: int, n: int, num_steps: int = 1,
                 dtype=torch.float64, mode: str = "plain", lda: Optional[int] = None,
                 ldb: Optional[int] = None, ldc: Optional[int] = None, x: float = 1e-8,
                 alpha: float = 1.0, beta: float = 0.0, opts=None):
        if dtype not in _DT:
            raise TypeError("dtype must be float64 or float32")
        self.shape = (int(m), int(k), int(n))
        self.dtype = dtype
        self.lds = (int(lda or max(1, m)), int(ldb or max(1, k)), int(ldc or max(1, m)))
        h = ctypes.c_void_p()
        check(LIB.fm_plan_create(alg.encode() if alg else None, _DT[dtype], _lib.MODES[mode],
                                 int(m), int(k), int(n), self.lds[0], self.lds[1], self.lds[2],
                                 int(num_steps), float(x), float(alpha), float(beta),
                                 _options_ref(opts), ctypes.byref(h)))
        self._h = h

    def workspace_bytes(self) -> int:
        out = ctypes.c_size_t(0)
        check(LIB.fm_plan_workspace(self._h, ctypes.byref(out)))
        return out.value

    def run(self, A: torch.Tensor, B: torch.Tensor, C: torch.Tensor,
            workspace: Optional[torch.Tensor] = None, timed: bool = False):
        A, B, C, m, k, n = _operands(A, B, C, 0.0)
        if (m, k, n) != self.shape or A.dtype != self.dtype:
            raise ValueError("plan is for %s %s, not %s %s" % (self.shape, self.dtype, (m, k, n), A.dtype))
        if (_ld(A), _ld(B), _ld(C)) != self.lds:
            raise ValueError("plan is for leading dimensions %s, not %s" % (self.lds, (_ld(A), _ld(B), _ld(C))))
        ws_ptr, ws_bytes = _workspace(workspace, A)
        ms = ctypes.c_float(0.0)
        check(LIB.fm_plan_run(self._h, ctypes.c_void_p(A.data_ptr()), ctypes.c_void_p(B.data_ptr()),
                              ctypes.c_void_p(C.data_ptr()), _stream_handle(A), ws_ptr, ws_bytes,
                              ctypes.byref(ms) if timed else None))
        return (C, ms.value) if timed else C

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and LIB is not None:
            LIB.fm_plan_destroy(h)
            self._h = None


def fastmatmul_scaled(alg_name: str, A: torch.Tensor, B: torch.Tensor,
                      C: Optional[torch.Tensor] = None, num_steps: int = 1, x: float = 1e-8,
                      alpha: float = 1.0, beta: float = 0.0, mode: str = "plain",
                      scaling: str = "outside_inside", scaling_steps: int = 1,
                      timed: bool = False):
    """fastmatmul on diagonally scaled copies of A and B, scaled back
    (scaling/scaling.hpp:105-196): C := alpha diag(r) (A' B') diag(s) + beta C.
    ``scaling`` is one of outside_inside, inside_outside, inside, outside."""
    A, B, C, m, k, n = _operands(A, B, C, beta)
    if scaling not in _lib.SCALINGS:
        raise ValueError("scaling must be one of %s" % sorted(_lib.SCALINGS))
    ms = ctypes.c_float(0.0)
    check(LIB.fm_fastmatmul_scaled(alg_name.encode(), _DT[A.dtype], _lib.MODES[mode], m, k, n,
                                   ctypes.c_void_p(A.data_ptr()), _ld(A),
                                   ctypes.c_void_p(B.data_ptr()), _ld(B),
                                   ctypes.c_void_p(C.data_ptr()), _ld(C), int(num_steps), float(x),
                                   float(alpha), float(beta), _lib.SCALINGS[scaling],
                                   int(scaling_steps), _stream_handle(A), None, 0,
                                   ctypes.byref(ms) if timed else None))
    return (C, ms.value) if timed else C


def gemm(A: torch.Tensor, B: torch.Tensor, C: Optional[torch.Tensor] = None, alpha: float = 1.0,
         beta: float = 0.0, mode: str = "plain") -> torch.Tensor:
    """Classical C := alpha * A @ B + beta * C on the library's leaf kernel."""
    A, B, C, m, k, n = _operands(A, B, C, beta)
    st = LIB.fm_gemm(_DT[A.dtype], _lib.MODES[mode], m, k, n, ctypes.c_void_p(A.data_ptr()), _ld(A),
                     ctypes.c_void_p(B.data_ptr()), _ld(B), ctypes.c_void_p(C.data_ptr()), _ld(C),
                     float(alpha), float(beta), _stream_handle(A))
    check(st)
    return C


def workspace_bytes(alg_name: str, m: int, k: int, n: int, num_steps: int,
                    dtype=torch.float64, mode: str = "plain") -> int:
    out = ctypes.c_size_t(0)
    check(LIB.fm_workspace_query(alg_name.encode(), _DT[dtype], _lib.MODES[mode], m, k, n,
                                 num_steps, ctypes.byref(out)))
    return out.value


def scaled_workspace_bytes(alg_name: str, m: int, k: int, n: int, num_steps: int,
                           beta_nonzero: bool = False, dtype=torch.float64,
                           mode: str = "plain") -> int:
    """Workspace fastmatmul_scaled needs (fm_scaled_workspace_query)."""
    out = ctypes.c_size_t(0)
    check(LIB.fm_scaled_workspace_query(alg_name.encode(), _DT[dtype], _lib.MODES[mode], m, k, n,
                                        num_steps, 1 if beta_nonzero else 0, ctypes.byref(out)))
    return out.value


def release_caches():
    LIB.fm_release_caches()


def release_workspace(stream=None):
    """Free the library's cached workspace of ``stream`` (default: the
    current stream of the current device)."""
    s = stream if stream is not None else torch.cuda.current_stream()
    LIB.fm_release_workspace(ctypes.c_void_p(s.cuda_stream))


def set_gemm_tile(cfg: int = -1):
    """Force a leaf-GEMM tile configuration (-1: automatic choice)."""
    check(LIB.fm_set_gemm_tile(int(cfg)))


def plan_info(alg_name: str, m: int, k: int, n: int, num_steps: int = 1,
              dtype=torch.float64, mode: str = "plain") -> dict:
    """Structure of the plan fastmatmul would run for this shape (built on
    the host; nothing is launched)."""
    info = _lib.FmPlanInfo()
    check(LIB.fm_plan_query(alg_name.encode(), _DT[dtype], _lib.MODES[mode], int(m), int(k),
                            int(n), int(num_steps), ctypes.byref(info)))
    kinds = _lib.STAGES[:5]  # plan stages; plane preparation counts with LEAF
    return {"launches": {nm: info.launches[i] for i, nm in enumerate(kinds)},
            "stage_bytes": {nm: info.stage_bytes[i] for i, nm in enumerate(kinds)},
            "two_level_stages": info.two_level_stages,
            "leaf": [{"cfg": info.leaf_cfg[i], "products": info.leaf_products[i],
                      "tiles": info.leaf_tiles[i]} for i in range(min(2, info.launches[3]))],
            "workspace_bytes": info.workspace_bytes, "dfs_levels": info.dfs_levels,
            "prep_bytes": info.prep_bytes}


def set_schedule(dfs_levels: int = 0, workspace_limit: int = 0, split_remainder: bool = True):
    """Recursion schedule for plans built afterwards: the top ``dfs_levels``
    levels depth-first (one child sub-plan at a time, shared workspace), more
    depth-first levels while the workspace exceeds ``workspace_limit`` bytes
    (0: no limit), and the leaf launch's part-empty last round of tiles in a
    second launch with smaller tiles (``split_remainder``)."""
    check(LIB.fm_set_schedule(int(dfs_levels), int(workspace_limit), 1 if split_remainder else 0))


def set_stage_fusion(on=True):
    """Two-level S/T stages where the shape allows them (bit-identical
    results).  True or 1: on, and with split_bf16 leaves the last S/T stage
    writes the leaves' slice planes; 2: on, planes cut by a separate pass;
    False or 0: one level at a time."""
    LIB.fm_set_stage_fusion(2 if on == 2 else 1 if on else 0)


def set_ozaki_slices(slices: int = 8):
    """Digit count of the ozaki_int8 leaf mode (6, 7 or 8)."""
    check(LIB.fm_set_ozaki_slices(int(slices)))


def profile_enable(on: bool = True):
    """Bracket every library launch with HIP events on its stream."""
    LIB.fm_profile_enable(1 if on else 0)


_PEAK_KINDS = {"fp64": 0, "fp32": 1, "bf16": 2, "int8": 3}
# published dense peaks of MI355X (TFLOP/s, TOP/s for int8), for sizing the probe
_PUBLISHED = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0, "int8": 5000.0}


def probe_mfma_peak(kind: str, seconds: float = 0.05) -> float:
    """Measured matrix-unit peak (tera-ops/s) of one MFMA kind on this GPU,
    running about ``seconds`` (fm_probe_mfma_peak)."""
    # operations per probe iteration: 256 CUs x 8 work groups x 4 waves x 8 chains
    per_iter = {"fp64": 2048, "fp32": 2048, "bf16": 16384, "int8": 65536}[kind] * 8 * 4 * 8 * 256
    iters = max(1, int(_PUBLISHED[kind] * 1e12 * seconds / per_iter))
    out = ctypes.c_double(0.0)
    check(LIB.fm_probe_mfma_peak(_PEAK_KINDS[kind], iters, ctypes.byref(out)))
    return out.value


def probe_stream(reads: int, writes: int, elems: int) -> float:
    """Measured HBM rate (GB/s) of a stage kernel's access pattern: each lane
    reads one element position of ``reads`` FP64 slabs of ``elems`` elements
    and writes it to ``writes`` slabs (fm_probe_stream)."""
    out = ctypes.c_double(0.0)
    check(LIB.fm_probe_stream(int(reads), int(writes), int(elems), ctypes.byref(out)))
    return out.value


def profile_read():
    """{stage: (ms_total, launches)} accumulated since the last read."""
    p = _lib.FmProfile()
    check(LIB.fm_profile_read(ctypes.byref(p)))
    return {nm: (p.ms[i], p.launches[i]) for i, nm in enumerate(_lib.STAGES)}


class Batch:
    """Many independent C_i := alpha_i A_i B_i + beta_i C_i in one launch
    sequence (one grouped launch of all base-case products), optionally
    replayed from a captured hipGraph.  ``problems`` is a list of dicts with
    keys A, B, C (column-major GPU tensors of one dtype) and optional alg,
    num_steps, alpha, beta; a missing alg / num_steps is autotuned."""

    def __init__(self, problems, mode: str = "plain", candidates=None, max_steps: int = 2,
                 graph: bool = True, workspace: Optional[torch.Tensor] = None):
        if not problems:
            raise ValueError("empty batch")
        self._keep = []
        arr = (_lib.FmProblem * len(problems))()
        dtype = None
        device = None
        for i, p in enumerate(problems):
            try:
                A, B, C, _, _, _ = _operands(p["A"], p["B"], p["C"], float(p.get("beta", 0.0)))
            except (ValueError, TypeError) as e:
                raise type(e)("problem %d: %s" % (i, e)) from None
            dtype = dtype or A.dtype
            device = device or A.device
            if A.dtype != dtype:
                raise TypeError("problem %d: all matrices of a batch share one dtype" % i)
            if A.device != device:
                raise ValueError("problem %d: all matrices of a batch share one device" % i)
            self._keep += [A, B, C]
            alg = p.get("alg")
            arr[i] = _lib.FmProblem(alg.encode() if alg else None, int(p.get("num_steps", -1)),
                                    A.shape[0], A.shape[1], B.shape[1], A.data_ptr(), _ld(A),
                                    B.data_ptr(), _ld(B), C.data_ptr(), _ld(C),
                                    float(p.get("alpha", 1.0)), float(p.get("beta", 0.0)))
        cands = None
        ncand = 0
        if candidates:
            cands = (ctypes.c_char_p * len(candidates))(*[c.encode() for c in candidates])
            ncand = len(candidates)
        ws_ptr, ws_bytes = None, 0
        if workspace is not None:
            ws_ptr = ctypes.c_void_p(workspace.data_ptr())
            ws_bytes = workspace.numel() * workspace.element_size()
            self._keep.append(workspace)
        h = ctypes.c_void_p()
        check(LIB.fm_batch_create(_DT[dtype], _lib.MODES[mode], arr, len(problems), cands, ncand,
                                  max_steps, 1 if graph else 0, ws_ptr, ws_bytes, ctypes.byref(h)))
        self._h = h
        self.count = len(problems)
        self._device = self._keep[0].device

    def run(self, timed: bool = False):
        ms = ctypes.c_float(0.0)
        s = ctypes.c_void_p(torch.cuda.current_stream(self._device).cuda_stream)
        check(LIB.fm_batch_run(self._h, s, ctypes.byref(ms) if timed else None))
        return ms.value if timed else None

    def choice(self, i: int):
        a = ctypes.c_char_p()
        st = ctypes.c_int()
        check(LIB.fm_batch_choice(self._h, i, ctypes.byref(a), ctypes.byref(st)))
        return a.value.decode(), st.value

    def workspace_bytes(self) -> int:
        return LIB.fm_batch_workspace_bytes(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and LIB is not None:  # module globals may be gone at interpreter exit
            LIB.fm_batch_destroy(h)
            self._h = None


def autotune_load(path: str):
    """Use the cost-model table at ``path`` (as written by fastmm.calibrate)."""
    check(LIB.fm_autotune_load(path.encode()))


def autotune_table() -> str:
    """Path of the cost-model table in use ("built-in" without one)."""
    return LIB.fm_autotune_table().decode()


def autotune_estimate(alg: Optional[str], m: int, k: int, n: int, num_steps: int,
                      dtype=torch.float32, mode: str = "plain") -> float:
    """The autotuner's cost-model estimate in microseconds."""
    us = ctypes.c_double(0.0)
    check(LIB.fm_autotune_estimate(_DT[dtype], _lib.MODES[mode], alg.encode() if alg else None,
                                   m, k, n, num_steps, ctypes.byref(us)))
    return us.value


# ---------------------------------------------------------------------------
# Reference-shaped API
# ---------------------------------------------------------------------------
class Matrix:
    """Column-major view with a stride and a scalar multiplier
    (linalg/linalg.hpp:41-192), stored in a GPU tensor.

    ``Matrix(m, n)`` allocates (uninitialised, like ``new Scalar[m*n]``);
    ``Matrix(tensor)`` wraps a column-major 2-D tensor without copying."""

    def __init__(self, m, n: Optional[int] = None, multiplier: float = 1.0,
                 dtype=torch.float64, device="cuda"):
        if isinstance(m, torch.Tensor):
            self.data = _col_major(m, "data")
        else:
            n = m if n is None else n
            self.data = empty(m, n, dtype, device)
        self._multiplier = multiplier

    @classmethod
    def from_numpy(cls, a, device="cuda"):
        t = torch.as_tensor(a).to(device)
        return cls(col_major(t))

    def numpy(self):
        return self.data.cpu().numpy()

    def m(self):
        return self.data.shape[0]

    def n(self):
        return self.data.shape[1]

    def stride(self):
        return _ld(self.data)

    def multiplier(self):
        return self._multiplier

    def set_multiplier(self, v):
        self._multiplier = v

    def UpdateMultiplier(self, v):
        self._multiplier *= v

    def __call__(self, i, j):
        return self.data[i, j].item()

    def Submatrix(self, start_row, start_col, num_rows, num_cols):
        return Matrix(self.data[start_row:start_row + num_rows, start_col:start_col + num_cols],
                      multiplier=self._multiplier)

    def Subblock(self, num_block_rows, num_block_cols, row_ind, col_ind):
        """Block (row_ind, col_ind), 1-indexed, of floor size
        (linalg/linalg.hpp:97-102, 179-184)."""
        rs = self.m() // num_block_rows
        cs = self.n() // num_block_cols
        return self.Submatrix(rs * (row_ind - 1), cs * (col_ind - 1), rs, cs)


class _Algorithm:
    """The namespace of one generated algorithm (e.g. ``strassen``)."""

    def __init__(self, name):
        self.name = name
        m, k, n, r = (ctypes.c_int() for _ in range(4))
        check(LIB.fm_algorithm_info(name.encode(), ctypes.byref(m), ctypes.byref(k),
                                    ctypes.byref(n), ctypes.byref(r)))
        self.dims = (m.value, k.value, n.value)
        self.rank = r.value

    def FastMatmul(self, A: Matrix, B: Matrix, C: Matrix, num_steps: int, x: float = 1e-8,
                   alpha: float = 1.0, beta: float = 0.0) -> float:
        """C := alpha A B + beta C; returns elapsed device ms.  The leaf
        multiplier is alpha * B.multiplier * C.multiplier as in
        algorithms/strassen.hpp:669,380-383; no Matrix is modified."""
        eff = alpha * B.multiplier() * C.multiplier()
        _, ms = fastmatmul(self.name, A.data, B.data, C.data, num_steps, x, eff, beta, timed=True)
        return ms

    def __repr__(self):
        return "<fastmm algorithm %s <%d,%d,%d;%d>>" % ((self.name,) + self.dims + (self.rank,))


_algs = {}


def alg(name: str) -> _Algorithm:
    if name not in _algs:
        _algs[name] = _Algorithm(name)
    return _algs[name]


def MatMul(A: Matrix, B: Matrix, C: Matrix, beta: float = 0.0):
    """C <- C.multiplier * A B + beta C (linalg/linalg.hpp:261-269)."""
    gemm(A.data, B.data, C.data, C.multiplier(), beta)


def RunAlgorithm(algorithm: str, A: Matrix, B: Matrix, C: Matrix, num_steps: int) -> float:
    """Dispatch by the reference's enum name (algorithms/all_algorithms.hpp:93-243)."""
    if algorithm not in tables.ENUM_TO_NAMESPACE:
        raise KeyError("unknown algorithm %r" % algorithm)
    ns = tables.ENUM_TO_NAMESPACE[algorithm]
    if algorithm == "MKL":
        return alg(ns).FastMatmul(A, B, C, 0)
    return alg(ns).FastMatmul(A, B, C, num_steps, tables.default_x(algorithm, num_steps))


def __getattr__(name):
    if name.startswith("_"):
        raise AttributeError(name)
    try:
        return alg(name)
    except KeyError:
        raise AttributeError(name) from None
