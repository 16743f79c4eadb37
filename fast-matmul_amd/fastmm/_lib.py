"""ctypes binding of libfastmm.so (the C ABI declared in include/fastmm.h).

The library is built in-tree by fast-matmul_amd/build.py.  There is no
fallback: if the library is missing, importing the package raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfastmm.so")

FM_OK = 0
FM_ERR_INVALID_ARG = 1
FM_ERR_UNKNOWN_ALG = 2
FM_ERR_SHAPE = 3
FM_ERR_WORKSPACE = 4
FM_ERR_HIP = 5
FM_ERR_UNSUPPORTED = 6

FM_F64 = 0
FM_F32 = 1

MODES = {"plain": 0, "split_bf16": 1, "ozaki_int8": 2}

# every symbol include/fastmm.h declares
EXPORTS = (
    "fm_fastmatmul", "fm_workspace_query", "fm_gemm", "fm_num_algorithms",
    "fm_algorithm_name", "fm_algorithm_info", "fm_status_string", "fm_last_error",
    "fm_release_caches", "fm_profile_enable", "fm_profile_read", "fm_set_gemm_tile",
    "fm_set_ozaki_slices", "fm_batch_create", "fm_batch_run", "fm_batch_choice",
    "fm_batch_workspace_bytes", "fm_batch_destroy", "fm_autotune_estimate",
    "fm_set_stage_fusion", "fm_set_schedule", "fm_plan_query", "fm_fastmatmul_scaled",
    "fm_options_default", "fm_fastmatmul_ex", "fm_plan_create", "fm_plan_run",
    "fm_plan_workspace", "fm_plan_destroy", "fm_scaled_workspace_query", "fm_release_workspace",
    "fm_fastmatmul_noalloc", "fm_probe_mfma_peak", "fm_autotune_load", "fm_autotune_table",
    "fm_probe_stream",
)

# the fixed scratch array of fm_fastmatmul_noalloc (FM_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 256 << 20

# scaling rules of scaling/scaling.hpp:17-22
SCALINGS = {"outside_inside": 0, "inside_outside": 1, "inside": 2, "outside": 3}

STAGES = ("S", "T", "W", "LEAF", "PEEL", "PREP")


class FmPlanInfo(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int * 5), ("stage_bytes", ctypes.c_double * 5),
                ("two_level_stages", ctypes.c_int),
                ("leaf_cfg", ctypes.c_int * 2), ("leaf_products", ctypes.c_int * 2),
                ("leaf_tiles", ctypes.c_int * 2), ("workspace_bytes", ctypes.c_size_t),
                ("dfs_levels", ctypes.c_int), ("prep_bytes", ctypes.c_double)]


class FmOptions(ctypes.Structure):
    _fields_ = [("gemm_tile", ctypes.c_int), ("dfs_levels", ctypes.c_int),
                ("workspace_limit", ctypes.c_size_t), ("split_remainder", ctypes.c_int),
                ("stage_fusion", ctypes.c_int), ("ozaki_slices", ctypes.c_int)]


class FmProfile(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * 6), ("launches", ctypes.c_int64 * 6)]


class FmProblem(ctypes.Structure):
    _fields_ = [("alg", ctypes.c_char_p), ("num_steps", ctypes.c_int), ("m", ctypes.c_int),
                ("k", ctypes.c_int), ("n", ctypes.c_int), ("A", ctypes.c_void_p),
                ("lda", ctypes.c_int64), ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_int64), ("alpha", ctypes.c_double),
                ("beta", ctypes.c_double)]


class FastMMError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libfastmm.so is not built (expected at %s); run "
            "`python fast-matmul_amd/build.py` or __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    i64, vp, c_int, c_double = ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.fm_fastmatmul.restype = c_int
    lib.fm_fastmatmul.argtypes = [ctypes.c_char_p, c_int, c_int, c_int, c_int, c_int,
                                  vp, i64, vp, i64, vp, i64, c_int, c_double, c_double,
                                  c_double, vp, vp, ctypes.c_size_t,
                                  ctypes.POINTER(ctypes.c_float)]
    lib.fm_fastmatmul_ex.restype = c_int
    lib.fm_fastmatmul_ex.argtypes = [ctypes.c_char_p, c_int, c_int, c_int, c_int, c_int,
                                     vp, i64, vp, i64, vp, i64, c_int, c_double, c_double,
                                     c_double, ctypes.POINTER(FmOptions), vp, vp, ctypes.c_size_t,
                                     ctypes.POINTER(ctypes.c_float)]
    lib.fm_fastmatmul_noalloc.restype = c_int
    lib.fm_fastmatmul_noalloc.argtypes = [ctypes.c_char_p, c_int, c_int, c_int, c_int, c_int,
                                          vp, i64, vp, i64, vp, i64, c_int, c_double, c_double,
                                          c_double, vp, ctypes.POINTER(ctypes.c_float)]
    lib.fm_autotune_load.restype = c_int
    lib.fm_autotune_load.argtypes = [ctypes.c_char_p]
    lib.fm_autotune_table.restype = ctypes.c_char_p
    lib.fm_autotune_table.argtypes = []
    lib.fm_probe_mfma_peak.restype = c_int
    lib.fm_probe_mfma_peak.argtypes = [c_int, c_int, ctypes.POINTER(c_double)]
    lib.fm_probe_stream.restype = c_int
    lib.fm_probe_stream.argtypes = [c_int, c_int, i64, ctypes.POINTER(c_double)]
    lib.fm_options_default.restype = None
    lib.fm_options_default.argtypes = [ctypes.POINTER(FmOptions)]
    lib.fm_plan_create.restype = c_int
    lib.fm_plan_create.argtypes = [ctypes.c_char_p, c_int, c_int, c_int, c_int, c_int, i64, i64,
                                   i64, c_int, c_double, c_double, c_double,
                                   ctypes.POINTER(FmOptions), ctypes.POINTER(vp)]
    lib.fm_plan_run.restype = c_int
    lib.fm_plan_run.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_size_t,
                                ctypes.POINTER(ctypes.c_float)]
    lib.fm_plan_workspace.restype = c_int
    lib.fm_plan_workspace.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    lib.fm_plan_destroy.restype = None
    lib.fm_plan_destroy.argtypes = [vp]
    lib.fm_scaled_workspace_query.restype = c_int
    lib.fm_scaled_workspace_query.argtypes = [ctypes.c_char_p, c_int, c_int, c_int, c_int, c_int,
                                              c_int, c_int, ctypes.POINTER(ctypes.c_size_t)]
    lib.fm_release_workspace.restype = None
    lib.fm_release_workspace.argtypes = [vp]
    lib.fm_fastmatmul_scaled.restype = c_int
    lib.fm_fastmatmul_scaled.argtypes = [ctypes.c_char_p, c_int, c_int, c_int, c_int, c_int,
                                         vp, i64, vp, i64, vp, i64, c_int, c_double, c_double,
                                         c_double, c_int, c_int, vp, vp, ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_float)]
    lib.fm_workspace_query.restype = c_int
    lib.fm_workspace_query.argtypes = [ctypes.c_char_p, c_int, c_int, c_int, c_int, c_int,
                                       c_int, ctypes.POINTER(ctypes.c_size_t)]
    lib.fm_gemm.restype = c_int
    lib.fm_gemm.argtypes = [c_int, c_int, c_int, c_int, c_int, vp, i64, vp, i64, vp, i64,
                            c_double, c_double, vp]
    lib.fm_num_algorithms.restype = c_int
    lib.fm_algorithm_name.restype = ctypes.c_char_p
    lib.fm_algorithm_name.argtypes = [c_int]
    lib.fm_algorithm_info.restype = c_int
    lib.fm_algorithm_info.argtypes = [ctypes.c_char_p] + [ctypes.POINTER(c_int)] * 4
    lib.fm_status_string.restype = ctypes.c_char_p
    lib.fm_status_string.argtypes = [c_int]
    lib.fm_last_error.restype = ctypes.c_char_p
    lib.fm_release_caches.restype = None
    lib.fm_profile_enable.restype = None
    lib.fm_profile_enable.argtypes = [c_int]
    lib.fm_profile_read.restype = c_int
    lib.fm_profile_read.argtypes = [ctypes.POINTER(FmProfile)]
    lib.fm_set_gemm_tile.restype = c_int
    lib.fm_set_gemm_tile.argtypes = [c_int]
    lib.fm_set_ozaki_slices.restype = c_int
    lib.fm_set_ozaki_slices.argtypes = [c_int]
    lib.fm_set_stage_fusion.restype = None
    lib.fm_set_stage_fusion.argtypes = [c_int]
    lib.fm_set_schedule.restype = c_int
    lib.fm_plan_query.restype = c_int
    lib.fm_plan_query.argtypes = [ctypes.c_char_p, c_int, c_int, c_int, c_int, c_int, c_int,
                                  ctypes.POINTER(FmPlanInfo)]
    lib.fm_set_schedule.argtypes = [c_int, ctypes.c_size_t, c_int]
    lib.fm_batch_create.restype = c_int
    lib.fm_batch_create.argtypes = [c_int, c_int, ctypes.POINTER(FmProblem), c_int,
                                    ctypes.POINTER(ctypes.c_char_p), c_int, c_int, c_int, vp,
                                    ctypes.c_size_t, ctypes.POINTER(vp)]
    lib.fm_batch_run.restype = c_int
    lib.fm_batch_run.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_float)]
    lib.fm_batch_choice.restype = c_int
    lib.fm_batch_choice.argtypes = [vp, c_int, ctypes.POINTER(ctypes.c_char_p),
                                    ctypes.POINTER(c_int)]
    lib.fm_batch_workspace_bytes.restype = ctypes.c_size_t
    lib.fm_batch_workspace_bytes.argtypes = [vp]
    lib.fm_batch_destroy.restype = None
    lib.fm_batch_destroy.argtypes = [vp]
    lib.fm_autotune_estimate.restype = c_int
    lib.fm_autotune_estimate.argtypes = [c_int, c_int, ctypes.c_char_p, c_int, c_int, c_int,
                                         c_int, ctypes.POINTER(c_double)]
    return lib


LIB = _load()


def check(status: int):
    if status == FM_OK:
        return
    msg = "%s: %s" % (LIB.fm_status_string(status).decode(), LIB.fm_last_error().decode())
    if status == FM_ERR_UNKNOWN_ALG:
        raise KeyError(msg)
    if status in (FM_ERR_INVALID_ARG, FM_ERR_SHAPE, FM_ERR_UNSUPPORTED):
        raise ValueError(msg)
    raise FastMMError(status, msg)
