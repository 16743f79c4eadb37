"""The C ABI library (no GPU needed): it loads, exports every entry point
include/fastmm.h declares, lists the registry, and rejects bad arguments
before touching a device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "fastmm.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const char\*|fm_status|int|void|size_t)\s+\**(fm_\w+)\(", txt, re.M)))


def test_exports_every_declared_symbol():
    from fastmm import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 9
    for nm in names:
        assert hasattr(lib, nm), nm
    assert sorted(_lib.EXPORTS) == names


def test_registry_matches_tables():
    import fastmm
    from fastmm import tables
    assert sorted(fastmm.algorithm_names()) == tables.algorithm_names()
    for name in tables.algorithm_names():
        t = tables.get_table(name)
        a = fastmm.alg(name)
        assert a.dims == (t.m, t.k, t.n) and a.rank == t.R


def test_argument_errors():
    from fastmm import _lib
    L = _lib.LIB
    nul = ctypes.c_void_p(0)
    # unknown algorithm
    st = L.fm_fastmatmul(b"nope", 0, 0, 4, 4, 4, nul, 4, nul, 4, nul, 4, 1, 1e-8, 1.0, 0.0,
                         None, None, 0, None)
    assert st == _lib.FM_ERR_UNKNOWN_ALG
    # leading dimension too small
    st = L.fm_fastmatmul(b"strassen", 0, 0, 8, 8, 8, nul, 4, nul, 8, nul, 8, 1, 1e-8, 1.0, 0.0,
                         None, None, 0, None)
    assert st == _lib.FM_ERR_INVALID_ARG
    # null pointers
    st = L.fm_fastmatmul(b"strassen", 0, 0, 8, 8, 8, nul, 8, nul, 8, nul, 8, 1, 1e-8, 1.0, 0.0,
                         None, None, 0, None)
    assert st == _lib.FM_ERR_INVALID_ARG
    # negative steps / dims
    st = L.fm_fastmatmul(b"strassen", 0, 0, -1, 8, 8, nul, 8, nul, 8, nul, 8, 1, 1e-8, 1.0, 0.0,
                         None, None, 0, None)
    assert st == _lib.FM_ERR_INVALID_ARG
    # probe sizes out of range fail before any device call
    out = ctypes.c_double(0.0)
    assert L.fm_probe_stream(4, 0, 1 << 20, ctypes.byref(out)) == _lib.FM_ERR_INVALID_ARG
    assert L.fm_probe_stream(4096, 4096, 1 << 30, ctypes.byref(out)) == _lib.FM_ERR_INVALID_ARG
    assert L.fm_probe_stream(4, 16, 0, ctypes.byref(out)) == _lib.FM_ERR_INVALID_ARG
    # empty output is a no-op
    st = L.fm_fastmatmul(b"strassen", 0, 0, 0, 8, 8, nul, 1, nul, 8, nul, 1, 1, 1e-8, 1.0, 0.0,
                         None, None, 0, None)
    assert st == _lib.FM_OK
    assert L.fm_status_string(_lib.FM_ERR_SHAPE) == b"shape not supported at this depth"


def test_workspace_query_and_depth_limit():
    import fastmm
    # grey424_26_206 (<4,2,4;26>, 22 S and 20 T outputs, 4 A blocks and 6 B
    # blocks used unchanged) over 2 levels at 8192^3, in units of 8-byte
    # elements of the level-2 blocks (512 x 2048 S, 2048 x 512 T, 512 x 512 M)
    s2, t2, m2 = 512 * 2048, 2048 * 512, 512 * 512
    m1 = 2048 * 2048
    one_level = (22 * 2048 * 4096 + 20 * 4096 * 2048 + 26 * m1          # level 1
                 + 26 * (22 * s2 + 20 * t2) + 676 * m2)                 # level 2
    # two-level stages store neither the level-1 S/T operands nor the level-1
    # products (the two W stages run as one)
    two_level = 26 * (22 * s2 + 20 * t2) + 26 * (4 * s2 + 6 * t2) + 676 * m2
    try:
        fastmm.set_stage_fusion(False)
        assert fastmm.workspace_bytes("grey424_26_206", 8192, 8192, 8192, 2) == 8 * one_level
    finally:
        fastmm.set_stage_fusion(True)
    assert fastmm.workspace_bytes("grey424_26_206", 8192, 8192, 8192, 2) == 8 * two_level
    # a peeling level in between keeps the one-level stages: 8200 / 4 = 2050
    # rows per level-1 block, not a multiple of 4
    one = fastmm.workspace_bytes("grey424_26_206", 8200, 8192, 8192, 2)
    fastmm.set_stage_fusion(False)
    try:
        assert fastmm.workspace_bytes("grey424_26_206", 8200, 8192, 8192, 2) == one
    finally:
        fastmm.set_stage_fusion(True)
    assert fastmm.workspace_bytes("strassen", 64, 64, 64, 0) == 0
    with pytest.raises(ValueError):
        fastmm.workspace_bytes("strassen", 3, 3, 3, 2)
    with pytest.raises(KeyError):
        fastmm.workspace_bytes("nope", 8, 8, 8, 1)


def _build_shim_example(tmp_path):
    import subprocess
    exe = str(tmp_path / "shim_example")
    lib_dir = os.path.join(ROOT, "fast-matmul_amd", "fastmm")
    cmd = ["g++", "-std=c++14", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
           "-I", "/opt/rocm/include", os.path.join(ROOT, "tests", "cpp", "shim_example.cpp"),
           "-o", exe, "-L", lib_dir, "-lfastmm", "-L", "/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath," + lib_dir + ":/opt/rocm/lib"]
    subprocess.check_call(cmd)
    return exe


def test_cpp_header_builds_against_the_library(tmp_path):
    """include/fastmm/fast_matmul.hpp (the reference's Matrix/FastMatmul
    spelling) compiles with a host compiler and links to libfastmm.so."""
    exe = _build_shim_example(tmp_path)
    assert os.path.exists(exe)


def test_autotune_cost_model():
    """The batch autotuner's cost model (host only): classical wins small
    problems, a fast algorithm wins large ones, too-deep recursion is
    infeasible."""
    import math
    import torch
    import fastmm
    f32 = torch.float32
    assert fastmm.autotune_estimate(None, 512, 512, 512, 0, f32) < \
        fastmm.autotune_estimate("strassen", 512, 512, 512, 2, f32)
    assert fastmm.autotune_estimate("strassen", 8192, 8192, 8192, 1, f32) < \
        fastmm.autotune_estimate(None, 8192, 8192, 8192, 0, f32)
    assert math.isinf(fastmm.autotune_estimate("strassen", 3, 3, 3, 2, f32))


def test_options_defaults_and_rejects():
    """fm_options_default reads the process defaults; bad options are refused
    before anything is planned."""
    import fastmm
    from fastmm import _lib
    o = fastmm.options()
    assert (o.gemm_tile, o.dfs_levels, o.workspace_limit, o.split_remainder, o.stage_fusion,
            o.ozaki_slices) == (-1, 0, 0, 1, 1, 8)
    with pytest.raises(TypeError):
        fastmm.options(no_such_field=1)
    L = _lib.LIB
    nul = ctypes.c_void_p(0)
    for bad in (dict(gemm_tile=99), dict(dfs_levels=-2), dict(ozaki_slices=5)):
        ob = fastmm.options(**bad)
        st = L.fm_fastmatmul_ex(b"strassen", 0, 0, 8, 8, 8, ctypes.c_void_p(16), 8,
                                ctypes.c_void_p(16), 8, ctypes.c_void_p(16), 8, 1, 1e-8, 1.0, 0.0,
                                ctypes.byref(ob), None, None, 0, None)
        assert st == _lib.FM_ERR_INVALID_ARG, bad
    h = ctypes.c_void_p()
    assert L.fm_plan_create(b"nope", 0, 0, 8, 8, 8, 8, 8, 8, 1, 1e-8, 1.0, 0.0, None,
                            ctypes.byref(h)) == _lib.FM_ERR_UNKNOWN_ALG
    assert L.fm_plan_create(b"strassen", 0, 0, 8, 8, 8, 4, 8, 8, 1, 1e-8, 1.0, 0.0, None,
                            ctypes.byref(h)) == _lib.FM_ERR_INVALID_ARG
    assert not h.value
    assert L.fm_plan_run(None, nul, nul, nul, None, None, 0, None) == _lib.FM_ERR_INVALID_ARG


def test_scaled_workspace_query():
    """The scaled call's workspace: its plan's, the copies of A and B, the
    product when beta != 0 and the scaling vectors, each 256-byte aligned."""
    import fastmm
    al = lambda v: (v + 255) // 256 * 256
    m, k, n = 300, 200, 250
    plan = fastmm.workspace_bytes("strassen", m, k, n, 2)
    L = max(m, k, n)
    base = al(plan) + al(8 * m * k) + al(8 * k * n) + al(8 * m) + al(8 * n) + 2 * al(8 * L) + 2 * al(8 * L)
    assert fastmm.scaled_workspace_bytes("strassen", m, k, n, 2) == base
    assert fastmm.scaled_workspace_bytes("strassen", m, k, n, 2, beta_nonzero=True) == base + al(8 * m * n)


def test_cost_table_load_and_interpolation(tmp_path):
    """The autotuner's cost table: the packaged, GPU-measured one is in use;
    another file replaces it; leaf rates interpolate in log size; a
    malformed file is refused and leaves the table unchanged."""
    import fastmm
    import torch
    packaged = fastmm.autotune_table()
    assert packaged.endswith(os.path.join("tuning", "mi355x.txt")), packaged
    f32 = torch.float32
    try:
        t = tmp_path / "t.txt"
        t.write_text("leaf f32 plain 512 50\nleaf f32 plain 2048 200\nstage_gbps f32 1000\n"
                     "stage_gbps f64 1000\nlaunch_us 0\nlevel_us f32 0\nlevel_us f64 0\n")
        fastmm.autotune_load(str(t))
        assert fastmm.autotune_table() == str(t)
        # 2 * 512^3 at 50 TFLOP/s, 2 * 2048^3 at 200, 1024 halfway (125) in log size
        assert abs(fastmm.autotune_estimate(None, 512, 512, 512, 0, f32) - 2 * 512 ** 3 / 50e6) < 1e-6
        assert abs(fastmm.autotune_estimate(None, 2048, 2048, 2048, 0, f32) - 2 * 2048 ** 3 / 200e6) < 1e-6
        assert abs(fastmm.autotune_estimate(None, 1024, 1024, 1024, 0, f32) - 2 * 1024 ** 3 / 125e6) < 1e-6
        # one Strassen level at 2048: stage bytes at 1 TB/s plus 7 leaves of 1024^3
        s_bytes = 4 * (1024 ** 2 * (4 + 5) * 2 + 1024 ** 2 * (7 + 4))
        want = s_bytes / 1e6 + 7 * 2 * 1024 ** 3 / 125e6  # bytes / (1000 GB/s) in us
        assert abs(fastmm.autotune_estimate("strassen", 2048, 2048, 2048, 1, f32) - want) < 1e-3
        bad = tmp_path / "bad.txt"
        bad.write_text("leaf f32 plain 512\n")
        with pytest.raises(ValueError):
            fastmm.autotune_load(str(bad))
        assert fastmm.autotune_table() == str(t)
    finally:
        fastmm.autotune_load(packaged)
