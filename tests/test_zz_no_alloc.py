"""fm_fastmatmul_noalloc against fm_fastmatmul through the C ABI: the same
plans and kernels with the workspace and descriptors in the library's fixed
device scratch array, so the results must be identical bit for bit.  Two
calls in a row, the second smaller (the array keeps the first call's
contents), then modes whose leaves keep slice or digit planes in it, and a
multiply too large for the array, refused before anything runs."""
import ctypes

import numpy as np
import pytest
import torch

from cases import make_inputs

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _dev(a, fm):
    return fm.col_major(torch.from_numpy(np.asfortranarray(a)).to(DEV))


def _noalloc(fm, alg, A, B, C, steps, alpha, beta, mode="plain"):
    from fastmm import _lib
    m, k = A.shape
    n = B.shape[1]
    st = _lib.LIB.fm_fastmatmul_noalloc(
        alg.encode(), _lib.FM_F64 if A.dtype == torch.float64 else _lib.FM_F32,
        _lib.MODES[mode], m, k, n, ctypes.c_void_p(A.data_ptr()), fm._ld(A),
        ctypes.c_void_p(B.data_ptr()), fm._ld(B), ctypes.c_void_p(C.data_ptr()), fm._ld(C),
        steps, 1e-8, alpha, beta, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), None)
    return st


def _case(fm, alg, shape, steps, dtype, seed, alpha, beta, mode="plain"):
    m, k, n = shape
    A, B, C0 = make_inputs(m, k, n, seed, "u11", dtype, with_c=True)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    ref = fm.fastmatmul(alg, Ad, Bd, _dev(C0, fm), steps, alpha=alpha, beta=beta, mode=mode)
    C = _dev(C0, fm)
    st = _noalloc(fm, alg, Ad, Bd, C, steps, alpha, beta, mode)
    torch.cuda.synchronize()
    assert st == 0, fm._lib.LIB.fm_last_error()
    return ref, C


def test_noalloc_two_calls_second_smaller(fm):
    ref1, C1 = _case(fm, "grey424_26_206", (1024, 1024, 1024), 2, np.float64, 71, 1.5, -0.5)
    assert torch.equal(ref1, C1), float((ref1 - C1).abs().max())
    ref2, C2 = _case(fm, "strassen", (300, 200, 250), 2, np.float64, 72, 1.0, 0.0)
    assert torch.equal(ref2, C2), float((ref2 - C2).abs().max())


@pytest.mark.parametrize("mode,dtype,alg,shape", [
    ("split_bf16", np.float32, "strassen", (1024, 768, 512)),
    # leaves of 256^3: the last S/T stage writes the slice planes into the array
    ("split_bf16", np.float32, "strassen", (1024, 1024, 1024)),
    ("ozaki_int8", np.float64, "smirnov333_23_139", (486, 486, 486)),
    ("plain", np.float32, "grey234_20_144", (130, 290, 250)),   # peeling at both levels
])
def test_noalloc_leaf_modes(fm, mode, dtype, alg, shape):
    ref, C = _case(fm, alg, shape, 2, dtype, 73, 2.0, 0.25, mode)
    assert torch.equal(ref, C), float((ref - C).abs().max())


def test_noalloc_too_large_refused(fm):
    from fastmm import _lib
    n = 8192
    A = fm.empty(n, n, torch.float64, DEV)
    B = fm.empty(n, n, torch.float64, DEV)
    C = fm.empty(n, n, torch.float64, DEV)
    st = _noalloc(fm, "grey424_26_206", A, B, C, 2, 1.0, 0.0)
    assert st == _lib.FM_ERR_WORKSPACE
    assert b"scratch" in _lib.LIB.fm_last_error()
