"""FP64-accurate leaves from int8 digits (FM_MODE_OZAKI_INT8).

* The GPU digit product equals the restatement in oracle/ozaki.py bit for
  bit (every step is exact or one fixed-order FP64 rounding): classical
  products with peeling-sized edges, alpha/beta, 6-8 digits.
* Small integers (< 2^21: three digits, all pairs kept) give the exact
  product, also through two levels of a fast algorithm.
* On uniform(-1,1) data the fast algorithm with digit leaves stays within
  10x (8 digits) of the plain FP64 fast algorithm's normwise error against
  the quad-precision product.
"""
import numpy as np
import pytest
import torch

import ozaki
from cases import make_inputs

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a, fm):
    return fm.col_major(torch.from_numpy(np.asfortranarray(a)).to(DEV))


# FmTileCfg 1 (64 x 64) selects the 16 x 16 x 64 digit kernel, 2 (128 x 64)
# the 32 x 32 x 32 one on 64 x 64 tiles, -1 the default 32 x 32 x 32 one on
# 64 x 128 tiles
@pytest.mark.parametrize("kernel", [-1, 1, 2])
@pytest.mark.parametrize("S", [6, 7, 8])
@pytest.mark.parametrize("shape", [(64, 64, 64), (100, 130, 70), (257, 500, 55), (1, 64, 1),
                                   (130, 700, 96)])
def test_digit_gemm_matches_restatement(fm, S, shape, kernel):
    m, k, n = shape
    A, B, C0 = make_inputs(m, k, n, S * 7 + m, "u11", with_c=True)
    A[3 % m, :] *= 1e-9  # rows/columns with very different scales
    B[:, 2 % n] *= 1e6
    fm.set_ozaki_slices(S)
    fm.set_gemm_tile(kernel)
    try:
        C = fm.gemm(_dev(A, fm), _dev(B, fm), _dev(C0, fm), 1.5, -0.5,
                    mode="ozaki_int8").cpu().numpy()
    finally:
        fm.set_ozaki_slices(8)
        fm.set_gemm_tile(-1)
    want = ozaki.ozaki_matmul(A, B, S, 1.5, -0.5, C0)
    assert np.array_equal(C, want)


@pytest.mark.parametrize("alg,steps", [("strassen", 2), ("smirnov333_23_139", 2), ("grey424_26_206", 1)])
def test_digit_fast_exact_on_integers(fm, alg, steps):
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.integers(-(2 ** 12), 2 ** 12, (402, 300)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-(2 ** 12), 2 ** 12, (300, 562)).astype(np.float64))
    C = fm.fastmatmul(alg, _dev(A, fm), _dev(B, fm), None, steps, mode="ozaki_int8").cpu().numpy()
    assert np.array_equal(C, A @ B)


def test_digit_fast_accuracy(fm, oracle_mod):
    n = 729
    A, B, _ = make_inputs(n, n, n, 77, "u11")
    Q = oracle_mod.quad_matmul(A, B)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    Co = fm.fastmatmul("smirnov333_23_139", Ad, Bd, None, 2, mode="ozaki_int8").cpu().numpy()
    Cp = fm.fastmatmul("smirnov333_23_139", Ad, Bd, None, 2).cpu().numpy()
    eo = np.abs(Co - Q).max()
    ep = np.abs(Cp - Q).max()
    assert eo <= 10 * ep, (eo, ep)
