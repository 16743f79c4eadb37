"""Pins of the Ozaki restatement (oracle/ozaki.py) by the mathematics."""
import numpy as np
import pytest

import ozaki


@pytest.mark.parametrize("S", [6, 7, 8])
def test_digits_reconstruct_within_bound(S):
    rng = np.random.default_rng(S)
    A = rng.standard_normal((50, 80)) * 10.0 ** rng.integers(-8, 8, (50, 1))
    D, e = ozaki.split_rows(A, S)
    assert np.abs(D).max() <= 127
    rec = sum(np.ldexp(D[s].astype(np.float64), -7 * (s + 1)) for s in range(S))
    scale = np.ldexp(1.0, e)[:, None]
    assert np.all(np.abs(A / scale - rec) < 2.0 ** (-7 * S))
    # every digit has the sign of its value (truncation toward zero)
    assert np.all((D * np.sign(A)[None]) >= 0)


def test_small_integers_are_exact():
    rng = np.random.default_rng(1)
    A = rng.integers(-(2 ** 20), 2 ** 20, (40, 60)).astype(np.float64)
    B = rng.integers(-(2 ** 20), 2 ** 20, (60, 30)).astype(np.float64)
    # values below 2^21 need at most 3 digits; all digit pairs (s, t <= 3) are kept
    C = ozaki.ozaki_matmul(A, B, 7)
    exact = (A.astype(object) @ B.astype(object)).astype(np.float64)
    assert np.array_equal(C, exact)


@pytest.mark.parametrize("S", [6, 7, 8])
def test_error_bound_against_quad(oracle_mod, S):
    rng = np.random.default_rng(10 + S)
    A = rng.random((120, 700)) * 2 - 1
    B = rng.random((700, 90)) * 2 - 1
    Q = oracle_mod.quad_matmul(A, B)
    C = ozaki.ozaki_matmul(A, B, S)
    K = A.shape[1]
    rowmax = np.abs(A).max(1)[:, None]
    colmax = np.abs(B).max(0)[None, :]
    # dropped digit pairs: at most (S+1) 2^-7S per term (scales <= 2x the maxima),
    # plus the final FP64 roundings
    bound = K * 4 * rowmax * colmax * (S + 1) * 2.0 ** (-7 * S) + 4 * 2.0 ** -53 * np.abs(Q)
    assert np.all(np.abs(C - Q) <= bound)
    plain = A @ B
    if S >= 8:
        # FP64-level on uniform data: within 10x of the plain product's error
        assert np.abs(C - Q).max() <= 10 * np.abs(plain - Q).max()
    elif S == 7:
        assert np.abs(C - Q).max() <= 20 * np.abs(plain - Q).max()
