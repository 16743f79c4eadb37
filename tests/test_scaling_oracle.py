"""The oracle's restatement of the reference's diagonal scaling
(scaling/scaling.hpp:105-196), pinned by what the mathematics fixes:
outer scaling leaves every row of A and column of B with largest magnitude
1, inner scaling balances column k of A against row k of B, the factors
reconstruct the operands and the product, and power-of-two factors scale
exactly."""
import numpy as np
import pytest

from cases import make_inputs

import scaling as S


def _skewed(m, k, n, seed, dtype):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (m, k)) * 10.0 ** rng.uniform(-3, 3, (m, 1))
    B = rng.uniform(-1, 1, (k, n)) * 10.0 ** rng.uniform(-3, 3, (1, n))
    return np.asfortranarray(A.astype(dtype)), np.asfortranarray(B.astype(dtype))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_outer_scaling_normalises_rows_and_columns(dtype):
    A, B = _skewed(50, 40, 30, 1, dtype)
    As, Bs, r, s = S.scale(A, B, 1, "outside")
    eps = np.finfo(dtype).eps
    assert np.all(np.abs(np.abs(As).max(axis=1) - 1) <= 2 * eps)
    assert np.all(np.abs(np.abs(Bs).max(axis=0) - 1) <= 2 * eps)
    # diag(r) A' = A and B' diag(s) = B to a few ulps
    assert np.allclose(As * r[:, None], A, rtol=4 * eps, atol=0)
    assert np.allclose(Bs * s[None, :], B, rtol=4 * eps, atol=0)


def test_inner_scaling_balances_columns_against_rows():
    A, B = _skewed(60, 45, 35, 2, np.float64)
    As, Bs, r, s = S.scale(A, B, 1, "inside")
    assert np.all(r == 1) and np.all(s == 1)
    ca, rb = np.abs(As).max(axis=0), np.abs(Bs).max(axis=1)
    assert np.allclose(ca, rb, rtol=1e-14)
    # the product is unchanged up to rounding: column k of A / z, row k of B * z
    assert np.allclose(As @ Bs, A @ B, rtol=1e-12, atol=1e-12 * np.abs(A @ B).max())


@pytest.mark.parametrize("kind", sorted(S.TYPES))
def test_post_processing_reconstructs_the_product(kind):
    A, B = _skewed(64, 48, 40, 3, np.float64)
    As, Bs, r, s = S.scale(A, B, 2, kind)
    D = A @ B
    got = S.post(np.asfortranarray(As @ Bs), r, s)
    assert np.abs(got - D).max() <= 1e-13 * (np.abs(A).max(axis=1)[:, None]
                                             * np.abs(B).max(axis=0)[None, :] * 48).max()


def test_power_of_two_maxima_scale_exactly():
    """Rows / columns whose maxima are powers of two: every scaling
    multiplication is exact, so the scaled product times r and s equals the
    unscaled product bit for bit (power-of-two scaling commutes with rounding)."""
    A, B, _ = make_inputs(32, 24, 20, 4, "u11", np.float64)
    A[:, 0] = 2.0 ** np.arange(-16, 16)          # row i has max 2^(i-16)
    A[:, 1:] *= (2.0 ** np.arange(-16, 16))[:, None] * 0.5
    B[0, :] = 4.0
    As, Bs, r, s = S.scale(A, B, 1, "outside")
    assert np.all(np.log2(r) == np.round(np.log2(r)))
    assert np.array_equal(S.post(np.asfortranarray(As @ Bs), r, s), np.asfortranarray(A @ B))


def test_zero_rows_keep_factor_one():
    A, B, _ = make_inputs(8, 6, 5, 5, "u11", np.float64)
    A[3, :] = 0.0
    As, Bs, r, s = S.scale(A, B, 1, "outside_inside")
    assert r[3] == 1.0 and np.all(As[3] == 0.0) and np.all(np.isfinite(As))
