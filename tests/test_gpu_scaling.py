"""Diagonal scaling around the fast multiply (fm_fastmatmul_scaled) against
the oracle's restatement of scaling/scaling.hpp.

* parity: on row/column-skewed inputs, the GPU result is within twice the
  oracle's own error (both against the float64 product, measured per entry
  relative to max|A[i,:]| max|B[:,j]| K) plus a floor;
* exactness: when every scaling factor is a power of two and rows (columns)
  that the algorithm combines share their factor, scaling commutes with
  every rounding, so the scaled and unscaled products agree bit for bit;
* effect: on the skewed FP32 data of the split-bf16 workload the scaled
  error is orders of magnitude below the unscaled one (as the oracle shows).
"""
import numpy as np
import pytest
import torch

from cases import make_inputs

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a, fm):
    return fm.col_major(torch.from_numpy(np.asfortranarray(a)).to(DEV))


def _skewed(m, k, n, seed):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (m, k)) * 10.0 ** rng.uniform(-3, 3, (m, 1))
    B = rng.uniform(-1, 1, (k, n)) * 10.0 ** rng.uniform(-3, 3, (1, n))
    return np.asfortranarray(A), np.asfortranarray(B)


def _err(C, A, B):
    D = A @ B
    scale = np.abs(A).max(axis=1)[:, None] * np.abs(B).max(axis=0)[None, :] * A.shape[1]
    return float((np.abs(C.astype(np.float64) - D) / scale).max())


@pytest.fixture(scope="module")
def scaling_mod():
    import scaling
    return scaling


@pytest.mark.parametrize("kind", ["outside_inside", "inside_outside", "inside", "outside"])
@pytest.mark.parametrize("dtype,floor", [(np.float64, 1e-15), (np.float32, 1e-6)])
def test_scaled_parity_with_oracle(fm, oracle_mod, scaling_mod, kind, dtype, floor):
    A, B = _skewed(256, 200, 180, 11)
    A, B = A.astype(dtype), B.astype(dtype)
    C0 = np.asfortranarray(np.random.default_rng(12).uniform(-1, 1, (256, 180)).astype(dtype))
    want = scaling_mod.scaled_fastmatmul(oracle_mod, "strassen", A, B, C0, 2, 1.5, -0.5, kind, 2)
    got = fm.fastmatmul_scaled("strassen", _dev(A, fm), _dev(B, fm), _dev(C0, fm), 2, alpha=1.5,
                               beta=-0.5, scaling=kind, scaling_steps=2).cpu().numpy()
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    truth = 1.5 * (A64 @ B64) + (-0.5) * C0.astype(np.float64)
    sc = np.abs(A64).max(axis=1)[:, None] * np.abs(B64).max(axis=0)[None, :] * 200
    e_ref = float((np.abs(want.astype(np.float64) - truth) / sc).max())
    e = float((np.abs(got.astype(np.float64) - want.astype(np.float64)) / sc).max())
    assert e <= 2 * e_ref + floor, (e, e_ref)


def test_power_of_two_scaling_is_exact(fm):
    """Row i of A has largest magnitude 2^(i mod 8) / 2 and column j of B
    2^(j mod 8): the outer factors are powers of two, and rows 64 apart (the
    ones Strassen's two levels combine at 256^2) share them."""
    A, B, _ = make_inputs(256, 256, 256, 13, "u11", np.float64)
    ra = 2.0 ** (np.arange(256) % 8)
    A = np.asfortranarray(A * 0.25 * ra[:, None])
    A[:, 7] = 0.5 * ra                         # the row maximum, exactly
    cb = 2.0 ** (np.arange(256) % 8)
    B = np.asfortranarray(B * 0.5 * cb[None, :])
    B[3, :] = cb                               # the column maximum, exactly
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    plain = fm.fastmatmul("strassen", Ad, Bd, None, 2)
    scaled = fm.fastmatmul_scaled("strassen", Ad, Bd, None, 2, scaling="outside", scaling_steps=1)
    assert torch.equal(plain, scaled)


@pytest.mark.parametrize("alg", ["strassen", "grey424_26_206"])
def test_scaling_repairs_skewed_fp32(fm, alg):
    A, B = _skewed(512, 512, 512, 7)
    A32, B32 = A.astype(np.float32), B.astype(np.float32)
    Ad, Bd = _dev(A32, fm), _dev(B32, fm)
    A64, B64 = A32.astype(np.float64), B32.astype(np.float64)
    unscaled = _err(fm.fastmatmul(alg, Ad, Bd, None, 2).cpu().numpy(), A64, B64)
    for kind in ("outside_inside", "outside"):
        scaled = _err(fm.fastmatmul_scaled(alg, Ad, Bd, None, 2, scaling=kind).cpu().numpy(),
                      A64, B64)
        assert scaled < 1e-5 and unscaled > 1.0, (kind, scaled, unscaled)
    # the split-bf16 leaves take the same scaling
    sb = _err(fm.fastmatmul_scaled(alg, Ad, Bd, None, 2, mode="split_bf16").cpu().numpy(),
              A64, B64)
    assert sb < 1e-5, sb


@pytest.mark.parametrize("alg", ["strassen", "smirnov333_23_139"])
def test_scaling_with_digit_leaves(fm, alg):
    """FP64 rows/columns scaled by 10^+-3: with the reference's outside-inside
    scaling the int8-digit leaves reach the plain FP64 leaves' componentwise
    error (within 4x), below the unscaled fast multiply's."""
    A, B = _skewed(486, 486, 486, 9)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    unscaled = _err(fm.fastmatmul(alg, Ad, Bd, None, 2).cpu().numpy(), A, B)
    plain = _err(fm.fastmatmul_scaled(alg, Ad, Bd, None, 2).cpu().numpy(), A, B)
    digits = _err(fm.fastmatmul_scaled(alg, Ad, Bd, None, 2, mode="ozaki_int8").cpu().numpy(),
                  A, B)
    assert digits <= 4 * plain + 1e-15, (digits, plain)
    assert digits < 0.1 * unscaled, (digits, unscaled)
