"""Every leaf-GEMM tile configuration, forced through fm_set_gemm_tile, on
integer data (bit-exact: every partial sum is an exact integer, so the
result cannot depend on the order of k) and on random data against a
float64 torch reference (normwise, by dtype), over edge shapes that leave
partial tiles and partial K slices."""
import numpy as np
import pytest
import torch

from cases import make_inputs

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
# plain FP64 / FP32 configurations (24-29 and 31 are split_bf16 only)
PLAIN_CFGS = list(range(24)) + [30, 32, 33, 34]


def _dev(a, fm):
    return fm.col_major(torch.from_numpy(np.asfortranarray(a)).to(DEV))


@pytest.fixture(autouse=True)
def _reset(fm):
    yield
    fm.set_gemm_tile(-1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("cfg", PLAIN_CFGS)
def test_tile_config_exact_on_integers(fm, cfg, dtype):
    fm.set_gemm_tile(cfg)
    # whole-tile shapes (every configuration's full-tile variant), edges, and
    # edges with whole K slices (the spread kernels' edge variant)
    for shape in [(1, 1, 1), (17, 33, 65), (128, 128, 128), (256, 256, 256), (512, 160, 512),
                  (200, 300, 150), (257, 129, 383), (64, 1000, 48), (200, 160, 150),
                  (130, 64, 260)]:
        m, k, n = shape
        A, B, C0 = make_inputs(m, k, n, 31, "int", dtype, with_c=True)
        exact = A.astype(np.float64) @ B.astype(np.float64)
        C = fm.gemm(_dev(A, fm), _dev(B, fm), _dev(C0, fm), 2.0, -3.0).cpu().numpy()
        assert np.array_equal(C.astype(np.float64), 2 * exact - 3 * C0.astype(np.float64)), shape


@pytest.mark.parametrize("shape", [(515, 1031, 263), (512, 1024, 512), (204, 512, 190)])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 4e-16), (np.float32, 2e-7)])
@pytest.mark.parametrize("cfg", PLAIN_CFGS)
def test_tile_config_random(fm, cfg, dtype, tol, shape):
    fm.set_gemm_tile(cfg)
    m, k, n = shape
    A, B, _ = make_inputs(m, k, n, 32, "u11", dtype)
    C = fm.gemm(_dev(A, fm), _dev(B, fm)).double().cpu().numpy()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    err = np.abs(C - ref).max() / (np.abs(A).max() * np.abs(B).max() * k)
    assert err < tol, err
