"""The x rule for approximate (APA) algorithms, pinned on the oracle.

recommended_x(depth, dtype) = eps^(1/2^depth) rounded to a power of ten must
(1) reproduce the reference's FP64 table (algorithms/all_algorithms.hpp:
102-121) and (2) land near the minimum of the reference's lambda sweep
(tests/bini_stability.cpp:17-39) on the oracle, for FP64 and FP32 data: the
error there is within 10x of the sweep's minimum, at least 100x below the
rounding-dominated end (lambda = eps) and 10x below the approximation-
dominated end (lambda -> 1; FP32 at depth 2 is only ~40x better there)."""
import numpy as np
import pytest
import torch

from cases import make_inputs


def test_recommended_x_reproduces_the_reference_table():
    from fastmm import stability, tables
    for depth, x in {1: 1e-8, 2: 1e-4, 3: 1e-2, 4: 1e-1}.items():
        assert stability.recommended_x(depth, torch.float64) == x
        assert tables.default_x("BINI322_10_52_APPROX", depth) == x
    assert [stability.recommended_x(d, torch.float32) for d in (1, 2, 3)] == [1e-3, 1e-2, 1e-1]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("steps", [1, 2])
def test_recommended_x_near_sweep_minimum(oracle_mod, dtype, steps):
    from fastmm import stability
    A, B, _ = make_inputs(90, 80, 80, 3, "u01", np.float64)   # the reference test's shape
    A, B = A.astype(dtype), B.astype(dtype)
    C1 = A.astype(np.float64) @ B.astype(np.float64)

    def err(x):
        C2 = oracle_mod.fastmatmul("bini322_10_52_approx", A, B, None, steps, x=x)
        return float(np.abs((C1 - C2) / C1).max())

    eps = float(np.finfo(dtype).eps)
    lams = [eps * 2.0 ** i for i in range(int(-np.log2(eps)))]
    errs = [err(l) for l in lams]
    xr = stability.recommended_x(steps, torch.float64 if dtype == np.float64 else torch.float32)
    e = err(xr)
    assert e <= 10 * min(errs)
    assert e * 100 <= errs[0] and e * 10 <= errs[-1]
