"""The lambda sweep of an approximate algorithm on the GPU
(tests/bini_stability.cpp:17-39) against the same sweep on the oracle: the
GPU's best lambda is within a factor 8 of the oracle's and its best error
within 10x, for FP64, FP32 and split-bf16 leaves."""
import numpy as np
import pytest
import torch

from cases import make_inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,mode", [(np.float64, "plain"), (np.float32, "plain"),
                                        (np.float32, "split_bf16")])
@pytest.mark.parametrize("steps", [1, 2])
def test_lambda_sweep_matches_oracle(fm, oracle_mod, dtype, mode, steps):
    from fastmm import stability
    A, B, _ = make_inputs(90, 80, 80, 3, "u01", np.float64)
    A, B = A.astype(dtype), B.astype(dtype)
    Ad = fm.col_major(torch.from_numpy(np.asfortranarray(A)).cuda())
    Bd = fm.col_major(torch.from_numpy(np.asfortranarray(B)).cuda())
    sweep = stability.lambda_sweep("bini322_10_52_approx", Ad, Bd, steps, mode=mode)
    C1 = A.astype(np.float64) @ B.astype(np.float64)
    ref = [(l, float(np.abs((C1 - oracle_mod.fastmatmul("bini322_10_52_approx", A, B, None,
                                                            steps, x=l)) / C1).max()))
           for l, _ in sweep]
    gl, ge = min(sweep, key=lambda t: t[1])
    rl, re_ = min(ref, key=lambda t: t[1])
    assert 1 / 8 <= gl / rl <= 8, (gl, rl)
    assert ge <= 10 * re_, (ge, re_)
    text = stability.format_sweep(steps, sweep)
    assert text.startswith("%d\n" % steps) and text.count(";") == len(sweep)
