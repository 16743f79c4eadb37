"""Mismatched operands are refused in Python before anything reaches the
library: every entry point checks shapes, dtypes and devices (a C narrower
than B, a float32 operand under float64 A, or a host C would otherwise let a
kernel write past C's allocation)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _m(fm, m, n, dtype=torch.float64, device=DEV):
    return fm.empty(m, n, dtype, device)


def _calls(fm):
    return [
        ("gemm", lambda A, B, C: fm.gemm(A, B, C)),
        ("fastmatmul", lambda A, B, C: fm.fastmatmul("strassen", A, B, C, 1)),
        ("fastmatmul_scaled", lambda A, B, C: fm.fastmatmul_scaled("strassen", A, B, C, 1)),
        ("batch", lambda A, B, C: fm.Batch([dict(A=A, B=B, C=C, alg="strassen", num_steps=1)],
                                           graph=False)),
    ]


@pytest.mark.parametrize("which", range(4))
def test_mismatches_raise(fm, which):
    name, call = _calls(fm)[which]
    A, B = _m(fm, 64, 32), _m(fm, 32, 48)
    bad = [
        (A, _m(fm, 31, 48), _m(fm, 64, 48)),                      # inner dimensions
        (A, B, _m(fm, 64, 47)),                                   # C too narrow
        (A, B, _m(fm, 63, 48)),                                   # C too short
        (A, _m(fm, 32, 48, torch.float32), _m(fm, 64, 48)),       # B dtype
        (A, B, _m(fm, 64, 48, torch.float32)),                    # C dtype
        (A, B, torch.empty(48, 64, dtype=torch.float64).t()),     # host C
        (A, B.cpu(), _m(fm, 64, 48)),                             # host B
        (A, B, torch.empty(64, 48, dtype=torch.float64, device=DEV)),  # row-major C
    ]
    for i, (a, b, c) in enumerate(bad):
        with pytest.raises((ValueError, TypeError)):
            call(a, b, c)
            pytest.fail("%s accepted bad case %d" % (name, i))


def test_good_call_still_runs(fm):
    A = fm.col_major(torch.ones(64, 32, dtype=torch.float64, device=DEV))
    B = fm.col_major(torch.ones(32, 48, dtype=torch.float64, device=DEV))
    C = fm.gemm(A, B, _m(fm, 64, 48))
    assert torch.equal(C, torch.full((64, 48), 32.0, dtype=torch.float64, device=DEV))
