"""Approximate (APA) algorithms: the parameter x, its sweep, and its choice
for lower-precision data.

An APA table (Bini <3,2,2;10>, Schonhage <3,3,3;21>, ...) has coefficients
that are polynomials in x and its inverse; the product it computes differs
from A B by O(x) per recursion level, while the division by powers of x
amplifies rounding by O(eps / x^...).  The reference fixes x per depth for
FP64 (algorithms/all_algorithms.hpp:102-121: 1e-8, 1e-4, 1e-2, 1e-1 at
depths 1-4) and studies the trade-off with a sweep of lambda = x from
DBL_EPSILON, doubling up to 1 (tests/bini_stability.cpp:17-39), printing
"lambda, MaxRelativeDiff(C_classical, C_apa);" per depth.

That table is eps^(1 / 2^depth) rounded to a power of ten for
eps = DBL_EPSILON; `recommended_x` applies the same rule to the working
precision, which is what FP32 data and the split-bf16 leaves need
(FP32 eps = 2^-23).
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import torch

from . import fastmatmul, gemm, tables

_EPS = {torch.float64: 2.0 ** -52, torch.float32: 2.0 ** -23}


def recommended_x(num_steps: int, dtype=torch.float64) -> float:
    """eps^(1 / 2^num_steps) rounded to the nearest power of ten; for FP64 this
    is the reference's table (algorithms/all_algorithms.hpp:102-121)."""
    if num_steps <= 0:
        return 1e-8
    e = math.log10(_EPS[dtype]) / (2 ** num_steps)
    return 10.0 ** round(e)


def max_relative_diff(C1: torch.Tensor, C2: torch.Tensor) -> float:
    """max |C1 - C2| / |C1| over the entries (linalg/linalg.hpp:280-295)."""
    d = (C1.double() - C2.double()).abs() / C1.double().abs()
    return float(d.max())


def lambda_sweep(alg_name: str, A: torch.Tensor, B: torch.Tensor, num_steps: int,
                 lambdas: Optional[Sequence[float]] = None, mode: str = "plain"
                 ) -> List[Tuple[float, float]]:
    """[(lambda, MaxRelativeDiff(classical, fast with x = lambda)), ...] with
    the reference's sweep (eps, 2 eps, 4 eps, ... < 1) by default."""
    t = tables.get_table(alg_name)
    if not t.is_apa:
        raise ValueError("%s is not an approximate algorithm" % alg_name)
    if lambdas is None:
        lam, lambdas = _EPS[A.dtype], []
        while lam < 1:
            lambdas.append(lam)
            lam *= 2
    C1 = gemm(A, B)
    out = []
    for lam in lambdas:
        C2 = fastmatmul(alg_name, A, B, None, num_steps, x=lam, mode=mode)
        out.append((float(lam), max_relative_diff(C1, C2)))
    return out


def format_sweep(num_steps: int, sweep: Sequence[Tuple[float, float]]) -> str:
    """The reference's printout: the depth, then 'lambda, error; ' pairs."""
    return "%d\n%s\n" % (num_steps, "".join("%g, %g; " % p for p in sweep))
