"""Restatement of the int8-digit (Ozaki) FP64 product -- TEST INFRASTRUCTURE ONLY.

This is the arithmetic of the FM_MODE_OZAKI_INT8 leaves written in numpy with
exact integer products, operation for operation:
  e_i: |A[i,:]| < 2^e_i (frexp of the row maximum; -1074 for a zero row)
  digits: y = A[i,k] 2^-e_i; S times: y *= 128, d = trunc(y), y -= d
  P_u = sum_{s+t=u} D_s E_t (exact), u = 2..S+1 (digits numbered from 1)
  C = fl(fl(sum_{u=S+1..2} 2^-7u P_u) 2^(e_i+f_j) * alpha) (+ fl(beta C))
with the sum over u accumulated from the smallest scale up, each addition
rounded once.  The reference has no counterpart (its leaves are MKL dgemm);
this restatement is pinned by the mathematics in tests/test_ozaki_oracle.py:
digits reconstruct every value to within 2^-7S of its scale, small-integer
inputs give the exact product, and the error against the quad-precision
product obeys the bound (S+1) 2^-7S per term.
"""
from __future__ import annotations

import numpy as np


def exponents(X: np.ndarray, axis: int) -> np.ndarray:
    """e with max|X| along `axis` < 2^e (-1074 where the slice is all zero)."""
    mx = np.abs(X).max(axis=axis)
    _, e = np.frexp(mx)
    return np.where(mx > 0, e, -1074).astype(np.int64)


def digits(Y: np.ndarray, S: int) -> np.ndarray:
    """S truncated base-128 digits of Y (|Y| < 1), shape (S,) + Y.shape."""
    y = Y.astype(np.float64).copy()
    out = np.empty((S,) + Y.shape, dtype=np.int64)
    for s in range(S):
        y = y * 128.0
        t = np.trunc(y)
        out[s] = t.astype(np.int64)
        y = y - t
    return out


def split_rows(A, S):
    e = exponents(A, 1)
    return digits(np.ldexp(A, -e[:, None]), S), e


def split_cols(B, S):
    f = exponents(B, 0)
    return digits(np.ldexp(B, -f[None, :]), S), f


def ozaki_matmul(A, B, S=7, alpha=1.0, beta=0.0, C=None):
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    D, e = split_rows(A, S)
    E, f = split_cols(B, S)
    groups = [np.zeros((A.shape[0], B.shape[1]), dtype=np.int64) for _ in range(S)]
    for s in range(S):
        for t in range(S - s):
            groups[s + t] += D[s] @ E[t]
    acc = np.zeros((A.shape[0], B.shape[1]))
    for u in range(S - 1, -1, -1):  # group u holds digit pairs with s + t = u + 2
        acc = acc + np.ldexp(groups[u].astype(np.float64), -7 * (u + 2))
    r = np.ldexp(acc, e[:, None] + f[None, :]) * alpha
    if beta != 0.0:
        r = r + beta * np.asarray(C, dtype=np.float64)
    return r
