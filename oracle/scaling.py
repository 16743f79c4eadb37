"""Oracle (test infrastructure only): the reference's diagonal scaling,
restated in numpy with the reference's arithmetic types.

scaling/scaling.hpp:105-196: Scaling(A, B, steps, r, s, type) repeats outer
and/or inner scaling; PostProcessScaling(C, r, s) multiplies the product's
row i by r_i and column j by s_j.  The maxima are doubles (MaxAbsRow and
MaxAbsCol return double, scaling.hpp:58-83), reciprocals and square roots
are evaluated in double and stored in the Scalar type (InverseVector,
InnerScaling, scaling.hpp:46-52,111-121), every element update is one
Scalar multiplication, and the factor products r, s are Scalar (UpdateVals).
Zero rows/columns keep factor 1 here; the reference divides by zero.
"""
import numpy as np

TYPES = {"outside_inside": 0, "inside_outside": 1, "inside": 2, "outside": 3}


def _recip(v, dt):
    return np.where(v > 0, 1.0 / np.where(v > 0, v.astype(np.float64), 1.0), 1.0).astype(dt)


def scale(A, B, steps, kind="outside_inside"):
    """Scaled copies (A', B') and factors (r, s) with A B = diag(r) A' B' diag(s)
    up to the rounding of the scaling passes (scaling/scaling.hpp:150-196)."""
    dt = A.dtype
    A = np.array(A, dtype=dt, order="F", copy=True)
    B = np.array(B, dtype=dt, order="F", copy=True)
    r = np.ones(A.shape[0], dtype=dt)
    s = np.ones(B.shape[1], dtype=dt)

    def outer():  # OuterScaling, scaling.hpp:89-109
        nonlocal A, B, r, s
        x = np.abs(A).max(axis=1)
        A = (A * _recip(x, dt)[:, None]).astype(dt)
        r = np.where(x > 0, r * x, r).astype(dt)
        y = np.abs(B).max(axis=0)
        B = (B * _recip(y, dt)[None, :]).astype(dt)
        s = np.where(y > 0, s * y, s).astype(dt)

    def inner():  # InnerScaling, scaling.hpp:111-121
        nonlocal A, B
        ca = np.abs(A).max(axis=0).astype(np.float64)
        rb = np.abs(B).max(axis=1).astype(np.float64)
        ok = (ca > 0) & (rb > 0)
        z = np.where(ok, np.sqrt(np.where(ok, ca, 1.0) / np.where(ok, rb, 1.0)), 1.0).astype(dt)
        zi = (1.0 / z.astype(np.float64)).astype(dt)
        A = (A * zi[None, :]).astype(dt)
        B = (B * z[:, None]).astype(dt)

    t = TYPES[kind]
    for _ in range(steps):
        if t == 0:
            outer()
            inner()
        elif t == 1:
            inner()
            outer()
        elif t == 2:
            inner()
        else:
            outer()
    return np.asfortranarray(A), np.asfortranarray(B), r, s


def post(P, r, s, alpha=1.0, beta=0.0, C=None):
    """alpha * ((P * r_i) * s_j) + beta * C, one rounding per operation
    (PostProcessScaling, scaling.hpp:124-130, then alpha / beta)."""
    dt = P.dtype
    v = (P * r[:, None]).astype(dt)
    v = (v * s[None, :]).astype(dt)
    if alpha != 1.0:
        v = (dt.type(alpha) * v).astype(dt)
    if beta != 0.0:
        v = (v + (dt.type(beta) * C).astype(dt)).astype(dt)
    return np.asfortranarray(v)


def scaled_fastmatmul(oracle_mod, alg, A, B, C=None, steps=1, alpha=1.0, beta=0.0,
                      kind="outside_inside", scaling_steps=1):
    """The oracle's fast multiply on scaled copies, scaled back."""
    As, Bs, r, s = scale(A, B, scaling_steps, kind)
    P = oracle_mod.fastmatmul(alg, As, Bs, None, steps)
    return post(P, r, s, alpha, beta, C)
