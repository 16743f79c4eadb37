"""Pins of the CPU oracle (oracle/fm_oracle.c) -- the checker the GPU path is
compared with.

* golden.json holds SHA-256 digests of the reference's own outputs
  (oracle/_ref, MKL leaves, 1 thread) on seeded cases: the oracle must
  reproduce every one bit for bit.
* Where oracle/_ref is built (this container), oracle and reference are also
  compared directly on extra cases.
* Integer inputs give exact pins: every intermediate is a small dyadic
  rational, so any recursion, peeling or alpha/beta mistake changes bits.
* The quad oracle (linalg/quadmatmul.hpp:8-26) is checked against exact
  rational arithmetic; the error metrics against their definitions
  (linalg/linalg.hpp:280-345), including dividing by the FIRST argument.
* README.md:72-73's known answer (Strassen, 257x500x55, glibc rand() inputs,
  3.87253e-15 with the authors' MKL) is reproduced in magnitude.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from cases import golden_cases, make_inputs, sha, x_for


def _case_params():
    return [pytest.param(c, id=c[0]) for c in golden_cases()]


@pytest.mark.parametrize("case", _case_params())
def test_oracle_matches_reference_digest(case, oracle_mod, golden):
    (cid, alg, m, k, n, steps, alpha, beta, dist, dt, seed) = case
    rec = {r["id"]: r for r in golden["cases"]}[cid]
    dtype = np.float64 if dt == "f64" else np.float32
    A, B, C0 = make_inputs(m, k, n, seed, dist, dtype, with_c=beta != 0.0)
    if not oracle_mod.reproducible():
        pytest.skip("MKL reproducible mode unavailable: digests are CPU-specific")
    C = oracle_mod.fastmatmul(alg, A, B, C0, steps, alpha, beta, x_for(alg, steps))
    assert sha(C) == rec["sha256"]


def test_readme_known_answer(oracle_mod, golden):
    A, B, _ = make_inputs(257, 500, 55, 0, "rand")
    C = oracle_mod.fastmatmul("strassen", A, B, None, 1)
    D = oracle_mod.gemm(A, B)
    err = oracle_mod.max_rel_diff(C, D)
    # README.md:72-73 prints 3.87253e-15 (authors' MKL); same magnitude here
    assert 1e-15 < err < 1e-14
    assert abs(err - 3.87253e-15) / 3.87253e-15 < 0.5


@pytest.mark.parametrize("alg,steps,shape", [
    ("strassen", 1, (257, 500, 55)), ("strassen", 3, (300, 500, 100)),
    ("grey424_26_206", 2, (402, 500, 562)), ("smirnov333_23_139", 2, (366, 500, 200)),
    ("grey423_20_144", 2, (101, 500, 340)), ("hk323_15_84", 1, (600, 500, 671)),
    ("grey333_23_125", 2, (257, 500, 55)), ("classical333_27_81", 1, (100, 100, 100))])
def test_integer_inputs_are_exact(alg, steps, shape, oracle_mod):
    m, k, n = shape
    A, B, C0 = make_inputs(m, k, n, 7, "int", with_c=True)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    C = oracle_mod.fastmatmul(alg, A, B, None, steps)
    assert np.array_equal(C, exact)
    # alpha = 3, beta = -2: still exact
    C = oracle_mod.fastmatmul(alg, A, B, C0, steps, 3.0, -2.0)
    assert np.array_equal(C, 3 * exact - 2 * C0)


def test_steps0_ignores_beta(oracle_mod):
    """num_steps = 0 is the reference's base case MatMul(A, B, C) whose beta
    defaults to 0 (algorithms/strassen.hpp:385-388): beta is ignored."""
    A, B, C0 = make_inputs(40, 30, 20, 3, "int", with_c=True)
    C = oracle_mod.fastmatmul("strassen", A, B, C0, 0, 2.0, 5.0)
    assert np.array_equal(C, 2.0 * (A @ B))


@pytest.mark.parametrize("alg,steps,shape,x", [
    ("strassen", 2, (257, 500, 55), 1e-8), ("bini322_10_52_approx", 1, (300, 500, 100), 1e-8),
    ("schonhage333_21_117_approx", 1, (366, 500, 200), 1e-4),
    ("grey323_15_89", 2, (634, 700, 1000), 1e-8), ("smirnov336_40_960", 1, (402, 500, 562), 1e-8)])
def test_oracle_equals_reference_build(alg, steps, shape, x, oracle_mod):
    if oracle_mod.ref() is None:
        pytest.skip("oracle/_ref not built")
    m, k, n = shape
    A, B, C0 = make_inputs(m, k, n, 11, "u01", with_c=True)
    for a, b in ((1.0, 0.0), (7.0, -1.3)):
        C1 = oracle_mod.fastmatmul(alg, A, B, C0, steps, a, b, x)
        C2 = oracle_mod.ref_fastmatmul(alg, A, B, C0, steps, a, b, x)
        assert np.array_equal(C1, C2)


def test_quad_matmul_against_rationals(oracle_mod):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((6, 9)) * 10.0 ** rng.integers(-5, 5, (6, 9))
    B = rng.standard_normal((9, 4))
    Q = oracle_mod.quad_matmul(A, B)
    for i in range(6):
        for j in range(4):
            s = sum(Fraction(A[i, l]) * Fraction(B[l, j]) for l in range(9))
            assert abs(Q[i, j] - float(s)) <= abs(float(s)) * 2 ** -52
    rows, cols = np.array([0, 5, 3]), np.array([1, 3, 0])
    v, absab = oracle_mod.quad_entries(A, B, rows, cols)
    assert np.array_equal(v, Q[rows, cols])
    assert np.allclose(absab, (np.abs(A) @ np.abs(B))[rows, cols], rtol=1e-15)


def test_metric_definitions(oracle_mod):
    C = np.array([[1.0, 2.0], [4.0, -8.0]], order="F")
    D = np.array([[1.5, 2.0], [4.0, -6.0]], order="F")
    # divides by the first argument (linalg/linalg.hpp:288)
    assert oracle_mod.max_rel_diff(C, D) == 0.5
    assert oracle_mod.max_rel_diff(D, C) == pytest.approx(1 / 3)
    assert oracle_mod.max_abs_diff(C, D) == 2.0
    assert oracle_mod.frobenius_diff(C, D) == pytest.approx(math.sqrt(0.25 + 4.0))


def test_naive_leaf_agrees(oracle_mod):
    """Without MKL the oracle's own triple loop gives the same integer results."""
    A, B, _ = make_inputs(129, 65, 33, 9, "int")
    oracle_mod.use_naive()
    try:
        C = oracle_mod.fastmatmul("grey424_26_206", A, B, None, 2)
    finally:
        oracle_mod.use_mkl(1)
    assert np.array_equal(C, A @ B)
