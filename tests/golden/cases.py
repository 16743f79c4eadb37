"""Case list and seeded inputs shared by make_golden.py and the tests.

Inputs are column-major float64 (or float32) arrays from numpy's PCG64
generator with a fixed seed per case:
  "u01"  : uniform [0, 1), the distribution of the reference's RandomMatrix
           (linalg/linalg.hpp:422-433);
  "u11"  : uniform [-1, 1), the distribution of the bench workloads;
  "int"  : integers in [-8, 8], for exact (bit-for-bit) pins;
  "rand" : glibc rand()/RAND_MAX from the initial seed, in the reference's
           fill order, i.e. the exact inputs of tests/matmul_tests.cpp's first
           case (README.md:72-73).
Shape lists come from tests/matmul_tests.cpp:140-145 (FastMatmulTests) and
:49-51 (ExtraMatMulTests); alpha/beta pairs from tests/matmul_tests.cpp:59-129.
"""
from __future__ import annotations

import ctypes
import hashlib

import numpy as np

# tests/matmul_tests.cpp:140-145, first six shapes (the rest are 4000-scale)
TEST_SHAPES = [(257, 500, 55), (300, 500, 100), (366, 500, 200), (101, 500, 340),
               (402, 500, 562), (600, 500, 671)]

# algorithms of tests/matmul_tests.cpp:148-174 plus those of the benchmark workloads
PARITY_ALGS = ["strassen", "grey333_23_125", "grey322_11_50", "grey332_15_103",
               "grey323_15_103", "grey323_15_89", "grey333_23_152", "grey333_23_221",
               "grey432_20_144", "grey423_20_144", "grey324_20_144", "grey342_20_144",
               "grey234_20_144", "grey243_20_144", "grey422_14_84", "grey424_26_257",
               "grey424_26_206", "grey433_29_234", "grey522_18_99", "smirnov333_23_128",
               "smirnov333_23_139", "hk332_15_94", "hk323_15_84", "bini322_10_52_approx",
               "schonhage333_21_117_approx", "classical222_8_24", "classical333_27_81"]

# RunAlgorithm's per-depth x for the APA algorithms (all_algorithms.hpp:102-138)
APA_X = {("bini322_10_52_approx", 1): 1e-8, ("bini322_10_52_approx", 2): 1e-4,
         ("schonhage333_21_117_approx", 1): 1e-4, ("schonhage333_21_117_approx", 2): 1e-2}


def x_for(alg, steps):
    return APA_X.get((alg, steps), 1e-8)


def make_inputs(m, k, n, seed, dist="u01", dtype=np.float64, with_c=False):
    if dist == "rand":
        libc = ctypes.CDLL("libc.so.6")
        libc.srand(1)

        def rm(r, c):
            v = np.array([libc.rand() for _ in range(r * c)], dtype=np.float64) / 2147483647.0
            return np.asfortranarray(v.reshape((c, r)).T)

        A, B = rm(m, k), rm(k, n)
        C = rm(m, n) if with_c else None
    else:
        rng = np.random.default_rng(seed)

        def gen(r, c):
            if dist == "u01":
                x = rng.random((r, c))
            elif dist == "u11":
                x = rng.random((r, c)) * 2.0 - 1.0
            elif dist == "int":
                x = rng.integers(-8, 9, size=(r, c)).astype(np.float64)
            else:
                raise ValueError(dist)
            return np.asfortranarray(x)

        A, B = gen(m, k), gen(k, n)
        C = gen(m, n) if with_c else None
    A = np.asfortranarray(A.astype(dtype))
    B = np.asfortranarray(B.astype(dtype))
    if C is not None:
        C = np.asfortranarray(C.astype(dtype))
    return A, B, C


def sha(a) -> str:
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def golden_cases():
    """(id, alg, m, k, n, steps, alpha, beta, dist, dtype, seed)."""
    cases = [("readme_kat", "strassen", 257, 500, 55, 1, 1.0, 0.0, "rand", "f64", 0)]
    seed = 100
    for alg in PARITY_ALGS:
        for si, (m, k, n) in enumerate(TEST_SHAPES[:3]):
            for steps in (1, 2):
                seed += 1
                cases.append(("%s_%d_%dx%dx%d" % (alg, steps, m, k, n), alg, m, k, n, steps,
                              1.0, 0.0, "u01", "f64", seed))
    # alpha/beta cases of tests/matmul_tests.cpp:59-129 at two of its shapes
    for alg, a, b in [("strassen", 1.0, 2.5), ("grey424_26_257", 1.0, 2.5),
                      ("smirnov333_23_128", 7.0, -1.3), ("grey522_18_99", 7.0, -1.3),
                      ("smirnov333_23_128", -1.2, 0.0), ("grey522_18_99", -1.2, 0.0)]:
        for (m, k, n) in [(1001, 200, 309), (1007, 400, 679)]:
            for steps in (1, 2):
                seed += 1
                cases.append(("ab_%s_%g_%g_%d_%dx%dx%d" % (alg, a, b, steps, m, k, n), alg, m, k,
                              n, steps, a, b, "u01", "f64", seed))
    # float32 path
    for alg in ["strassen", "grey424_26_206", "smirnov333_23_139", "grey423_20_144"]:
        for steps in (1, 2):
            seed += 1
            cases.append(("f32_%s_%d" % (alg, steps), alg, 402, 500, 562, steps, 1.0, 0.0, "u01",
                          "f32", seed))
    # mean-zero inputs (the distribution of the bench workloads)
    for alg, steps, shp in [("strassen", 1, (1024, 1024, 1024)), ("grey424_26_206", 2, (512, 512, 512)),
                            ("smirnov333_23_139", 2, (729, 729, 729))]:
        seed += 1
        cases.append(("u11_%s_%d_%d" % (alg, steps, shp[0]), alg) + shp + (steps, 1.0, 0.0, "u11", "f64", seed))
    # the tall-skinny grey424_26_206 workload (131072 x 2400 x 2400) at an
    # eighth of its row and column sizes: the same aspect, leaves of
    # 4096 x 600 x 300 (one level) and 1024 x 300 x 75 (two)
    for steps in (1, 2):
        seed += 1
        cases.append(("ts_grey424_26_206_%d_16384x1200x1200" % steps, "grey424_26_206", 16384, 1200,
                      1200, steps, 1.0, 0.0, "u11", "f64", seed))
    return cases
