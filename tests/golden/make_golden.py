#!/usr/bin/env python3
"""Write golden.json: outputs of the reference's own fast-matmul path on the
seeded cases of cases.py, recorded as data.

The reference is compiled from its sources by oracle/Makefile into
oracle/_ref/libfmref.so (leaf dgemm = the image's MKL runtime, one thread).
For every case this records the SHA-256 of the output matrix bytes and the
reference's MaxRelativeDiff against its classical MatMul
(linalg/linalg.hpp:280-295, tests/matmul_tests.cpp:19-27).  Run in the
container that holds /root/reference:

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
from cases import golden_cases, make_inputs, sha, x_for  # noqa: E402


def main():
    if O.ref() is None:
        raise SystemExit("oracle/_ref/libfmref.so not built")
    if not O.use_mkl(1, reproducible=True) or not O.reproducible():
        raise SystemExit("MKL runtime (with CBWR COMPATIBLE mode) not found")
    out = {"blas": "MKL 2021.4 (libmkl_rt.so.1), 1 thread, CBWR=COMPATIBLE", "cases": []}
    for (cid, alg, m, k, n, steps, alpha, beta, dist, dt, seed) in golden_cases():
        dtype = np.float64 if dt == "f64" else np.float32
        A, B, C0 = make_inputs(m, k, n, seed, dist, dtype, with_c=beta != 0.0)
        x = x_for(alg, steps)
        C = O.ref_fastmatmul(alg, A, B, C0, steps, alpha, beta, x)
        A64, B64 = A.astype(np.float64), B.astype(np.float64)
        D = O.gemm(A64, B64, None if C0 is None else C0.astype(np.float64), alpha, beta)
        rec = {"id": cid, "alg": alg, "m": m, "k": k, "n": n, "steps": steps, "alpha": alpha,
               "beta": beta, "dist": dist, "dtype": dt, "seed": seed, "x": x,
               "sha256": sha(C),
               "max_rel_diff": O.max_rel_diff(C.astype(np.float64), D),
               "normwise": O.normwise_err(C, D, A64, B64)}
        out["cases"].append(rec)
        print("%-48s %s %.3e" % (cid, rec["sha256"][:12], rec["max_rel_diff"]))
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
