#!/usr/bin/env python3
"""Host CPU baseline for bench.py -- TEST/REPORTING INFRASTRUCTURE ONLY.

Times the reference's own FastMatmul (oracle/_ref/libfmref.so: the
reference's headers compiled from its sources, sequential build, MKL leaf
dgemm threaded over the given cores -- the reference's DFS mode without its
OpenMP add loops) on a seeded uniform(-1,1) problem, median of `reps` runs
after one warm-up, like benchmarks/matmul_benchmarks.cpp:24-41.  Falls back
to the C restatement (oracle/liboracle.so) when the reference build is
absent.  Prints one JSON line.  Run in its own process: it loads MKL's
OpenMP runtime, which must not share a process with torch's.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--alg", default="grey424_26_206")
    ap.add_argument("--m", type=int, default=8192)
    ap.add_argument("--k", type=int, default=8192)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()

    import oracle as O

    threads = a.threads or len(os.sched_getaffinity(0))
    rng = np.random.default_rng(0xFA57)
    A = np.asfortranarray(rng.random((a.m, a.k)) * 2 - 1)
    B = np.asfortranarray(rng.random((a.k, a.n)) * 2 - 1)
    C = np.zeros((a.m, a.n), order="F")
    have_mkl = O.use_mkl(threads, reproducible=False)
    ref = O.ref()
    kind = "reference" if ref is not None and have_mkl else "port"

    def run():
        if kind == "reference":
            t0 = time.perf_counter()
            ref.fmref_fastmatmul_d(a.alg.encode(), a.m, a.k, a.n, ctypes.c_void_p(A.ctypes.data), a.m,
                                   ctypes.c_void_p(B.ctypes.data), a.k, ctypes.c_void_p(C.ctypes.data),
                                   a.m, a.steps, ctypes.c_double(1e-8), ctypes.c_double(1.0),
                                   ctypes.c_double(0.0))
            return time.perf_counter() - t0
        t0 = time.perf_counter()
        O.fastmatmul(a.alg, A, B, None, a.steps)
        return time.perf_counter() - t0

    run()  # warm-up (MKL initialisation, page faults)
    times = sorted(run() for _ in range(a.reps))
    t = times[len(times) // 2]
    flops = 2.0 * a.m * a.n * a.k
    print(json.dumps({
        "value": flops / t / 1e12, "unit": "TFLOP/s", "ms": t * 1e3, "cores": threads,
        "kind": kind,
        "blas": "MKL 2021.4 (libmkl_rt.so.1)" if have_mkl else "oracle triple loop",
        "sample": "%s numsteps=%d FP64 %dx%dx%d uniform(-1,1), median of %d after 1 warm-up" % (
            a.alg, a.steps, a.m, a.k, a.n, a.reps),
    }))


if __name__ == "__main__":
    main()
