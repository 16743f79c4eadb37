#!/usr/bin/env python3
"""Host CPU baseline for bench.py -- TEST/REPORTING INFRASTRUCTURE ONLY.

Times the reference's CPU path on this host's cores, on seeded uniform(-1,1)
problems, median of `reps` runs after one warm-up (benchmarks/
matmul_benchmarks.cpp:24-41), in four modes:

  ref_seq_1core      the reference's own FastMatmul (oracle/_ref/libfmref.so:
                     its headers compiled from its sources, sequential build),
                     MKL leaves on one thread -- the reference's sequential
                     mode; at --n1 (default 4096) to bound the run time;
  ref_seq_threaded   the same build with MKL threaded over all cores, at --n;
  port_dfs_parallel  the C restatement built with OpenMP add loops
                     (oracle/liboracle_par.so) and threaded MKL leaves, at
                     --n: the semantics of the reference's _PARALLEL_ = 1
                     (DFS) build, whose adds are OpenMP parallel-for loops
                     and whose leaves use threaded MKL (algorithms/
                     strassen.hpp:17-19, 691-693).  That build itself needs
                     mkl.h, which this image does not have;
  classical_dgemm    MKL dgemm on all cores, at --n.

The headline `value` is ref_seq_threaded (the reference's own code on all
cores at the GPU workload's size).  Prints one JSON line.  Runs in its own
process: it loads MKL and an OpenMP runtime, which must not share a process
with torch's.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

# MKL's GNU threading layer: one OpenMP runtime (libgomp) for MKL and the
# parallel add loops; must be set before MKL loads
os.environ.setdefault("MKL_THREADING_LAYER", "GNU")

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def _median(fn, reps):
    fn()  # warm-up (MKL initialisation, page faults)
    times = sorted(fn() for _ in range(reps))
    return times[len(times) // 2]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--alg", default="grey424_26_206")
    ap.add_argument("--n", type=int, default=8192, help="square size of the all-core modes")
    ap.add_argument("--n1", type=int, default=4096, help="square size of the one-core mode")
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--modes", default="ref_seq_1core,ref_seq_threaded,port_dfs_parallel,classical_dgemm")
    a = ap.parse_args()

    import oracle as O

    threads = a.threads or len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    O.use_parallel_adds()
    have_mkl = O.use_mkl(threads, reproducible=False)
    ref = O.ref()
    if not have_mkl:
        print(json.dumps({"error": "MKL runtime not found"}))
        return

    def inputs(n, seed=0xFA57):
        rng = np.random.default_rng(seed)
        A = np.asfortranarray(rng.random((n, n)) * 2 - 1)
        B = np.asfortranarray(rng.random((n, n)) * 2 - 1)
        return A, B, np.zeros((n, n), order="F")

    def ref_run(A, B, C, n):
        def f():
            t0 = time.perf_counter()
            ref.fmref_fastmatmul_d(a.alg.encode(), n, n, n, ctypes.c_void_p(A.ctypes.data), n,
                                   ctypes.c_void_p(B.ctypes.data), n, ctypes.c_void_p(C.ctypes.data),
                                   n, a.steps, ctypes.c_double(1e-8), ctypes.c_double(1.0),
                                   ctypes.c_double(0.0))
            return time.perf_counter() - t0
        return f

    modes = {}
    wanted = a.modes.split(",")

    def record(name, n, cores, kind, t):
        fl = 2.0 * n ** 3
        modes[name] = {"value": fl / t / 1e12, "ms": t * 1e3, "n": n, "cores": cores, "kind": kind}

    if "ref_seq_1core" in wanted and ref is not None:
        O.use_mkl(1, reproducible=False)
        A, B, C = inputs(a.n1)
        record("ref_seq_1core", a.n1, 1, "reference", _median(ref_run(A, B, C, a.n1), a.reps))
        O.use_mkl(threads, reproducible=False)
    A, B, C = inputs(a.n)
    if "ref_seq_threaded" in wanted and ref is not None:
        record("ref_seq_threaded", a.n, threads, "reference", _median(ref_run(A, B, C, a.n), a.reps))
    if "port_dfs_parallel" in wanted:
        def port():
            t0 = time.perf_counter()
            O.fastmatmul(a.alg, A, B, None, a.steps)
            return time.perf_counter() - t0
        record("port_dfs_parallel", a.n, threads, "port", _median(port, a.reps))
    if "classical_dgemm" in wanted:
        def dgemm():
            t0 = time.perf_counter()
            O.gemm(A, B)
            return time.perf_counter() - t0
        record("classical_dgemm", a.n, threads, "mkl", _median(dgemm, a.reps))

    head = modes.get("ref_seq_threaded") or modes.get("port_dfs_parallel")
    if head is None:
        print(json.dumps({"error": "no fast-multiply mode ran", "modes": modes}))
        return
    print(json.dumps({
        "value": head["value"], "unit": "TFLOP/s", "ms": head["ms"], "cores": head["cores"],
        "kind": head["kind"], "blas": "MKL 2021.4 (libmkl_rt.so.1, GNU threading layer)",
        "sample": "%s numsteps=%d FP64 %d^3 uniform(-1,1), reference sequential build with MKL "
                  "leaves on %d threads, median of %d after 1 warm-up" % (
                      a.alg, a.steps, head["n"], head["cores"], a.reps),
        "modes": modes,
    }))


if __name__ == "__main__":
    main()
