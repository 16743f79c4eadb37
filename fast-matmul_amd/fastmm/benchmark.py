"""Benchmark sets in the reference's MATLAB-ready output format.

The reference's benchmark driver (benchmarks/matmul_benchmarks.cpp:19-60)
prints, per algorithm and recursion depth, one line

    NAME_steps = [ P Q R steps ms ;  P Q R steps ms ; ];

where P x Q times Q x R is the multiply, steps the recursion depth and ms
the median of five trials, and NAME is the enum name of
algorithms/all_algorithms.hpp:247-326 (Alg2Str).  The format is described in
data/README.md:15-27, and the reference's MATLAB scripts plot it with
eff_gflops = P * R * (2Q - 1) / 1e9 / (ms / 1e3).

Here each trial runs on the GPU through RunAlgorithm on fresh uniform [0, 1)
matrices (RandomMatrix of the reference), and the time of a trial is the
device time of the multiply (fm_fastmatmul's ms_out), the counterpart of the
reference's std::chrono timing around its recursion
(algorithms/strassen.hpp:708-721).

    python -m fastmm.benchmark --alg FAST424_26_206 --steps 1,2 --sizes 4096,8192

``--mode split_bf16`` (FP32 data) or ``--mode ozaki_int8`` (FP64 data) runs
the multiplies with those leaf products (fm_fastmatmul's mode argument).
"""
from __future__ import annotations

import argparse
import re
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from . import Matrix, RunAlgorithm, fastmatmul, tables


def enum_name(algorithm: str) -> str:
    """Alg2Str: the enum name of an algorithm given by enum or namespace name."""
    if algorithm in tables.ENUM_TO_NAMESPACE:
        return algorithm
    for enum, ns in tables.ENUM_TO_NAMESPACE.items():
        if ns == algorithm and enum not in ("MKL", "CLASSICAL222", "FAST323_15_89"):
            return enum
    return algorithm.upper()



def single_benchmark(algorithm: str, m: int, k: int, n: int, num_steps: int,
                     trials: int = 5, dtype=torch.float64, seed: int = 0,
                     mode: str = "plain") -> float:
    """Median device ms of `trials` multiplies (benchmarks/matmul_benchmarks.cpp:23-41)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    times = []
    for _ in range(trials):
        A = Matrix(torch.rand((k, m), generator=g, device="cuda", dtype=dtype).t())
        B = Matrix(torch.rand((n, k), generator=g, device="cuda", dtype=dtype).t())
        C = Matrix(m, n, dtype=dtype)
        if mode != "plain":
            ns = tables.ENUM_TO_NAMESPACE.get(algorithm, algorithm)
            x = tables.default_x(algorithm, num_steps) if algorithm in tables.ENUM_TO_NAMESPACE else 1e-8
            _, ms = fastmatmul(ns, A.data, B.data, C.data, num_steps, x, mode=mode, timed=True)
        elif algorithm in tables.ENUM_TO_NAMESPACE:
            ms = RunAlgorithm(algorithm, A, B, C, num_steps)
        else:
            from . import alg
            ms = alg(algorithm).FastMatmul(A, B, C, num_steps)
        times.append(ms)
    times.sort()
    return times[trials // 2]


def benchmark_set(algorithm: str, shapes: Sequence[Tuple[int, int, int]],
                  steps: Sequence[int], trials: int = 5, dtype=torch.float64,
                  mode: str = "plain") -> List[str]:
    """One output line per depth (benchmarks/matmul_benchmarks.cpp:45-60)."""
    lines = []
    for s in steps:
        pts = []
        for (m, k, n) in shapes:
            ms = single_benchmark(algorithm, m, k, n, s, trials, dtype, mode=mode)
            pts.append(" %d %d %d %d %s ; " % (m, k, n, s, _fmt(ms)))
        lines.append("%s_%d = [%s];" % (enum_name(algorithm), s, "".join(pts)))
    return lines


def _fmt(ms: float) -> str:
    # std::cout's default 6 significant digits
    return "%.6g" % ms


_LINE = re.compile(r"^\s*([A-Za-z0-9_]+)_(\d+)\s*=\s*\[(.*)\]\s*;\s*$")


def parse(text: str) -> Dict[str, List[Tuple[int, int, int, int, float]]]:
    """Read lines of the format back: {NAME_steps: [(P, Q, R, steps, ms), ...]}."""
    out = {}
    for line in text.splitlines():
        mt = _LINE.match(line)
        if not mt:
            continue
        pts = []
        for chunk in mt.group(3).split(";"):
            f = chunk.split()
            if not f:
                continue
            if len(f) != 5:
                raise ValueError("bad data point %r" % chunk)
            pts.append((int(f[0]), int(f[1]), int(f[2]), int(f[3]), float(f[4])))
        out["%s_%s" % (mt.group(1), mt.group(2))] = pts
    return out


def eff_gflops(pt: Tuple[int, int, int, int, float]) -> float:
    """data/README.md:30: P R (2Q - 1) / 1e9 / (ms / 1e3)."""
    P, Q, R, _, ms = pt
    return P * R * (2 * Q - 1) / 1e9 / (ms / 1e3)


def main(argv: Optional[Sequence[str]] = None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--alg", required=True, help="enum (FAST424_26_206) or namespace name")
    ap.add_argument("--steps", default="1", help="comma-separated recursion depths")
    ap.add_argument("--sizes", default="4096",
                    help="comma-separated square sizes, or PxQxR shapes")
    ap.add_argument("--trials", type=int, default=5)
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--mode", default="plain", choices=["plain", "split_bf16", "ozaki_int8"],
                    help="leaf products (split_bf16 needs --fp32, ozaki_int8 FP64 data)")
    a = ap.parse_args(argv)
    shapes = []
    for tok in a.sizes.split(","):
        dims = [int(v) for v in tok.lower().split("x")]
        shapes.append(tuple(dims) if len(dims) == 3 else (dims[0],) * 3)
    steps = [int(v) for v in a.steps.split(",")]
    dtype = torch.float32 if a.fp32 else torch.float64
    for line in benchmark_set(a.alg, shapes, steps, a.trials, dtype, mode=a.mode):
        print(line, flush=True)


if __name__ == "__main__":
    main()
