"""Calibrate the autotuner's cost model on this GPU and write its table.

    python -m fastmm.calibrate [--out PATH]

Measures, with the library's own kernels:
* the grouped leaf GEMM rate (TFLOP/s of 2mnk) against leaf size, for FP32
  and FP64 plain leaves, split-bf16 (FP32) and int8-digit (FP64) leaves: a
  batch of equal classical products of size s^3 per size, timed from a
  captured hipGraph (the form a batch's leaf launch takes);
* the S/T/W stage bandwidth: algorithmic stage bytes of one Strassen level
  at 4096^3 over the stages' device time (the library's per-launch events);
* the cost of one launch: a batch of one 64^3 product;
* the time per unit of K of one 64 x 64 tile (one 64 x 4096 x 64 product):
  a peeling launch lasts at least its longest tile, however few its flops;
* the cost of one recursion level of a batched problem beyond its bytes and
  launches: 48 uneven problems at one level of Strassen and of <4,2,3;20>,
  measured time minus the model's estimate without that term (less the
  model's error on the same problems done classically), per problem.

The table (default: tuning/mi355x.txt beside libfastmm.so, which the library
loads at first use) is small text, one rate per line; fm_autotune_load reads
another.
"""
from __future__ import annotations

import argparse
import json
import os
import tempfile
import time

import torch

import fastmm
from fastmm import _lib

SIZES = (256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096)
MODES = (("f32", torch.float32, "plain"), ("f32", torch.float32, "split_bf16"),
         ("f64", torch.float64, "plain"), ("f64", torch.float64, "ozaki_int8"))
DEFAULT_OUT = os.path.join(os.path.dirname(_lib.LIB_PATH), "tuning", "mi355x.txt")


def _time(fn, reps=5):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def _batch(shapes, dtype, mode, alg=None, steps=0, dev="cuda:0"):
    g = torch.Generator(device=dev)
    g.manual_seed(0xCA1)
    probs = []
    for (m, k, n) in shapes:
        A = fastmm.col_major(torch.rand((m, k), generator=g, device=dev, dtype=dtype) * 2 - 1)
        B = fastmm.col_major(torch.rand((k, n), generator=g, device=dev, dtype=dtype) * 2 - 1)
        probs.append({"A": A, "B": B, "C": fastmm.empty(m, n, dtype, dev), "alg": alg,
                      "num_steps": steps})
    return probs


def leaf_rates():
    out = []
    for dt, tdt, mode in MODES:
        for s in SIZES:
            count = int(max(1, min(256, round(2 * (4096 / s) ** 3))))
            probs = _batch([(s, s, s)] * count, tdt, mode)
            b = fastmm.Batch(probs, mode=mode, graph=True)
            t = _time(b.run)
            out.append((dt, mode, s, 2.0 * s ** 3 * count / t / 1e12))
            del b, probs
            torch.cuda.empty_cache()
    return out


def stage_rate(dtype):
    n = 4096
    g = torch.Generator(device="cuda:0")
    g.manual_seed(0xCA2)
    A = fastmm.col_major(torch.rand((n, n), generator=g, device="cuda:0", dtype=dtype) * 2 - 1)
    B = fastmm.col_major(torch.rand((n, n), generator=g, device="cuda:0", dtype=dtype) * 2 - 1)
    C = fastmm.empty(n, n, dtype, "cuda:0")
    fastmm.fastmatmul("strassen", A, B, C, 1)
    torch.cuda.synchronize()
    fastmm.profile_read()
    fastmm.profile_enable(True)
    try:
        for _ in range(5):
            fastmm.fastmatmul("strassen", A, B, C, 1)
        torch.cuda.synchronize()
    finally:
        fastmm.profile_enable(False)
    prof = fastmm.profile_read()
    plan = fastmm.plan_info("strassen", n, n, n, 1, dtype=dtype)
    bytes_ = sum(plan["stage_bytes"][k] for k in ("S", "T", "W"))
    ms = sum(prof[k][0] for k in ("S", "T", "W")) / 5
    return bytes_ / (ms * 1e-3) / 1e9


def launch_cost():
    b = fastmm.Batch(_batch([(64, 64, 64)], torch.float32, "plain"), graph=True)
    return _time(b.run, 50) * 1e6


def tile_k_cost(dtype, launch_us):
    """Time per unit of K of one 64 x 64 tile (a launch of one tile, K = 4096),
    less the launch: the floor under a peeling launch's duration."""
    k = 4096
    b = fastmm.Batch(_batch([(64, k, 64)], dtype, "plain"), graph=True)
    return max(0.0, (_time(b.run, 20) * 1e6 - launch_us) / k)


def write_table(path, rates, gbps, launch_us, level_us, tile_k, note):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write("# fastmm autotuner cost model, MI355X (gfx950); %s\n" % note)
        f.write("# leaf <dtype> <mode> <leaf size = cbrt(m n k)> <TFLOP/s of 2mnk>\n")
        for dt, mode, s, tf in rates:
            f.write("leaf %s %s %d %.2f\n" % (dt, mode, s, tf))
        for dt in ("f64", "f32"):
            f.write("stage_gbps %s %.0f\n" % (dt, gbps[dt]))
        f.write("launch_us %.2f\n" % launch_us)
        for dt in ("f64", "f32"):
            f.write("level_us %s %.2f\n" % (dt, level_us[dt]))
        for dt in ("f64", "f32"):
            f.write("tile_k_us %s %.4f\n" % (dt, tile_k[dt]))


def level_cost(path, rates, gbps, launch_us, tile_k):
    """Per-problem cost of one batched recursion level beyond bytes and
    launches: measured minus estimated (with that term zero), over batches of
    uneven problems (every dimension a multiple of 64 in [512, 4096]) run
    at one level of Strassen and of <4,2,3;20>, whose odd block sizes bring
    peeling and edge tiles."""
    import numpy as np
    rng = np.random.default_rng(0xCA3)
    vals = np.arange(512, 4097, 64)
    shapes = [tuple(int(v) for v in rng.choice(vals, 3)) for _ in range(48)]
    level = {}
    with tempfile.TemporaryDirectory() as d:
        tmp = os.path.join(d, "t.txt")
        write_table(tmp, rates, gbps, launch_us, {"f64": 0.0, "f32": 0.0}, tile_k, "level term zero")
        fastmm.autotune_load(tmp)
        for dt, tdt in (("f32", torch.float32), ("f64", torch.float64)):
            extra = []
            for alg in ("strassen", "grey423_20_144"):
                probs = _batch(shapes, tdt, "plain", alg, 1)
                b = fastmm.Batch(probs, graph=True)
                t = _time(b.run) * 1e6
                # the classical batch of the same shapes, measured and estimated
                cprobs = _batch(shapes, tdt, "plain", None, 0)
                cb = fastmm.Batch(cprobs, graph=True)
                tc = _time(cb.run) * 1e6
                est = sum(fastmm.autotune_estimate(alg, m, k, n, 1, dtype=tdt) for m, k, n in shapes)
                estc = sum(fastmm.autotune_estimate(None, m, k, n, 0, dtype=tdt) for m, k, n in shapes)
                # the level's measured cost beyond the model, with the model's
                # error on the classical batch taken out
                extra.append(((t - est) - (tc - estc)) / len(shapes))
                del b, probs, cb, cprobs
                torch.cuda.empty_cache()
            level[dt] = max(0.0, float(np.mean(extra)))
    return level


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args(argv)
    rates = leaf_rates()
    gbps = {"f64": stage_rate(torch.float64), "f32": stage_rate(torch.float32)}
    launch_us = launch_cost()
    tile_k = {"f64": tile_k_cost(torch.float64, launch_us), "f32": tile_k_cost(torch.float32, launch_us)}
    level_us = level_cost(a.out, rates, gbps, launch_us, tile_k)
    note = "python -m fastmm.calibrate, %s" % time.strftime("%Y-%m-%d")
    write_table(a.out, rates, gbps, launch_us, level_us, tile_k, note)
    fastmm.autotune_load(a.out)
    print(json.dumps({"table": a.out, "stage_gbps": gbps, "launch_us": launch_us,
                      "level_us": level_us, "tile_k_us": tile_k,
                      "leaf": {"%s/%s/%d" % (dt, m, s): round(tf, 2) for dt, m, s, tf in rates}}))


if __name__ == "__main__":
    main()
